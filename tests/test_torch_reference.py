"""Pins the FP64 torch restatement of the forward map (tests/torch_pm.py), the
reference gradient of the GPU adjoint: its forward output against the C++
oracle, and its autograd gradient against central finite differences."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import torch_pm as tp  # noqa: E402

L = 100.0


@pytest.mark.parametrize("dim,mbits,smooth", [(2, 5, False), (3, 4, False), (3, 4, True)])
def test_ic_fields_match_oracle(orc, dim, mbits, smooth):
    w = orc.white_noise(3, (1 << mbits) ** dim)
    d, phi = tp.ic_fields(torch.tensor(w), dim, mbits, L, -2.0, smooth=smooth)
    dref = orc.ic_scale_free(dim, mbits, L, -2.0, w, smooth=smooth)
    pref = orc.compute_potential(dim, mbits, L, dref)
    assert np.max(np.abs(d.numpy() - dref)) < 1e-12 * np.max(np.abs(dref))
    assert np.max(np.abs(phi.numpy() - pref)) < 1e-12 * np.max(np.abs(pref))


@pytest.mark.parametrize("dim,mbits,f,n", [(2, 5, 2, 8), (3, 3, 2, 6)])
def test_forward_matches_oracle_run(orc, dim, mbits, f, n):
    Nm = 1 << mbits
    Nf = f * Nm
    a0, da = 0.05, 0.05
    w = orc.white_noise(11, Nm ** dim)
    rho, Y, P = tp.forward(torch.tensor(w), dim, mbits, f, L, -2.0, a0, da, n)
    dens = orc.ic_scale_free(dim, mbits, L, -2.0, w)
    phi = orc.compute_potential(dim, mbits, L, dens)
    sim = orc.Sim(dim, Nm, Nf, L, wrap=1)  # gather on folded coordinates
    sim.init(phi, a0, da)
    sim.steps(n)
    X, Po, _ = sim.state()
    rho_o = orc.md_cic(dim, Nf, L, X, m=(Nf / Nm) ** dim)
    assert np.max(np.abs(rho.numpy() - rho_o)) < 1e-10 * np.max(rho_o)
    dX = (Y.numpy().T - X + L / 2) % L - L / 2
    assert np.max(np.abs(dX)) < 1e-10 * L
    assert np.max(np.abs(P.numpy().T - Po)) < 1e-10 * np.max(np.abs(Po))


def test_autograd_matches_finite_differences(orc):
    dim, mbits, f, n = 2, 3, 2, 3
    Nm = 1 << mbits
    a0, da = 0.05, 0.1
    w = torch.tensor(orc.white_noise(5, Nm ** dim), requires_grad=True)
    target, _, _ = tp.forward(torch.tensor(orc.white_noise(6, Nm ** dim)), dim, mbits, f, L, -2.0, a0, da, n)
    target = target.detach()

    def loss(x):
        r, _, _ = tp.forward(x, dim, mbits, f, L, -2.0, a0, da, n)
        return ((r - target) ** 2).sum()

    lw = loss(w)
    lw.backward()
    g = w.grad.numpy()
    rng = np.random.default_rng(0)
    eps = 1e-6
    for _ in range(3):
        v = torch.tensor(rng.standard_normal(w.numel()))
        with torch.no_grad():
            fd = (loss(w + eps * v) - loss(w - eps * v)).item() / (2 * eps)
        ad = float(np.dot(g, v.numpy()))
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_autograd_matches_finite_differences_clustered_3d(orc):
    # 16^3 particles, a = 0.05 -> 1 in 10 steps (strongly clustered): the
    # central difference converges to the autograd gradient as eps -> 0 (the
    # CIC map is piecewise linear, so eps must be small enough that no
    # particle changes cell).
    dim, mbits, f, n = 3, 4, 1, 10
    Nm = 1 << mbits
    a0, da = 0.05, 0.095
    w = torch.tensor(orc.white_noise(1, Nm ** dim), requires_grad=True)
    target, _, _ = tp.forward(torch.tensor(orc.white_noise(2, Nm ** dim)), dim, mbits, f, L, -2.0, a0, da, n)
    target = target.detach()

    def loss(x):
        r, _, _ = tp.forward(x, dim, mbits, f, L, -2.0, a0, da, n)
        return ((r - target) ** 2).sum()

    loss(w).backward()
    g = w.grad.numpy()
    rng = np.random.default_rng(1)
    for _ in range(4):
        v = torch.tensor(rng.standard_normal(w.numel()))
        ad = float(np.dot(g, v.numpy()))
        with torch.no_grad():
            fd = (loss(w + 1e-7 * v) - loss(w - 1e-7 * v)).item() / 2e-7
        assert abs(fd - ad) <= 1e-4 * abs(ad)
