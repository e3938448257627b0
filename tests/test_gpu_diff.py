"""Differentiable PM on the GPU (pm_diff_*): the forward map equals the plain
step sequence bit for bit, the backward pass reverses the leapfrog exactly,
and dL/d(noise) matches the FP64 torch autograd reference and central finite
differences."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 100.0


def _engine(dim, mbits, f, ns, a0, da, n, k, **kw):
    from conan_amd.diff import DiffPM

    print(f"[case] DiffPM dim={dim} mbits={mbits} f={f} ns={ns} a0={a0} da={da} n={n} ckpt={k}", flush=True)
    return DiffPM(dim, mbits, f, L, ns, a0, da, n, ckpt_every=k, **kw)


@pytest.mark.parametrize("dim,mbits,f,n", [(2, 6, 2, 8), (3, 4, 2, 6)])
def test_forward_equals_step_sequence(pm, orc, dim, mbits, f, n):
    a0, da, ns = 0.05, 0.05, -2.0
    Nm = 1 << mbits
    w = orc.white_noise(7, Nm ** dim)
    eng = _engine(dim, mbits, f, ns, a0, da, n, 4)
    rho = eng.forward(w)
    _, phi = pm.ic_scale_free(dim, mbits, L, ns, w)
    g = pm.Gravity(dim, Nm, f * Nm, L, precision="fp32", wrap_gather=True)  # the differentiable map's gather
    g.init(phi, a0, a0 + da / 2)
    g.steps(a0, da, n)
    # the same fixed-point sums: FP64 here, rounded to FP32 by the step's deposit
    assert np.array_equal(rho.astype(np.float32), g.deposit().astype(np.float32))
    assert np.max(np.abs(rho - g.deposit())) <= 2.0 ** -24 * np.max(rho)


def _loss_and_grad(eng, w, target):
    rho = eng.forward(w)
    gw = eng.backward(2 * (rho - target))
    return float(np.sum((rho - target) ** 2)), gw


@pytest.mark.parametrize("k", [1, 3, 10])
def test_reverse_leapfrog_reconstructs_checkpoints(orc, k):
    dim, mbits, f, n = 3, 5, 1, 10
    # value-by-value comparison with every checkpoint
    eng = _engine(dim, mbits, f, -2.0, 0.05, 0.095, n, k, verify_checkpoints=True)
    w = orc.white_noise(3, (1 << mbits) ** dim)
    target = np.ones((f << mbits) ** dim)
    _, gw = _loss_and_grad(eng, w, target)
    s = eng.stats()
    assert s["ncheckpoints"] == (n + k - 1) // k
    # every position and momentum bit recovered: the drift is an integer step
    # and the kick adds a multiple of the momentum quantum
    assert s["recon_pos"] == 0
    assert s["recon_mom"] == 0.0
    assert np.all(np.isfinite(gw))


@pytest.mark.parametrize("dim,mbits,f,n,k", [(3, 5, 1, 9, 4), (2, 6, 2, 7, 3)])
def test_backward_modes_give_the_same_gradient(orc, dim, mbits, f, n, k):
    """Stored or recomputed potentials, checksum or value-by-value checkpoint
    checks: the same numbers (the latter restores the checkpoint's storage
    order, which no result depends on)."""
    w = orc.white_noise(5, (1 << mbits) ** dim)
    target = np.ones((f << mbits) ** dim)
    out = []
    for store in (True, False):
        for verify in (False, True):
            eng = _engine(dim, mbits, f, -2.0, 0.05, 0.1, n, k, store_potentials=store, verify_checkpoints=verify)
            lv, g = _loss_and_grad(eng, w, target)
            s = eng.stats()
            assert s["recon_pos"] == 0 and s["recon_mom"] == 0.0
            out.append((lv, g))
    for lv, g in out[1:]:
        assert lv == out[0][0]
        assert np.array_equal(g, out[0][1])


def test_checkpoint_interval_does_not_change_the_gradient(orc):
    dim, mbits, f, n = 3, 4, 2, 8
    w = orc.white_noise(4, (1 << mbits) ** dim)
    target = np.ones((f << mbits) ** dim)
    grads = []
    for k in (1, 3, 8):
        eng = _engine(dim, mbits, f, -2.0, 0.05, 0.1, n, k)
        grads.append(_loss_and_grad(eng, w, target)[1])
    for g in grads[1:]:
        assert np.max(np.abs(g - grads[0])) <= 1e-9 * np.max(np.abs(grads[0]))


@pytest.mark.parametrize("dim,mbits,f,n,k", [(3, 5, 1, 10, 5), (2, 6, 2, 10, 5)])
def test_gradient_matches_fp64_torch_autograd(orc, dim, mbits, f, n, k):
    torch = pytest.importorskip("torch")
    import torch_pm as tp

    a0, da, ns = 0.05, 0.095, -2.0
    Nm = 1 << mbits
    w = orc.white_noise(1, Nm ** dim)
    w2 = orc.white_noise(2, Nm ** dim)
    eng = _engine(dim, mbits, f, ns, a0, da, n, k)
    target = eng.forward(w2)
    lg, g = _loss_and_grad(eng, w, target)
    wt = torch.tensor(w, requires_grad=True)
    rho_t, _, _ = tp.forward(wt, dim, mbits, f, L, ns, a0, da, n)
    lt = ((rho_t - torch.tensor(target)) ** 2).sum()
    lt.backward()
    gt = wt.grad.numpy()
    rel = np.linalg.norm(g - gt) / np.linalg.norm(gt)
    print(f"loss fp32 {lg:.9e} fp64 {lt.item():.9e}; gradient rel L2 {rel:.3e}", flush=True)
    assert abs(lg - lt.item()) <= 1e-4 * lt.item()
    assert rel < 1e-3


def _fd_errors(eng, w, target, g, eps, ndir, seed):
    rng = np.random.default_rng(seed)
    fd, ad = [], []
    for d in range(ndir):
        v = rng.standard_normal(w.size)
        lp = float(np.sum((eng.forward(w + eps * v) - target) ** 2))
        lm = float(np.sum((eng.forward(w - eps * v) - target) ** 2))
        fd.append((lp - lm) / (2 * eps))
        ad.append(float(np.dot(g, v)))
        print(f"direction {d}: fd {fd[-1]:.9e} adjoint {ad[-1]:.9e} rel {abs(fd[-1] - ad[-1]) / abs(ad[-1]):.2e}",
              flush=True)
    fd, ad = np.array(fd), np.array(ad)
    return np.linalg.norm(fd - ad) / np.linalg.norm(ad), np.median(np.abs(fd - ad) / np.abs(ad))


def test_gradient_matches_central_differences_128(orc):
    # 128^3 particles on a 128^3 mesh, P(k) ~ k^-2, 20 steps a = 0.05 -> 1,
    # checkpoints every 5 steps; loss = sum (rho - rho_target)^2 with the
    # target from a second seed; 16 seeded directions, eps = 1e-3.
    # The CIC map is piecewise linear in the positions: at eps = 1e-3 a
    # fraction of the particles changes cell between w - eps v and w + eps v,
    # so the central difference itself is off by O(1e-2) in this clustered
    # state -- the exact FP64 model shows the same (tests/torch_pm.py at 32^3:
    # 5e-4 .. 9e-2 at eps = 1e-3, converging to 1e-7 at eps = 1e-7).  The
    # adjoint is pinned against FP64 autograd above; here the bound guards
    # against gross errors (a sign, a factor, a missing term).
    dim, mbits, f, n, k = 3, 7, 1, 20, 5
    a0, da, ns = 0.05, 0.0475, -2.0
    Np = (1 << mbits) ** dim
    eng = _engine(dim, mbits, f, ns, a0, da, n, k)
    w = orc.white_noise(1, Np)
    target = eng.forward(orc.white_noise(2, Np))
    _, g = _loss_and_grad(eng, w, target)
    rel, med = _fd_errors(eng, w, target, g, 1e-3, 16, 2024)
    print(f"eps 1e-3: |fd - adjoint| / |adjoint| over 16 directions {rel:.3e}, median per direction {med:.3e}")
    assert rel < 5e-2 and med < 5e-2


def test_gradient_matches_central_differences_linear_regime(orc):
    # The same 128^3 problem stopped at a = 0.12 (20 steps from a = 0.02),
    # where few particles change cell over +-eps: FD and adjoint agree to 1e-3.
    dim, mbits, f, n, k = 3, 7, 1, 20, 5
    a0, da, ns = 0.02, 0.005, -2.0
    Np = (1 << mbits) ** dim
    eng = _engine(dim, mbits, f, ns, a0, da, n, k)
    w = orc.white_noise(1, Np)
    target = eng.forward(orc.white_noise(2, Np))
    _, g = _loss_and_grad(eng, w, target)
    rel, med = _fd_errors(eng, w, target, g, 1e-4, 16, 7)
    print(f"eps 1e-4: |fd - adjoint| / |adjoint| over 16 directions {rel:.3e}, median per direction {med:.3e}")
    assert rel < 2e-3 and med < 1e-3


def test_torch_autograd_function_on_device(orc):
    torch = pytest.importorskip("torch")
    from conan_amd.diff import density

    dim, mbits, f, n, k = 3, 4, 2, 6, 3
    Nm = 1 << mbits
    eng = _engine(dim, mbits, f, -2.0, 0.05, 0.1, n, k)
    w = orc.white_noise(9, Nm ** dim)
    target = eng.forward(orc.white_noise(10, Nm ** dim))
    _, g_host = _loss_and_grad(eng, w, target)
    wd = torch.tensor(w, device="cuda", requires_grad=True)
    td = torch.tensor(target, device="cuda")
    loss = ((density(wd, eng) - td) ** 2).sum()
    loss.backward()
    assert np.array_equal(wd.grad.cpu().numpy(), g_host)


def test_torch_autograd_interleaved_forwards(orc):
    """Two density() graphs on one engine, the second forward before the
    first backward: the first graph's backward replays its own forward, so
    both gradients equal those of separate runs."""
    torch = pytest.importorskip("torch")
    from conan_amd.diff import density

    dim, mbits, f, n, k = 3, 4, 2, 6, 3
    Nm = 1 << mbits
    eng = _engine(dim, mbits, f, -2.0, 0.05, 0.1, n, k)
    w1, w2 = orc.white_noise(11, Nm ** dim), orc.white_noise(12, Nm ** dim)
    target = eng.forward(orc.white_noise(10, Nm ** dim))
    _, g1 = _loss_and_grad(eng, w1, target)
    _, g2 = _loss_and_grad(eng, w2, target)
    td = torch.tensor(target, device="cuda")
    a = torch.tensor(w1, device="cuda", requires_grad=True)
    b = torch.tensor(w2, device="cuda", requires_grad=True)
    la = ((density(a, eng) - td) ** 2).sum()
    lb = ((density(b, eng) - td) ** 2).sum()
    with torch.no_grad():
        density(b, eng)  # a validation pass in between
    la.backward()
    lb.backward()
    assert np.array_equal(a.grad.cpu().numpy(), g1)
    assert np.array_equal(b.grad.cpu().numpy(), g2)


@pytest.mark.parametrize("store", [False, True])
def test_config_c5_512(orc, store):
    """C5: 512^3 particles on a 512^3 mesh, ns = -1.5, a = 0.05 -> 1 in 40
    steps, checkpoints every 8, the potentials recomputed (store = False) or
    kept from the forward pass: every checkpoint is reconstructed bit for bit
    by the reversed leapfrog, the gradient is finite and the same in both
    modes, and the device memory is reported (recomputing: 5 checkpoint states
    of 36 bytes per particle plus meshes)."""
    torch = pytest.importorskip("torch")
    dim, mbits, f, n, k = 3, 9, 1, 40, 8
    Np = (1 << mbits) ** dim
    eng = _engine(dim, mbits, f, -1.5, 0.05, 0.02375, n, k, store_potentials=store, verify_checkpoints=True)
    try:
        w = torch.tensor(np.random.default_rng(1).standard_normal(Np), device="cuda")
        target = eng.forward_device(torch.tensor(np.random.default_rng(2).standard_normal(Np), device="cuda"))
        rho = eng.forward_device(w)
        g = eng.backward_device(2 * (rho - target))
        torch.cuda.synchronize()
        s = eng.stats()
        print(f"[case] C5 store={store}: forward {s['forward_ms']:.1f} ms, backward {s['backward_ms']:.1f} ms, "
              f"ratio {(s['forward_ms'] + s['backward_ms']) / s['forward_ms']:.3f}, "
              f"{s['device_bytes'] / 1e9:.1f} GB, checkpoints {s['ncheckpoints']}", flush=True)
        assert s["ncheckpoints"] == 5
        assert s["recon_pos"] == 0 and s["recon_mom"] == 0.0
        assert bool(torch.isfinite(g).all()) and float(g.abs().max()) > 0
        assert s["device_bytes"] < (100e9 if store else 70e9)
        gsum = float(g.double().sum())
        gnorm = float(g.double().norm())
        _C5_GRADS[store] = (gsum, gnorm)
        if len(_C5_GRADS) == 2:
            assert _C5_GRADS[True] == _C5_GRADS[False]
    finally:
        eng.close()


_C5_GRADS = {}
