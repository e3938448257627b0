"""FP64 conservation reductions (pm_diagnostics): each sum against numpy sums
of the state and the deposit read back through the C ABI, and the energies
against the Layzer-Irvine cosmic energy equation over a run.

With P = a^2 dx/dt (the reference's momentum: dx/da = P / (a^2 adot),
src/run/nbody.hh:122-133) and the peculiar potential phi of the kick
(nabla^2 phi = G delta / a, nbody.hh:99-110), the kinetic energy K = m sum
|P|^2 / (2 a^2) and the potential energy W = 1/2 sum rho phi obey
d[a (K + W)] / da = -K (Layzer-Irvine): a (K + W) + int K da is constant.
The leapfrog of the reference is first order in da (its factors are taken at
the start of each interval), so the residual falls about linearly with da;
with the factors at the interval middles (PM_FLAG_MIDPOINT) faster.  A wrong
sign or factor of W, or a K taken at the wrong a, breaks the balance by O(1)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 2e-5)])
def test_diagnostics_match_numpy(pm, orc, prec, tol):
    dim, Nm, Nf, L, a0, da = 3, 32, 64, 100.0, 0.02, 0.0098
    noise = orc.white_noise(23, Nm ** dim)
    phi0 = orc.compute_potential(dim, 5, L, orc.ic_scale_free(dim, 5, L, -2.0, noise))
    say("diagnostics vs numpy", prec)
    with pm.Gravity(dim, Nm, Nf, L, precision=prec) as g:
        g.init(phi0, a0, a0 + da / 2)
        g.steps(a0, da, 30)
        d = g.diagnostics()
        X, P = g.get_state()
        a_pos, a_mom = g.times()
        rho = g.deposit()
        phi = g.potential(a_pos)
    m = (Nf / Nm) ** dim
    assert (d["a_pos"], d["a_mom"]) == (a_pos, a_mom)
    assert abs(d["mass"] / rho.sum() - 1) < 1e-13
    assert np.max(np.abs(np.array(d["momentum"]) - P.sum(axis=0))) < 1e-12 * np.abs(P).sum()
    p2 = np.sum(P ** 2)
    assert abs(d["p2"] / p2 - 1) < 1e-12
    assert abs(d["kinetic"] / (m * p2 / (2 * a_mom ** 2)) - 1) < 1e-12
    W = 0.5 * np.sum(rho * phi)
    say(f"  W {d['potential']:.10e} numpy {W:.10e} rel {abs(d['potential'] / W - 1):.3g}")
    assert W < 0 and abs(d["potential"] / W - 1) < tol


def _li_residual(pm, orc, da, midpoint, prec="fp64"):
    dim, Nm, Nf, L, a0, a1 = 3, 64, 64, 100.0, 0.02, 0.5
    noise = orc.white_noise(29, Nm ** dim)
    phi = orc.compute_potential(dim, 6, L, orc.ic_scale_free(dim, 6, L, -2.0, noise))
    n = int(round((a1 - a0) / da))
    a, K, W = [], [], []
    with pm.Gravity(dim, Nm, Nf, L, precision=prec, midpoint=midpoint) as g:
        g.init(phi, a0, a0 + da / 2)
        for i in range(n + 1):
            if i:
                g.steps(a0 + (i - 1) * da, da, 1)
            d = g.diagnostics()
            a.append(d["a_pos"])
            K.append(d["kinetic"])
            W.append(d["potential"])
    a, K, W = map(np.array, (a, K, W))
    # K is known at the interval middles a_i + da/2: at a_i (i >= 1) it is
    # the mean of its neighbours, and int K da over [a_j, a_j+1] is K_j da
    Ka = 0.5 * (K[:-1] + K[1:])
    E = a[1:] * (Ka + W[1:])
    intK = np.concatenate([[0.0], np.cumsum(K[1:-1] * da)])
    r = E - E[0] + intK
    return float(np.max(np.abs(r)) / np.max(np.abs(a[1:] * W[1:]))), float(K[1] / -W[1])


@pytest.mark.parametrize("midpoint", [False, True])
def test_layzer_irvine(pm, orc, midpoint):
    """64^3 particles on a 64^3 mesh, ns = -2, a = 0.02 -> 0.5; the residual
    of a (K + W) + int K da relative to max |a W|, at two step sizes."""
    r1, kw = _li_residual(pm, orc, 0.008, midpoint)
    r2, _ = _li_residual(pm, orc, 0.004, midpoint)
    say(f"Layzer-Irvine residual, midpoint={midpoint}: da 0.008 -> {r1:.3g}, da 0.004 -> {r2:.3g}; "
        f"K/|W| at the start {kw:.3f}")
    assert r2 < 0.02
    assert r2 < 0.6 * r1
