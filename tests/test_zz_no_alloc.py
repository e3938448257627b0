"""pm_power_spectrum_noalloc: the P(k) estimator with no allocation, its
scratch in one file-scope device array.  Called through the C ABI and
checked against pm_power_spectrum on the same handle state: mode counts
equal, mean |k| and P(k) within 1e-12 relative (the two sum in different
fixed orders).  Two calls in a row, the second on a smaller run with fewer
bins, so the second sees the first's leftovers in the shared array."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(pm, g, name, nbins, edges):
    k, p, n = np.empty(nbins), np.empty(nbins), np.empty(nbins)
    e = None if edges is None else np.ascontiguousarray(edges, dtype=np.float64)
    rc = getattr(g._lib, name)(g._h, nbins, pm.ptr(e), pm.ptr(k), pm.ptr(p), pm.ptr(n))
    pm.check(rc, g._h)
    return k, p, n


@pytest.mark.parametrize("first", ["noalloc", "existing"])
def test_power_spectrum_noalloc_matches(pm, orc, first):
    cases = [(3, 64, 128, "fp32", 64, None),
             (2, 32, 64, "fp64", 16, 2 * np.pi / 100.0 * np.linspace(1.0, 20.0, 17))]
    for dim, Nm, Nf, prec, nbins, edges in cases:
        L, a0, da = 100.0, 0.02, 0.0196
        mbits = int(np.log2(Nm))
        noise = orc.white_noise(17 + dim, Nm ** dim)
        phi = orc.compute_potential(dim, mbits, L, orc.ic_scale_free(dim, mbits, L, -2.0, noise))
        print("[case] noalloc P(k)", dim, Nm, Nf, prec, nbins, first, flush=True)
        with pm.Gravity(dim, Nm, Nf, L, precision=prec) as g:
            g.init(phi, a0, a0 + da / 2)
            g.steps(a0, da, 5)
            order = ["pm_power_spectrum_noalloc", "pm_power_spectrum"]
            if first == "existing":
                order.reverse()
            res = {name: _call(pm, g, name, nbins, edges) for name in order}
        ka, pa, na = res["pm_power_spectrum_noalloc"]
        kb, pb, nb = res["pm_power_spectrum"]
        assert np.array_equal(na, nb) and nb.sum() > 0
        m = nb > 0
        assert np.max(np.abs(ka[m] / kb[m] - 1)) < 1e-12
        assert np.max(np.abs(pa[m] / pb[m] - 1)) < 1e-12


def test_power_spectrum_noalloc_rejects_too_many_bins(pm, orc):
    with pm.Gravity(2, 16, 16, 100.0) as g:
        g.set_state(np.random.default_rng(3).uniform(0, 100.0, (256, 2)))
        with pytest.raises(pm._lib.PMError) as ei:
            _call(pm, g, "pm_power_spectrum_noalloc", 129, None)
        assert ei.value.code == -1 and "128" in str(ei.value)
