"""Failure handling through the C ABI, forced from the host: non-finite
inputs and reductions return PM_ENONFINITE, out-of-order calls PM_ESTATE,
bad arguments PM_EINVAL, and a device allocation beyond the cap set by
CONANPM_MAX_DEVICE_BYTES (what an oversized mesh meets on a full GPU)
PM_EDEVICE, without an abort; the library keeps working afterwards.  No
kernel is asked to run on invalid data."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENONFINITE, EDEVICE, ESTATE, EINVAL = -5, -3, -4, -1


def say(*a):
    print("[case]", *a, flush=True)


def code(pm, fn, *a):
    with pytest.raises(pm._lib.PMError) as ei:
        fn(*a)
    return ei.value.code


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_non_finite_inputs(pm, orc, prec):
    dim, Nm, Nf, L = 3, 8, 16, 100.0
    rng = np.random.default_rng(5)
    X = rng.uniform(0, L, (Nm ** dim, dim))
    say("non-finite inputs", prec)
    with pm.Gravity(dim, Nm, Nf, L, precision=prec) as g:
        assert code(pm, g.kick, 0.5, 0.01) == ESTATE  # no state yet
        Xb = X.copy()
        Xb[3, 1] = np.nan
        assert code(pm, g.set_state, Xb) == ENONFINITE
        g.set_state(X)
        Pb = np.zeros_like(X)
        Pb[5, 0] = np.inf
        assert code(pm, g.set_state, X, Pb) == ENONFINITE
        phi = np.zeros(Nf ** dim)
        phi[7] = np.nan
        assert code(pm, g.gather, phi) == ENONFINITE
        assert code(pm, g.poisson, phi + 1.0, 0.5) == ENONFINITE
        assert code(pm, g.steps, 0.5, np.nan, 2) == ENONFINITE
        assert code(pm, g.drift, np.inf, 0.01) == ENONFINITE
        assert code(pm, g.kick, -0.5, 0.01) == EINVAL
        phim = np.zeros(Nm ** dim)
        phim[0] = np.inf
        assert code(pm, g.init, phim, 0.02, 0.03) == ENONFINITE
        # the handle still works
        g.set_state(X)
        g.steps(0.5, 0.01, 2)
        assert np.all(np.isfinite(g.get_state()[0]))
    noise = np.zeros(Nm ** dim)
    noise[1] = np.nan
    assert code(pm, pm.ic_scale_free, dim, 3, L, -2.0, noise) == ENONFINITE


def test_overflowing_state_is_reported(pm):
    """A finite state whose FP64 reductions overflow (|P|^2 beyond the double
    range) is reported by pm_diagnostics as non-finite, not returned."""
    dim, Nm, Nf, L = 3, 8, 16, 100.0
    X = np.random.default_rng(6).uniform(0, L, (Nm ** dim, dim))
    P = np.zeros_like(X)
    P[9] = 1e300
    with pm.Gravity(dim, Nm, Nf, L) as g:
        g.set_state(X, P)
        assert code(pm, g.diagnostics) == ENONFINITE
        g.set_state(X, np.zeros_like(X))
        assert np.isfinite(g.diagnostics()["p2"])


def test_allocation_failure_is_clean():
    """In a child process with CONANPM_MAX_DEVICE_BYTES = 64 MB: a 256^3 mesh
    handle fails with PM_EDEVICE and a message naming the cap, the process
    goes on, and a small handle afterwards works."""
    script = r'''
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
from conan_amd import pm, _lib
try:
    pm.Gravity(3, 128, 256, 100.0)
    print("created"); sys.exit(1)
except _lib.PMError as e:
    print("code", e.code, str(e)); assert e.code == -3 and "CONANPM_MAX_DEVICE_BYTES" in str(e)
try:
    pm.fof(3, 100.0, 1.0, np.random.default_rng(0).uniform(0, 100.0, (4000000, 3)))
    print("fof created"); sys.exit(1)
except _lib.PMError as e:
    print("fof code", e.code); assert e.code == -3
with pm.Gravity(3, 8, 16, 100.0) as g:
    g.set_state(np.random.default_rng(1).uniform(0, 100.0, (512, 3)))
    g.steps(0.5, 0.01, 2)
print("ok")
'''
    env = dict(os.environ, CONANPM_MAX_DEVICE_BYTES=str(64 << 20))
    say("allocation cap 64 MB in a child process")
    r = subprocess.run([sys.executable, "-c", script, ROOT], env=env, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok")
