"""ctypes binding of libconanpm.so (the C ABI declared in include/conanpm.h).

The library is loaded from this package directory only.  There is no
fallback: when the shared object is missing or no GPU is present the calls
raise, so a test or benchmark can never silently run on something else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libconanpm.so")

PM_FP64 = 0
PM_FP32 = 1
PM_FLAG_WRAP_GATHER = 1
PM_FLAG_MIDPOINT = 2
PM_FLAG_SERIAL_DEPOSIT = 4
PM_DIFF_STORE_POTENTIALS = 1
PM_DIFF_VERIFY_CHECKPOINTS = 2
PM_NPHASES = 10
PHASE_NAMES = ["keys+sort", "deposit", "fft rows fwd", "fft y fwd", "green (z fwd*G*bwd)",
               "fft y bwd", "fft rows bwd", "gather+kick(+drift)", "drift", "nyquist-plane passes"]

# Every exported entry point of include/conanpm.h.
EXPORTS = [
    "pm_last_error", "pm_create", "pm_destroy", "pm_init_zeldovich", "pm_drift", "pm_kick",
    "pm_steps", "pm_get_state", "pm_set_state", "pm_get_times", "pm_get_momentum_quantum",
    "pm_set_momentum_quantum", "pm_deposit", "pm_potential",
    "pm_forces", "pm_gather_mesh", "pm_poisson_mesh", "pm_power_spectrum", "pm_power_spectrum_noalloc",
    "pm_diagnostics", "pm_profile", "pm_sync",
    "pm_ic_scale_free", "pm_ic_cdm", "pm_ic_potential", "pm_ic_displacement", "pm_fft_r2c", "pm_fft_c2r",
    "pm_cic", "pm_diff_create", "pm_diff_destroy", "pm_diff_last_error", "pm_diff_forward", "pm_diff_backward",
    "pm_diff_get_stats", "pm_fof_create", "pm_fof_destroy", "pm_fof_last_error", "pm_fof_run", "pm_fof_get_stats",
    "pm_fof",
]


class PMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class pm_config(C.Structure):
    _fields_ = [("dim", C.c_int), ("Nm", C.c_int), ("Nf", C.c_int), ("L", C.c_double),
                ("H0", C.c_double), ("Omega_m", C.c_double), ("Omega_L", C.c_double),
                ("precision", C.c_int), ("flags", C.c_uint)]


class pm_ic_params(C.Structure):
    _fields_ = [("dim", C.c_int), ("mbits", C.c_int), ("L", C.c_double), ("ns", C.c_double),
                ("smooth", C.c_int), ("scale", C.c_double), ("sigma", C.c_double), ("nyquist_cutoff", C.c_int)]


class pm_diag(C.Structure):
    _fields_ = [("mass", C.c_double), ("momentum", C.c_double * 3), ("p2", C.c_double), ("kinetic", C.c_double),
                ("potential", C.c_double), ("a_pos", C.c_double), ("a_mom", C.c_double)]


class pm_diff_params(C.Structure):
    _fields_ = [("dim", C.c_int), ("mbits", C.c_int), ("f", C.c_int), ("L", C.c_double), ("ns", C.c_double),
                ("smooth", C.c_int), ("scale", C.c_double), ("sigma", C.c_double), ("H0", C.c_double),
                ("Omega_m", C.c_double), ("Omega_L", C.c_double), ("a0", C.c_double), ("da", C.c_double),
                ("nsteps", C.c_int), ("ckpt_every", C.c_int), ("precision", C.c_int), ("flags", C.c_int)]


class pm_diff_stats(C.Structure):
    _fields_ = [("forward_ms", C.c_double), ("backward_ms", C.c_double), ("device_bytes", C.c_double),
                ("ncheckpoints", C.c_int), ("recon_pos", C.c_longlong), ("recon_mom", C.c_double)]


class pm_fof_stats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("particles", C.c_longlong), ("ngroups", C.c_longlong),
                ("candidates", C.c_longlong), ("links", C.c_longlong), ("cells", C.c_int * 3),
                ("table_bits", C.c_int)]


_lib = None

_dp = C.POINTER(C.c_double)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m conan_amd.build` first")
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
    # Loaded first, it is the one the library binds to, so torch tensors and
    # the library share one runtime; loaded after the library, torch would
    # bring a second runtime into the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    V = C.c_void_p
    sig = {
        "pm_last_error": (C.c_char_p, [V]),
        "pm_create": (C.c_int, [C.POINTER(pm_config), C.POINTER(V)]),
        "pm_destroy": (None, [V]),
        "pm_init_zeldovich": (C.c_int, [V, _dp, C.c_double, C.c_double]),
        "pm_drift": (C.c_int, [V, C.c_double, C.c_double]),
        "pm_kick": (C.c_int, [V, C.c_double, C.c_double]),
        "pm_steps": (C.c_int, [V, C.c_double, C.c_double, C.c_int, _dp]),
        "pm_get_state": (C.c_int, [V, _dp, _dp]),
        "pm_set_state": (C.c_int, [V, _dp, _dp]),
        "pm_get_times": (C.c_int, [V, _dp, _dp]),
        "pm_get_momentum_quantum": (C.c_int, [V, _dp]),
        "pm_set_momentum_quantum": (C.c_int, [V, C.c_double]),
        "pm_deposit": (C.c_int, [V, _dp, C.POINTER(C.c_int32)]),
        "pm_potential": (C.c_int, [V, C.c_double, _dp]),
        "pm_forces": (C.c_int, [V, C.c_double, _dp]),
        "pm_gather_mesh": (C.c_int, [V, _dp, _dp]),
        "pm_poisson_mesh": (C.c_int, [V, _dp, C.c_double, _dp]),
        "pm_power_spectrum": (C.c_int, [V, C.c_int, _dp, _dp, _dp, _dp]),
        "pm_power_spectrum_noalloc": (C.c_int, [V, C.c_int, _dp, _dp, _dp, _dp]),
        "pm_diagnostics": (C.c_int, [V, C.POINTER(pm_diag)]),
        "pm_profile": (C.c_int, [V, C.c_int, _dp, C.POINTER(C.c_longlong)]),
        "pm_sync": (C.c_int, [V]),
        "pm_ic_scale_free": (C.c_int, [C.POINTER(pm_ic_params), _dp, _dp, _dp]),
        "pm_ic_cdm": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                                _dp]),
        "pm_ic_potential": (C.c_int, [C.c_int, C.c_int, C.c_double, _dp]),
        "pm_ic_displacement": (C.c_int, [C.c_int, C.c_int, C.c_double, _dp, _dp]),
        "pm_fft_r2c": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp]),
        "pm_fft_c2r": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp]),
        "pm_cic": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp]),
        "pm_diff_create": (C.c_int, [C.POINTER(pm_diff_params), C.POINTER(V)]),
        "pm_diff_destroy": (None, [V]),
        "pm_diff_last_error": (C.c_char_p, [V]),
        "pm_diff_forward": (C.c_int, [V, V, V, C.c_int]),
        "pm_diff_backward": (C.c_int, [V, V, V, C.c_int]),
        "pm_diff_get_stats": (C.c_int, [V, C.POINTER(pm_diff_stats)]),
        "pm_fof_create": (C.c_int, [C.c_int, C.c_double, C.c_double, C.POINTER(V)]),
        "pm_fof_destroy": (None, [V]),
        "pm_fof_last_error": (C.c_char_p, [V]),
        "pm_fof_run": (C.c_int, [V, C.c_longlong, V, V, C.POINTER(C.c_longlong), C.c_int]),
        "pm_fof_get_stats": (C.c_int, [V, C.POINTER(pm_fof_stats)]),
        "pm_fof": (C.c_int, [C.c_int, C.c_longlong, C.c_double, C.c_double, _dp, C.POINTER(C.c_longlong),
                             C.POINTER(C.c_longlong)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(C.POINTER(C.c_int32))
    raise TypeError(a.dtype)


def check(rc: int, h=None) -> None:
    if rc != 0:
        msg = load().pm_last_error(h)
        raise PMError(rc, msg.decode() if msg else "unknown error")
