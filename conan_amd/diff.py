"""Differentiable particle-mesh runs over libconanpm.so (pm_diff_* of
include/conanpm.h).

``DiffPM`` maps a white-noise field w (the input of ``conan ic --scale-free``,
N^dim values, axis 0 fastest) through the scale-free initial conditions, the
Zel'dovich start and ``nsteps`` leapfrog steps of the FP32 form to the final
CIC density on the force mesh, and returns dL/dw for a given dL/drho with the
library's adjoint kernels.  ``density(noise, engine)`` wraps it as a
``torch.autograd.Function`` on device tensors, so a loss written in torch
differentiates back to the noise.

The reference has no gradients; the forward map is the one of
src/ic/ic.cc:37-117 and src/run/{nbody.hh,leapfrog.hh}.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PM_FP32


def _check(rc: int, h) -> None:
    if rc != 0:
        msg = _lib.load().pm_diff_last_error(h)
        raise _lib.PMError(rc, msg.decode() if msg else "unknown error")


class DiffPM:
    """One differentiable PM configuration (device state lives in the handle)."""

    def __init__(self, dim: int, mbits: int, f: int, L: float, ns: float, a0: float, da: float, nsteps: int,
                 ckpt_every: int = 5, sigma: float = 0.83, scale: float = 8.0, smooth: bool = False,
                 H0: float = 68.5, Omega_m: float = 1.0, Omega_L: float = 0.0, store_potentials: bool = True,
                 verify_checkpoints: bool = False):
        lib = _lib.load()
        self.dim, self.mbits, self.f, self.L = dim, mbits, f, L
        self.N = 1 << mbits
        self.Nf = f * self.N
        self.Np = self.N ** dim
        self.M = self.Nf ** dim
        flags = (_lib.PM_DIFF_STORE_POTENTIALS if store_potentials else 0) | \
            (_lib.PM_DIFF_VERIFY_CHECKPOINTS if verify_checkpoints else 0)
        p = _lib.pm_diff_params(dim, mbits, f, L, ns, 1 if smooth else 0, scale, sigma, H0, Omega_m, Omega_L,
                                a0, da, nsteps, ckpt_every, PM_FP32, flags)
        h = C.c_void_p()
        _check(lib.pm_diff_create(C.byref(p), C.byref(h)), None)
        self._h = h
        self._lib = lib
        # forward passes run so far: the handle's backward differentiates the
        # last one only, so a graph whose forward is no longer the last one
        # runs it again before its backward (see density)
        self.forwards = 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.pm_diff_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # host arrays --------------------------------------------------------
    def forward(self, noise: np.ndarray) -> np.ndarray:
        w = np.ascontiguousarray(noise, dtype=np.float64).reshape(-1)
        if w.size != self.Np:
            raise ValueError(f"noise must hold {self.Np} values")
        rho = np.empty(self.M)
        self.forwards += 1
        _check(self._lib.pm_diff_forward(self._h, w.ctypes.data, rho.ctypes.data, 0), self._h)
        return rho

    def backward(self, grad_rho: np.ndarray) -> np.ndarray:
        g = np.ascontiguousarray(grad_rho, dtype=np.float64).reshape(-1)
        if g.size != self.M:
            raise ValueError(f"grad_rho must hold {self.M} values")
        gw = np.empty(self.Np)
        _check(self._lib.pm_diff_backward(self._h, g.ctypes.data, gw.ctypes.data, 0), self._h)
        return gw

    # device tensors -----------------------------------------------------
    def forward_device(self, noise):
        import torch

        w = noise.detach().to(torch.float64).contiguous().reshape(-1)
        if w.numel() != self.Np or not w.is_cuda:
            raise ValueError(f"noise must be a device tensor of {self.Np} values")
        rho = torch.empty(self.M, dtype=torch.float64, device=w.device)
        torch.cuda.current_stream(w.device).synchronize()
        self.forwards += 1
        _check(self._lib.pm_diff_forward(self._h, w.data_ptr(), rho.data_ptr(), 1), self._h)
        return rho

    def backward_device(self, grad_rho):
        import torch

        g = grad_rho.detach().to(torch.float64).contiguous().reshape(-1)
        gw = torch.empty(self.Np, dtype=torch.float64, device=g.device)
        torch.cuda.current_stream(g.device).synchronize()
        _check(self._lib.pm_diff_backward(self._h, g.data_ptr(), gw.data_ptr(), 1), self._h)
        return gw

    def stats(self) -> dict:
        s = _lib.pm_diff_stats()
        _check(self._lib.pm_diff_get_stats(self._h, C.byref(s)), self._h)
        return {"forward_ms": s.forward_ms, "backward_ms": s.backward_ms, "device_bytes": s.device_bytes,
                "ncheckpoints": s.ncheckpoints, "recon_pos": s.recon_pos, "recon_mom": s.recon_mom}


def _autograd_function():
    import torch

    class PMDensity(torch.autograd.Function):
        """noise (device, N^dim) -> final CIC density (device, Nf^dim)."""

        @staticmethod
        def forward(ctx, noise, engine):
            ctx.engine = engine
            ctx.shape = noise.shape
            rho = engine.forward_device(noise)
            ctx.generation = engine.forwards
            ctx.save_for_backward(noise)
            return rho

        @staticmethod
        def backward(ctx, grad_rho):
            engine = ctx.engine
            if engine.forwards != ctx.generation:
                # another forward ran on this engine since this graph's one
                # (e.g. a second density() call, or a validation pass): the
                # handle holds that run's checkpoints, so replay this one
                (noise,) = ctx.saved_tensors
                engine.forward_device(noise)
            return engine.backward_device(grad_rho).reshape(ctx.shape), None

    return PMDensity


_PMDensity = None


def density(noise, engine: DiffPM):
    """Final CIC density of the run started from ``noise``, differentiable in torch."""
    global _PMDensity
    if _PMDensity is None:
        _PMDensity = _autograd_function()
    return _PMDensity.apply(noise, engine)
