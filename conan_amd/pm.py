"""Python mirror of conan's particle-mesh Solver interface over libconanpm.so.

``Gravity`` follows Gravity<R> (src/run/nbody.hh:53-219) behind the Solver
seam of src/run/leapfrog.hh:6-21 (init / kick / drift), and ``LeapFrog``
follows LeapFrog::run (src/run/leapfrog.hh:24-46); ``fof`` / ``FofPlan`` give
the friends-of-friends groups of conan fof (src/fof/fof.cc:157-214).  Arrays use the reference
layout: particles AoS (Np, dim) float64 in mass-grid order, meshes N^dim
float64 with axis 0 fastest (returned as flat arrays).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PM_FLAG_MIDPOINT, PM_FLAG_WRAP_GATHER, PM_FP32, PM_FP64, check, ptr


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class Gravity:
    """One PM run on the GPU: particles and meshes stay resident in HBM."""

    def __init__(self, dim: int, Nm: int, Nf: int, L: float, H0: float = 68.5,
                 Omega_m: float = 1.0, Omega_L: float = 0.0, precision: str = "fp64",
                 wrap_gather: bool = False, midpoint: bool = False, serial_deposit: bool = False):
        lib = _lib.load()
        self.dim, self.Nm, self.Nf, self.L = dim, Nm, Nf, L
        self.Np = Nm ** dim
        self.M = Nf ** dim
        cfg = _lib.pm_config(dim, Nm, Nf, L, H0, Omega_m, Omega_L,
                             PM_FP64 if precision == "fp64" else PM_FP32,
                             (PM_FLAG_WRAP_GATHER if wrap_gather else 0) | (PM_FLAG_MIDPOINT if midpoint else 0)
                             | (_lib.PM_FLAG_SERIAL_DEPOSIT if serial_deposit else 0))
        h = _lib.C.c_void_p()
        check(lib.pm_create(_lib.C.byref(cfg), _lib.C.byref(h)))
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc):
        check(rc, self._h)

    # Solver seam
    def init(self, phi, a_pos: float, a_mom: float):
        self._ck(self._lib.pm_init_zeldovich(self._h, ptr(_c(phi).ravel()), a_pos, a_mom))

    def drift(self, a: float, da: float):
        self._ck(self._lib.pm_drift(self._h, a, da))

    def kick(self, a: float, da: float):
        self._ck(self._lib.pm_kick(self._h, a, da))

    def steps(self, a: float, da: float, n: int) -> float:
        out = _lib.C.c_double()
        self._ck(self._lib.pm_steps(self._h, a, da, n, _lib.C.byref(out)))
        return out.value

    # state
    def get_state(self):
        X = np.empty((self.Np, self.dim))
        P = np.empty((self.Np, self.dim))
        self._ck(self._lib.pm_get_state(self._h, ptr(X), ptr(P)))
        return X, P

    def set_state(self, X, P=None):
        X = _c(X).reshape(self.Np, self.dim)
        Pp = None if P is None else _c(P).reshape(self.Np, self.dim)
        self._ck(self._lib.pm_set_state(self._h, ptr(X), ptr(Pp)))

    def times(self):
        a, b = _lib.C.c_double(), _lib.C.c_double()
        self._ck(self._lib.pm_get_times(self._h, _lib.C.byref(a), _lib.C.byref(b)))
        return a.value, b.value

    def momentum_quantum(self) -> float:
        """FP32 form: the power of two that momenta and kicks are rounded to (0 for FP64)."""
        q = _lib.C.c_double()
        self._ck(self._lib.pm_get_momentum_quantum(self._h, _lib.C.byref(q)))
        return q.value

    def set_momentum_quantum(self, pq: float) -> None:
        self._ck(self._lib.pm_set_momentum_quantum(self._h, float(pq)))

    # operators
    def deposit(self, keys: bool = False):
        rho = np.empty(self.M)
        k = np.empty(self.Np, dtype=np.int32) if keys else None
        self._ck(self._lib.pm_deposit(self._h, ptr(rho), ptr(k)))
        return (rho, k) if keys else rho

    def keys(self):
        k = np.empty(self.Np, dtype=np.int32)
        self._ck(self._lib.pm_deposit(self._h, None, ptr(k)))
        return k

    def potential(self, a: float):
        phi = np.empty(self.M)
        self._ck(self._lib.pm_potential(self._h, a, ptr(phi)))
        return phi

    def forces(self, a: float):
        F = np.empty((self.Np, self.dim))
        self._ck(self._lib.pm_forces(self._h, a, ptr(F)))
        return F

    def gather(self, phi):
        F = np.empty((self.Np, self.dim))
        self._ck(self._lib.pm_gather_mesh(self._h, ptr(_c(phi).ravel()), ptr(F)))
        return F

    def poisson(self, rho, a: float):
        phi = np.empty(self.M)
        self._ck(self._lib.pm_poisson_mesh(self._h, ptr(_c(rho).ravel()), a, ptr(phi)))
        return phi

    def power_spectrum(self, nbins: int = 64, edges=None):
        k = np.empty(nbins)
        p = np.empty(nbins)
        n = np.empty(nbins)
        e = None if edges is None else _c(edges)
        self._ck(self._lib.pm_power_spectrum(self._h, nbins, ptr(e), ptr(k), ptr(p), ptr(n)))
        return k, p, n

    def power_spectrum_noalloc(self, nbins: int = 64, edges=None):
        """power_spectrum through pm_power_spectrum_noalloc (no allocation; nbins <= 128)."""
        k = np.empty(nbins)
        p = np.empty(nbins)
        n = np.empty(nbins)
        e = None if edges is None else _c(edges)
        self._ck(self._lib.pm_power_spectrum_noalloc(self._h, nbins, ptr(e), ptr(k), ptr(p), ptr(n)))
        return k, p, n

    def diagnostics(self):
        d = _lib.pm_diag()
        self._ck(self._lib.pm_diagnostics(self._h, _lib.C.byref(d)))
        return {"mass": d.mass, "momentum": list(d.momentum)[: self.dim], "p2": d.p2, "kinetic": d.kinetic,
                "potential": d.potential, "a_pos": d.a_pos, "a_mom": d.a_mom}

    def profile(self, enable: int = -1):
        ms = np.zeros(_lib.PM_NPHASES)
        cnt = np.zeros(_lib.PM_NPHASES, dtype=np.int64)
        self._ck(self._lib.pm_profile(self._h, enable, ptr(ms),
                                      cnt.ctypes.data_as(_lib.C.POINTER(_lib.C.c_longlong))))
        return ms, cnt

    def sync(self):
        self._ck(self._lib.pm_sync(self._h))


class LeapFrog:
    """LeapFrog::run (src/run/leapfrog.hh:33-46): init(a0, a0+da/2), then n
    iterations of drift(a, da); kick(a+da/2, da); a += da; snapshot hook."""

    def __init__(self, a0: float, da: float, n: int):
        self.a0, self.da, self.n = a0, da, n

    def run(self, solver: Gravity, phi, snapshot=None, every: int = 10):
        solver.init(phi, self.a0, self.a0 + self.da / 2)
        a = self.a0
        i = 0
        while i < self.n:
            chunk = self.n - i
            if snapshot is not None:
                chunk = min(chunk, every - (i % every))
            a = solver.steps(a, self.da, chunk)
            i += chunk
            if snapshot is not None and i % every == 0:
                snapshot(i - 1, a)
        return a


def ic_cdm(dim: int, mbits: int, L: float, noise, ns: float = 0.96, Omega0: float = 0.32, H0: float = 69.0,
           nyquist_cutoff: bool = False):
    """conan ic without --scale-free (the CDM branch, src/ic/ic.cc:14-35,84-97) on the GPU."""
    lib = _lib.load()
    dens = np.empty((1 << mbits) ** dim)
    check(lib.pm_ic_cdm(dim, mbits, L, ns, Omega0, H0, 1 if nyquist_cutoff else 0, ptr(_c(noise).ravel()),
                        ptr(dens)))
    return dens


def ic_scale_free(dim: int, mbits: int, L: float, ns: float, noise, smooth: bool = False,
                  scale: float = 8.0, sigma: float = 0.83, potential: bool = True, nyquist_cutoff: bool = False):
    """conan ic --scale-free [-p] [--nyquist-cutoff] on the GPU (src/ic/ic.cc:37-83,100-117)."""
    lib = _lib.load()
    N = 1 << mbits
    p = _lib.pm_ic_params(dim, mbits, L, ns, 1 if smooth else 0, scale, sigma, 1 if nyquist_cutoff else 0)
    dens = np.empty(N ** dim)
    phi = np.empty(N ** dim) if potential else None
    check(lib.pm_ic_scale_free(_lib.C.byref(p), ptr(_c(noise).ravel()), ptr(dens), ptr(phi)))
    return dens, phi


def ic_potential(dim: int, mbits: int, L: float, field):
    lib = _lib.load()
    f = _c(field).ravel().copy()
    check(lib.pm_ic_potential(dim, mbits, L, ptr(f)))
    return f


def ic_displacement(dim: int, mbits: int, L: float, potential):
    lib = _lib.load()
    N = 1 << mbits
    psi = np.empty((N ** dim, dim))
    check(lib.pm_ic_displacement(dim, mbits, L, ptr(_c(potential).ravel()), ptr(psi)))
    return psi


def fft_r2c(field, dim: int, precision: str = "fp64"):
    """Forward real FFT of an N^dim mesh; returns the full-layout complex
    spectrum array (N/2+1 along axis 0) built from the library's half layout."""
    lib = _lib.load()
    f = _c(field).ravel()
    N = int(round(f.size ** (1.0 / dim)))
    H = N // 2
    nrows = N ** (dim - 1)
    out = np.empty(2 * (H * nrows + nrows))
    check(lib.pm_fft_r2c(dim, N, PM_FP64 if precision == "fp64" else PM_FP32, ptr(f), ptr(out)))
    z = out[0::2] + 1j * out[1::2]
    main = z[: H * nrows].reshape(nrows, H)
    nyq = z[H * nrows:].reshape(nrows, 1)
    return np.concatenate([main, nyq], axis=1)  # (rows, N/2+1), rows index (y, z) with y fastest


def fft_c2r(spec, dim: int, N: int, precision: str = "fp64"):
    lib = _lib.load()
    H = N // 2
    nrows = N ** (dim - 1)
    spec = np.asarray(spec).reshape(nrows, H + 1)
    z = np.concatenate([spec[:, :H].ravel(), spec[:, H]])
    inp = np.empty(2 * z.size)
    inp[0::2] = z.real
    inp[1::2] = z.imag
    out = np.empty(N ** dim)
    check(lib.pm_fft_c2r(dim, N, PM_FP64 if precision == "fp64" else PM_FP32, ptr(inp), ptr(out)))
    return out


def cic(dim: int, Nm: int, Nf: int, L: float, X, s: int = 1):
    """conan cic's deposit (src/cic/cic.hh:49-81)."""
    lib = _lib.load()
    rho = np.empty(Nf ** dim)
    check(lib.pm_cic(dim, Nm, Nf, L, s, ptr(_c(X).ravel()), ptr(rho)))
    return rho


def _fof_check(rc: int, h=None) -> None:
    if rc != 0:
        msg = _lib.load().pm_fof_last_error(h)
        raise _lib.PMError(rc, msg.decode() if msg else "unknown error")


def fof(dim: int, L: float, ll: float, X):
    """Friends-of-friends groups of positions X (n x dim), the grouping of
    conan fof (src/fof/fof.cc:157-214): returns (group per particle, numbered
    in the order of each group's lowest index; number of groups)."""
    lib = _lib.load()
    X = _c(X).reshape(-1, dim)
    g = np.empty(len(X), dtype=np.int64)
    ng = C.c_longlong(0)
    _fof_check(lib.pm_fof(dim, len(X), L, ll, ptr(X), g.ctypes.data_as(C.POINTER(C.c_longlong)), C.byref(ng)))
    return g, int(ng.value)


class FofPlan:
    """Friends-of-friends on the GPU with device buffers kept between runs.
    run() takes a float64 array (host) or a contiguous float64 torch tensor on
    the GPU (device memory, complete before the call); the groups come back
    in the same kind of memory, int64."""

    def __init__(self, dim: int, L: float, ll: float):
        self._lib = _lib.load()
        self.dim = dim
        h = C.c_void_p()
        _fof_check(self._lib.pm_fof_create(dim, L, ll, C.byref(h)))
        self._h = h

    def run(self, X):
        ng = C.c_longlong(0)
        if isinstance(X, np.ndarray):
            X = _c(X).reshape(-1, self.dim)
            g = np.empty(len(X), dtype=np.int64)
            _fof_check(self._lib.pm_fof_run(self._h, len(X), X.ctypes.data, g.ctypes.data, C.byref(ng), 0), self._h)
            return g, int(ng.value)
        import torch
        if X.dtype != torch.float64 or not X.is_cuda or not X.is_contiguous():
            raise TypeError("device positions must be a contiguous float64 tensor on the GPU")
        n = X.numel() // self.dim
        g = torch.empty(n, dtype=torch.int64, device=X.device)
        torch.cuda.synchronize(X.device)
        _fof_check(self._lib.pm_fof_run(self._h, n, X.data_ptr(), g.data_ptr(), C.byref(ng), 1), self._h)
        return g, int(ng.value)

    def stats(self) -> dict:
        s = _lib.pm_fof_stats()
        _fof_check(self._lib.pm_fof_get_stats(self._h, C.byref(s)), self._h)
        return {"device_ms": s.device_ms, "particles": s.particles, "ngroups": s.ngroups,
                "candidates": s.candidates, "links": s.links, "cells": list(s.cells), "table_bits": s.table_bits}

    def close(self):
        if self._h:
            self._lib.pm_fof_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
