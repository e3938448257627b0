"""ctypes binding of the oracle (oracle/liboracle.so): the CPU FP64
restatement of the reference used as the parity checker.  Test
infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ODIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ODIR, "liboracle.so")

_dp = C.POINTER(C.c_double)
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", ODIR], check=True, stdout=subprocess.DEVNULL)
    _lib = _bind(C.CDLL(LIB))
    return _lib


def build_planted(define: str, outdir: str):
    """A copy of the oracle compiled with the planted mistake ORC_PLANT_<define>
    (see pm_oracle.cpp), for tests that show a pin fails on it."""
    out = os.path.join(outdir, f"liboracle_{define.lower()}.so")
    subprocess.run(["g++", "-std=c++14", "-O2", "-ffp-contract=off", "-frounding-math", "-fopenmp", "-fPIC",
                    "-shared", f"-DORC_PLANT_{define}", "-o", out, os.path.join(ODIR, "pm_oracle.cpp")], check=True)
    return _bind(C.CDLL(out))


def _bind(lib):
    V = C.c_void_p
    i, d, sz = C.c_int, C.c_double, C.c_size_t
    sig = {
        "orc_modulus_i": (i, [i, i]),
        "orc_modulus_d": (d, [d, d]),
        "orc_white_noise": (None, [C.c_ulong, sz, _dp]),
        "orc_mt19937_nth": (C.c_uint32, [C.c_ulong, sz]),
        "orc_fft": (None, [i, i, _dp, _dp, i, i]),
        "orc_cic_keys": (None, [i, i, d, _dp, sz, C.POINTER(C.c_int32)]),
        "orc_md_cic": (None, [i, i, d, _dp, sz, d, _dp, i]),
        "orc_poisson": (None, [i, i, d, d, d, d, d, _dp, _dp, i]),
        "orc_gradient_2nd_order": (None, [i, i, d, _dp, _dp]),
        "orc_gather": (None, [i, i, d, _dp, _dp, sz, i, _dp]),
        "orc_zeldovich": (None, [i, i, d, _dp, d, d, _dp, _dp]),
        "orc_adot": (d, [d, d, d, d]),
        "orc_sim_create": (V, [i, i, i, d, d, d, d, i, i]),
        "orc_sim_destroy": (None, [V]),
        "orc_sim_set_midpoint": (None, [V, i]),
        "orc_sim_set_serial_deposit": (None, [V, i]),
        "orc_sim_init": (None, [V, _dp, d, d]),
        "orc_sim_steps": (d, [V, i]),
        "orc_sim_drift": (None, [V, d, d]),
        "orc_sim_kick": (None, [V, d, d]),
        "orc_sim_get": (None, [V, _dp, _dp, _dp]),
        "orc_sim_set": (None, [V, _dp, _dp]),
        "orc_sim_mesh": (None, [V, _dp]),
        "orc_ic_scale_free": (None, [i, i, d, d, i, d, d, i, _dp, _dp, i]),
        "orc_ic_cdm": (None, [i, i, d, d, d, d, i, _dp, _dp, i]),
        "orc_cdm_amplitude": (d, [d, d, d, d]),
        "orc_cic_subsample": (None, [i, i, d, i, _dp, _dp]),
        "orc_compute_potential": (None, [i, i, d, _dp, i]),
        "orc_compute_displacement": (None, [i, i, d, _dp, _dp, i]),
        "orc_cic_tool": (None, [i, i, i, d, i, _dp, _dp]),
        "orc_power_spectrum": (None, [i, i, d, _dp, i, _dp, _dp, _dp, C.POINTER(C.c_int64), i]),
        "orc_fof": (C.c_long, [i, C.c_long, d, d, _dp, C.POINTER(C.c_long)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def modulus_i(a, b):
    return load().orc_modulus_i(a, b)


def modulus_d(a, b):
    return load().orc_modulus_d(a, b)


def white_noise(seed, n):
    out = np.empty(n)
    load().orc_white_noise(seed, n, _p(out))
    return out


def mt19937_nth(seed, n):
    return load().orc_mt19937_nth(seed, n)


def fft(a, dim, sign=-1, nthreads=1):
    """Unnormalised complex DFT of an N^dim array given flat (axis 0 fastest)."""
    a = np.ascontiguousarray(a, dtype=np.complex128).ravel()
    N = int(round(a.size ** (1.0 / dim)))
    out = np.empty_like(a)
    load().orc_fft(dim, N, a.view(np.float64).ctypes.data_as(_dp), out.view(np.float64).ctypes.data_as(_dp),
                   sign, nthreads)
    return out


def cic_keys(dim, N, L, X):
    X = _c(X)
    k = np.empty(X.shape[0], dtype=np.int32)
    load().orc_cic_keys(dim, N, L, _p(X), X.shape[0], k.ctypes.data_as(C.POINTER(C.c_int32)))
    return k


def md_cic(dim, N, L, X, m=1.0, nthreads=1):
    X = _c(X)
    rho = np.empty(N ** dim)
    load().orc_md_cic(dim, N, L, _p(X), X.shape[0], m, _p(rho), nthreads)
    return rho


def poisson(dim, N, L, rho, a, H0=68.5, Om=1.0, OL=0.0, nthreads=1):
    rho = _c(rho).ravel()
    phi = np.empty_like(rho)
    load().orc_poisson(dim, N, L, H0, Om, OL, a, _p(rho), _p(phi), nthreads)
    return phi


def gradient_2nd_order(dim, N, L, f):
    f = _c(f).ravel()
    out = np.empty((f.size, dim))
    load().orc_gradient_2nd_order(dim, N, L, _p(f), _p(out))
    return out


def gather(dim, N, L, phi, X, wrap=0):
    X = _c(X)
    F = np.empty_like(X)
    load().orc_gather(dim, N, L, _p(_c(phi).ravel()), _p(X), X.shape[0], wrap, _p(F))
    return F


def zeldovich(dim, Nm, L, phi, a_pos, a_mom):
    n = Nm ** dim
    X = np.empty((n, dim))
    P = np.empty((n, dim))
    load().orc_zeldovich(dim, Nm, L, _p(_c(phi).ravel()), a_pos, a_mom, _p(X), _p(P))
    return X, P


def adot(H0, Om, OL, a):
    return load().orc_adot(H0, Om, OL, a)


class Sim:
    """Gravity<R> + LeapFrog on the CPU (src/run/nbody.hh, leapfrog.hh)."""

    def __init__(self, dim, Nm, Nf, L, H0=68.5, Om=1.0, OL=0.0, nthreads=1, wrap=0, midpoint=False,
                 serial_deposit=False):
        self.lib = load()
        self.dim, self.Nm, self.Nf = dim, Nm, Nf
        self.h = self.lib.orc_sim_create(dim, Nm, Nf, L, H0, Om, OL, nthreads, wrap)
        if midpoint:
            self.lib.orc_sim_set_midpoint(self.h, 1)
        if serial_deposit:
            self.lib.orc_sim_set_serial_deposit(self.h, 1)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_sim_destroy(self.h)
            self.h = None

    def init(self, phi, a0, da):
        self.lib.orc_sim_init(self.h, _p(_c(phi).ravel()), a0, da)

    def steps(self, n):
        return self.lib.orc_sim_steps(self.h, n)

    def drift(self, a, da):
        self.lib.orc_sim_drift(self.h, a, da)

    def kick(self, a, da):
        self.lib.orc_sim_kick(self.h, a, da)

    def state(self):
        n = self.Nm ** self.dim
        X = np.empty((n, self.dim))
        P = np.empty((n, self.dim))
        t = np.empty(3)
        self.lib.orc_sim_get(self.h, _p(X), _p(P), _p(t))
        return X, P, t

    def set_state(self, X, P):
        self.lib.orc_sim_set(self.h, _p(_c(X)), _p(_c(P)))

    def mesh(self):
        out = np.empty(self.Nf ** self.dim)
        self.lib.orc_sim_mesh(self.h, _p(out))
        return out


def ic_scale_free(dim, mbits, L, ns, noise, smooth=False, scale=8.0, sigma=0.83, nthreads=1, nyquist_cutoff=False,
                  lib=None):
    out = np.empty((1 << mbits) ** dim)
    (lib or load()).orc_ic_scale_free(dim, mbits, L, ns, 1 if smooth else 0, scale, sigma, 1 if nyquist_cutoff else 0,
                                      _p(_c(noise).ravel()), _p(out), nthreads)
    return out


def ic_cdm(dim, mbits, L, ns, Omega0, H0, noise, nyquist_cutoff=False, nthreads=1, lib=None):
    out = np.empty((1 << mbits) ** dim)
    (lib or load()).orc_ic_cdm(dim, mbits, L, ns, Omega0, H0, 1 if nyquist_cutoff else 0, _p(_c(noise).ravel()),
                               _p(out), nthreads)
    return out


def cdm_amplitude(k, ns, Omega0, h):
    return load().orc_cdm_amplitude(k, ns, Omega0, h)


def cic_subsample(dim, Nm, L, s, X, lib=None):
    Y = np.empty(((Nm * s) ** dim, dim))
    (lib or load()).orc_cic_subsample(dim, Nm, L, s, _p(_c(X)), _p(Y))
    return Y


def compute_potential(dim, mbits, L, dens, nthreads=1):
    f = _c(dens).ravel().copy()
    load().orc_compute_potential(dim, mbits, L, _p(f), nthreads)
    return f


def compute_displacement(dim, mbits, L, phi, nthreads=1):
    n = (1 << mbits) ** dim
    psi = np.empty((n, dim))
    load().orc_compute_displacement(dim, mbits, L, _p(_c(phi).ravel()), _p(psi), nthreads)
    return psi


def cic_tool(dim, Nm, Nf, L, s, X):
    rho = np.empty(Nf ** dim)
    load().orc_cic_tool(dim, Nm, Nf, L, s, _p(_c(X)), _p(rho))
    return rho


def power_spectrum(dim, N, L, delta, edges, nthreads=1):
    nb = len(edges) - 1
    k = np.empty(nb)
    p = np.empty(nb)
    c = np.empty(nb, dtype=np.int64)
    load().orc_power_spectrum(dim, N, L, _p(_c(delta).ravel()), nb, _p(_c(edges)), _p(k), _p(p),
                              c.ctypes.data_as(C.POINTER(C.c_int64)), nthreads)
    return k, p, c


def log_edges(N, L, nbins=64):
    kf = 2 * np.pi / L
    kn = np.pi * N / L
    return kf * (kn / kf) ** (np.arange(nbins + 1) / nbins)


def fof(dim, L, ll, X):
    """Friends-of-friends groups of positions X (n x dim): (group per
    particle, numbered by each group's lowest index; group count)."""
    X = _c(X).reshape(-1, dim)
    g = np.empty(len(X), dtype=np.int64)
    ng = load().orc_fof(dim, len(X), L, ll, _p(X), g.ctypes.data_as(C.POINTER(C.c_long)))
    return g, int(ng)


def fof_brute(dim, L, ll, X):
    """Connected components of the all-pairs Proximity graph (the same
    predicate, evaluated for every pair with numpy), labelled like fof():
    an independent check of the cell-grid flood fill on small inputs."""
    X = np.asarray(X, dtype=np.float64).reshape(-1, dim)
    n = len(X)
    parent = np.arange(n)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for i in range(n):
        b = X - X[i]
        d = np.where(b > L / 2, b - L, np.where(b < -L / 2, b + L, b))
        v = np.zeros(n)
        for k in range(dim):
            v = v + d[:, k] * d[:, k]
        for j in np.nonzero(v <= ll * ll)[0]:
            ri, rj = find(i), find(int(j))
            if ri != rj:
                parent[max(ri, rj)] = min(ri, rj)
    roots = np.array([find(i) for i in range(n)])
    uniq = np.unique(roots)  # the roots are the component minima, so sorted = discovery order
    return np.searchsorted(uniq, roots).astype(np.int64), len(uniq)
