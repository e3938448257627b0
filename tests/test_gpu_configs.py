"""GPU parity against the oracle over two whole runs small enough for the
oracle to follow, and of the initial-condition branches that feed them.

C1: 2-D, 128^2 particles on a 128^2 mesh, ns = -1 with the Nyquist cutoff,
    normalised to an rms Zel'dovich displacement of 0.1 mesh cell at a0,
    a = 0.02 -> 1 in 50 steps of 0.0196.
C2: 3-D, 128^3 particles on a 128^3 mesh, ns = -2, a = 0.02 -> 0.9996 in 100
    steps of 0.0098; FP64 form against the oracle run on the box's cores.

Same inputs give bit-exact cell keys and deposits (with
PM_FLAG_SERIAL_DEPOSIT, each cell summed in the reference's serial order, as
the oracle does; the default fixed-point deposit within 1e-13) and forces
equal to FFT rounding; the two trajectories, which see different FFT roundings every
step, are compared by the rms density difference, the rms force difference
and P(k) per logarithmic bin (the oracle's estimator on both densities)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, n))


# ------------------------------------------------------------------- IC branches
@pytest.mark.parametrize("dim,mbits,cut", [(2, 6, False), (2, 6, True), (3, 5, False), (3, 5, True)])
def test_ic_cdm_matches_oracle(pm, orc, dim, mbits, cut):
    """conan ic without --scale-free: the CDM branch (src/ic/ic.cc:14-35,
    84-97), physical k, normalised by size res^(R/2)."""
    L = 200.0
    noise = orc.white_noise(9, (1 << mbits) ** dim)
    say("ic cdm", dim, mbits, cut)
    got = pm.ic_cdm(dim, mbits, L, noise, nyquist_cutoff=cut)
    want = orc.ic_cdm(dim, mbits, L, 0.96, 0.32, 69.0, noise, nyquist_cutoff=cut)
    assert rel(got, want) < 1e-12


@pytest.mark.parametrize("dim,mbits,smooth", [(2, 6, False), (2, 6, True), (3, 5, False)])
def test_ic_scale_free_nyquist_cutoff_matches_oracle(pm, orc, dim, mbits, smooth):
    L = 100.0
    noise = orc.white_noise(13, (1 << mbits) ** dim)
    say("ic scale-free, nyquist cutoff", dim, mbits, smooth)
    dens, phi = pm.ic_scale_free(dim, mbits, L, -1.0, noise, smooth=smooth, scale=1.0, nyquist_cutoff=True)
    want = orc.ic_scale_free(dim, mbits, L, -1.0, noise, smooth=smooth, scale=1.0, nyquist_cutoff=True)
    assert rel(dens, want) < 1e-12
    assert rel(phi, orc.compute_potential(dim, mbits, L, want)) < 1e-12
    plain = orc.ic_scale_free(dim, mbits, L, -1.0, noise, smooth=smooth, scale=1.0)
    assert rel(plain, want) > 1e-3  # the cutoff removes power


# ------------------------------------------------------------------- configurations
def _log_edges(Nf, L, nbins=64):
    kf, kn = 2 * np.pi / L, np.pi * Nf / L
    return kf * (kn / kf) ** (np.arange(nbins + 1) / nbins)


def _gpu_vs_oracle(pm, orc, dim, Nm, Nf, L, phi, a0, da, n):
    nt = threads()
    say("oracle run", dim, Nm, Nf, n, "steps on", nt, "threads")
    sim = orc.Sim(dim, Nm, Nf, L, nthreads=nt, serial_deposit=True)
    sim.init(phi, a0, da)
    sim.steps(n)
    Xo, Po, t = sim.state()
    a_pos = t[0]
    say("GPU run", dim, Nm, Nf, n, "steps")
    with pm.Gravity(dim, Nm, Nf, L) as g:
        g.init(phi, a0, a0 + da / 2)
        a = g.steps(a0, da, n)
        assert a == t[2] and g.times() == (t[0], t[1])
        X, P = g.get_state()
        rho = g.deposit()
        F = g.forces(a_pos)
        # the same inputs: the oracle's end state
        g.set_state(Xo, Po)
        rho_same, keys_same = g.deposit(keys=True)
        F_same = g.forces(a_pos)
    with pm.Gravity(dim, Nm, Nf, L, serial_deposit=True) as g:
        g.set_state(Xo, Po)
        rho_serial = g.deposit()
    m = (Nf / Nm) ** dim
    rho_o = orc.md_cic(dim, Nf, L, Xo, m)
    assert np.array_equal(keys_same, orc.cic_keys(dim, Nf, L, Xo))
    assert np.array_equal(rho_serial, rho_o)
    assert np.all(np.abs(rho_same - rho_o) <= 1e-13 * np.abs(rho_o) + 1e-24)
    F_o = orc.gather(dim, Nf, L, orc.poisson(dim, Nf, L, rho_o, a_pos, nthreads=nt), Xo, 0)
    edges = _log_edges(Nf, L)
    _, pk, c = orc.power_spectrum(dim, Nf, L, rho - 1, edges, nthreads=nt)
    _, pk_o, c_o = orc.power_spectrum(dim, Nf, L, rho_o - 1, edges, nthreads=nt)
    assert np.array_equal(c, c_o)
    b = (c > 0) & (pk_o > 0)
    dx = np.abs(X - Xo)
    dx = np.minimum(dx, np.abs(dx - L))
    out = {"force_same_rel": rel(F_same, F_o),
           "dx_max_cells": float(dx.max() / (L / Nf)),
           "rho_rms": float(np.sqrt(np.mean((rho - rho_o) ** 2))),
           "force_rms_rel": float(np.sqrt(np.mean((F - F_o) ** 2) / np.mean(F_o ** 2))),
           "pk_max_rel": float(np.max(np.abs(pk[b] / pk_o[b] - 1))), "bins": int(b.sum()),
           "mom_rel": rel(P, Po), "rho_max": float(rho_o.max())}
    say(out)
    return out


def test_config_c1_2d(pm, orc):
    """C1: 2-D 128^2 on 128^2, ns = -1, Nyquist cutoff, rms displacement
    a0 |grad phi| = 0.1 cell (the reference's ic is linear in --sigma, so one
    rescale sets it), f = 1, 50 steps a = 0.02 -> 1."""
    dim, Nm, mbits, L, a0, da, n = 2, 128, 7, 100.0, 0.02, 0.0196, 50
    noise = orc.white_noise(1, Nm ** dim)
    phi = orc.compute_potential(dim, mbits, L, orc.ic_scale_free(dim, mbits, L, -1.0, noise, sigma=1.0,
                                                                   nyquist_cutoff=True))
    u = orc.gradient_2nd_order(dim, Nm, L, phi).reshape(-1, dim)
    phi *= 0.1 * (L / Nm) / np.sqrt(np.mean(np.sum((a0 * u) ** 2, axis=1)))
    r = _gpu_vs_oracle(pm, orc, dim, Nm, Nm, L, phi, a0, da, n)
    # measured: 7e-15, 1.6e-13 cell, 2e-14, 6e-15, 7e-15
    assert r["force_same_rel"] < 1e-12
    assert r["dx_max_cells"] < 1e-10
    assert r["rho_rms"] < 1e-11
    assert r["force_rms_rel"] < 1e-11
    assert r["pk_max_rel"] < 1e-11


def test_config_c2_3d(pm, orc):
    """C2: 3-D 128^3 on 128^3, ns = -2, 100 steps a = 0.02 -> 0.9996."""
    dim, Nm, mbits, L, a0, da, n = 3, 128, 7, 100.0, 0.02, 0.0098, 100
    noise = orc.white_noise(2, Nm ** dim)
    nt = threads()
    phi = orc.compute_potential(dim, mbits, L, orc.ic_scale_free(dim, mbits, L, -2.0, noise, nthreads=nt),
                                nthreads=nt)
    r = _gpu_vs_oracle(pm, orc, dim, Nm, Nm, L, phi, a0, da, n)
    assert r["rho_max"] > 50  # clustered
    # measured: 2.4e-15, 3.8e-13 cell, 2.4e-14, 1.4e-14, 2.7e-15
    assert r["force_same_rel"] < 1e-12
    assert r["dx_max_cells"] < 1e-10
    assert r["rho_rms"] < 1e-11
    assert r["force_rms_rel"] < 1e-11
    assert r["pk_max_rel"] < 1e-11
