#!/usr/bin/env python3
"""Benchmark of the MI355X particle-mesh step (conan's `run` hot loop).

Workload (the configuration the headline metric is quoted on): 3-D PM,
256^3 particles on a 512^3 mesh, box L = 100, Einstein-de Sitter, scale-free
P(k) ~ k^-1.5 initial conditions from seeded white noise, Zel'dovich start at
a = 0.02, leapfrog in a with da = 0.0049 (200 steps reach a = 0.99).

A step is one leapfrog iteration (drift + deposit + Poisson solve + gather +
kick).  The timed region runs --steps steps of the FP32 form (32-bit
fixed-point positions, FP32 mesh/FFT/gather, FP64 momenta) with particles and
meshes resident in HBM; the same is measured for the FP64 plain form.  The
differentiable runs (noise -> final density and its adjoint, at 128^3 and
512^3 particles) give the gradient-to-forward time ratio.  One JSON line goes
to stdout; progress goes to stderr.

conan fof's grouping runs on the final snapshot (b = 0.2): particles/s.

Usage: python bench.py [--gpus 1] [--steps 200] [--warmup 5] [--no-cpu] [--no-cpu1] [--no-fp64] [--no-grad]
                       [--no-fof]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "particle-steps/s & %HBM peak at 256³/512³; FP32 vs FP64; grad/fwd time ratio"

DIM, MBITS, F, L, NS = 3, 8, 2, 100.0, -1.5
A0, DA = 0.02, 0.0049
NM = 1 << MBITS
NF = F * NM
NP = NM ** DIM
M = NF ** DIM


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def phase_bytes(prec: str) -> dict:
    """Algorithmic HBM bytes of one launch of each phase (reads + writes the
    step needs, not what an implementation moves).  Sum over one step:
    FP32 84 Np + 48 M, FP64 120 Np + 96 M."""
    b_pos = 4 if prec == "fp32" else 8  # position component
    b_mesh = 4 if prec == "fp32" else 8
    nrows = M // NF
    spec = 2 * (2 * b_mesh) * (M // 2)  # one read + one write of the M/2 complex modes
    return {
        1: DIM * b_pos * NP + b_mesh * M,  # deposit: positions in, mesh out
        2: b_mesh * M + b_mesh * M + 2 * b_mesh * nrows,  # rows r2c: real in, half spectrum out
        3: spec,
        4: spec,
        5: spec,
        6: b_mesh * M + b_mesh * M + 2 * b_mesh * nrows,  # rows c2r
        7: 2 * DIM * (b_pos + 8) * NP + b_mesh * M,  # gather+kick+drift: X, P in and out, phi in
    }


def step_bytes(prec: str) -> int:
    return (84 * NP + 48 * M) if prec == "fp32" else (120 * NP + 96 * M)


def read_traffic(kernel_tag: str):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None when absent)."""
    path = os.path.join(ROOT, "profiles", "traffic_pm_step_c3.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel_tag:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def run_form(pm, phi, prec, steps, warmup, wrap=False, profile=True, pk_bins=64, keep_state=False):
    g = pm.Gravity(DIM, NM, NF, L, precision=prec, wrap_gather=wrap)
    try:
        if warmup > 0:
            g.init(phi, A0, A0 + DA / 2)
            g.steps(A0, DA, warmup)
        g.init(phi, A0, A0 + DA / 2)
        g.sync()
        if profile:
            g.profile(1)
        # the first and the last 10 steps timed apart (nearly uniform vs clustered)
        chunks = [10, steps - 20, 10] if steps >= 30 else [steps]
        t0 = time.perf_counter()
        a, tc = A0, []
        for n in chunks:
            c0 = time.perf_counter()
            a = g.steps(a, DA, n)  # returns after the stream is synchronised
            tc.append((time.perf_counter() - c0) / n)
        t1 = time.perf_counter()
        ms, cnt = g.profile(0) if profile else (None, None)
        k, p, c = g.power_spectrum(pk_bins)
        diag = g.diagnostics()
        state = g.get_state() if keep_state else None
        out = {"seconds": t1 - t0, "a": a, "phase_ms": ms, "phase_cnt": cnt, "pk": (k, p, c), "diag": diag,
               "state": state}
        if len(tc) == 3:  # only a run of >= 30 steps has a first and a last 10 timed apart
            out["first_ms"], out["last_ms"] = tc[0] * 1e3, tc[-1] * 1e3
        return out
    finally:
        g.close()


def cpu_baseline(phi, nsteps: int, threads: int, late=None):
    """The oracle (FP64 CPU restatement of the reference, OpenMP loops and the
    per-add critical section of the reference's deposit) on a bounded sample:
    the first step(s) of the run, and, given late = (X, P, a), one step from
    that clustered state."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as orc

    sim = orc.Sim(DIM, NM, NF, L, nthreads=threads)
    sim.init(phi, A0, DA)
    secs = sim.steps(nsteps)  # steady_clock around the leapfrog loop only
    out = {"value": NP * nsteps / secs, "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": f"{nsteps} leapfrog step(s) (drift+deposit+FFT Poisson+gather+kick) of the same "
                     f"256^3/512^3 FP64 run from the same potential, {threads} OpenMP threads, "
                     f"{secs:.1f} s"}
    if late is not None:
        X, P, a = late
        sim.set_state(X, P)
        t0 = time.perf_counter()
        sim.drift(a, DA)
        sim.kick(a + DA / 2, DA)
        s2 = time.perf_counter() - t0
        out["late_value"] = NP / s2
        regime = "clustered" if a > 0.5 else "still close to the linear regime"
        out["late_sample"] = (f"1 step from the end state of the timed FP32 run, a = {a:.3f} ({regime}), "
                              f"{threads} OpenMP threads, {s2:.1f} s")
    return out


FOF_B = 0.2  # linking length in units of the mean particle spacing
FOF_BYTES = 32  # per particle: read its position (3 doubles), write its group (int64)


def fof_bench(pm, X, runs=5, cpu=True):
    """conan fof's grouping (pm_fof) on the final snapshot of the timed FP32
    run (256^3 particles, a ~ 1), linking length b = 0.2 of the mean spacing,
    positions resident in HBM; the oracle's flood fill timed on a slab of the
    same snapshot as the CPU baseline."""
    import torch

    ll = FOF_B * L / NM
    plan = pm.FofPlan(DIM, L, ll)
    try:
        Xd = torch.from_numpy(np.ascontiguousarray(X)).to("cuda")
        g, ng = plan.run(Xd)  # warm-up: buffers allocated here
        dev_ms, wall = [], []
        for _ in range(runs):
            t0 = time.perf_counter()
            g, ng = plan.run(Xd)
            wall.append(time.perf_counter() - t0)
            dev_ms.append(plan.stats()["device_ms"])
        st = plan.stats()
        mass = torch.bincount(g).cpu().numpy()
    finally:
        plan.close()
    t = float(np.median(dev_ms)) * 1e-3
    out = {"workload": f"friends-of-friends, b = {FOF_B} (ll = {ll:.6g}), {NP} particles of the final "
                       "FP32 snapshot", "particles": NP, "ngroups": int(ng), "groups_ge_10": int((mass >= 10).sum()),
           "largest": int(mass.max()), "ms": round(t * 1e3, 4), "wall_ms": round(float(np.median(wall)) * 1e3, 4),
           "value": round(NP / t, 1), "unit": "particles/s",
           "candidates_per_particle": round(st["candidates"] / NP, 3), "links_per_particle": round(st["links"] / NP, 3),
           "cells": st["cells"], "table_bits": st["table_bits"],
           "hbm_frac_min_bytes": round(FOF_BYTES * NP / t / PEAK_HBM, 5)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle as orc

        sl = np.mod(X[:, 0], L) < L / 8
        Xs = np.ascontiguousarray(X[sl])
        t0 = time.perf_counter()
        _, ngs = orc.fof(DIM, L, ll, Xs)
        secs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(len(Xs) / secs, 1), "unit": "particles/s", "cores": 1, "kind": "port",
                               "sample": f"oracle flood fill (cell grid) over the {len(Xs)} particles of the slab "
                                         f"x < L/8 of the same snapshot, {ngs} groups, {secs:.1f} s"}
    return out


# Differentiable runs: (name, mbits, f, ns, a0, da, nsteps, ckpt_every, store_potentials)
DIFF_CONFIGS = [
    ("128^3 particles on 128^3, ns=-2, 20 steps a=0.05->1, checkpoints every 5", 7, 1, -2.0, 0.05, 0.0475, 20, 5,
     True),
    ("512^3 particles on 512^3, ns=-1.5, 40 steps a=0.05->1, checkpoints every 8, potentials recomputed", 9, 1, -1.5,
     0.05, 0.02375, 40, 8, False),
    ("512^3 particles on 512^3, ns=-1.5, 40 steps a=0.05->1, checkpoints every 8, potentials stored", 9, 1, -1.5,
     0.05, 0.02375, 40, 8, True),
]


def diff_bench(name, mbits, f, ns, a0, da, nsteps, ckpt, store, repeats=2):
    """Forward (noise -> final density) and backward (dL/drho -> dL/dnoise) of
    the differentiable FP32 run, loss = sum (rho - rho_target)^2 with the
    target from a second seed, all buffers on the device.  store: keep every
    step's potential from the forward pass, or recompute it in the backward
    pass from the reversed state."""
    import torch

    from conan_amd.diff import DiffPM

    Np = (1 << mbits) ** DIM
    eng = DiffPM(DIM, mbits, f, L, ns, a0, da, nsteps, ckpt_every=ckpt, store_potentials=store)
    try:
        w = torch.tensor(np.random.default_rng(1).standard_normal(Np), device="cuda")
        target = eng.forward_device(torch.tensor(np.random.default_rng(2).standard_normal(Np), device="cuda"))
        best = None
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rho = eng.forward_device(w)
            t1 = time.perf_counter()
            g = eng.backward_device(2 * (rho - target))
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            st = eng.stats()
            if best is None or st["forward_ms"] + st["backward_ms"] < best["forward_ms"] + best["backward_ms"]:
                best = dict(st, wall_forward_s=t1 - t0, wall_backward_s=t2 - t1)
        assert bool(torch.isfinite(g).all())
        f_ms, b_ms = best["forward_ms"], best["backward_ms"]
        return {"workload": name, "particles": Np, "mesh": f << mbits, "steps": nsteps, "ckpt_every": ckpt,
                "potentials_stored": store, "checkpoint_check": "checksum",
                "forward_ms": round(f_ms, 3), "backward_ms": round(b_ms, 3),
                "backward_over_forward": round(b_ms / f_ms, 3),
                "grad_fwd_ratio": round((f_ms + b_ms) / f_ms, 3),
                "device_GB": round(best["device_bytes"] / 1e9, 2), "checkpoints": best["ncheckpoints"],
                "reversal_position_mismatches": best["recon_pos"], "reversal_momentum_rel": best["recon_mom"]}
    finally:
        eng.close()


def copy_probe(nbytes=1 << 30, reps=5):
    """Device-to-device copy of 1 GiB (torch's copy kernel): the practical
    streaming bandwidth of this GPU, and the reference kernel of known size
    for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return {"bytes_per_copy": nbytes, "ms": round(ms, 4), "GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu1", action="store_true", help="skip the single-thread CPU baseline")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-grad", action="store_true", help="skip the differentiable runs")
    ap.add_argument("--no-fof", action="store_true", help="skip the friends-of-friends run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box here has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from conan_amd import pm

    log(f"IC: {NM}^3 white noise (seed 1), scale-free ns={NS}, FP64 on the GPU")
    noise = np.random.default_rng(1).standard_normal(NP)
    _, phi = pm.ic_scale_free(DIM, MBITS, L, NS, noise)
    del noise

    log(f"FP32 form: {args.warmup} warmup + {args.steps} timed steps")
    r32 = run_form(pm, phi, "fp32", args.steps, args.warmup, keep_state=not (args.no_fof and args.no_cpu))
    log(f"  {r32['seconds']:.3f} s -> {r32['seconds'] / args.steps * 1e3:.3f} ms/step")
    r64 = None
    if not args.no_fp64:
        log(f"FP64 plain form: {args.warmup} warmup + {args.steps} timed steps")
        r64 = run_form(pm, phi, "fp64", args.steps, args.warmup)
        log(f"  {r64['seconds']:.3f} s -> {r64['seconds'] / args.steps * 1e3:.3f} ms/step")

    from conan_amd._lib import PHASE_NAMES

    def phase_report(r, prec):
        pb = phase_bytes(prec)
        out = {}
        for i, name in enumerate(PHASE_NAMES):
            if r["phase_cnt"][i] == 0:
                continue
            avg = r["phase_ms"][i] / r["phase_cnt"][i] * 1e-3
            e = {"ms_per_launch": round(avg * 1e3, 4), "launches": int(r["phase_cnt"][i])}
            if i in pb:
                e["GBps"] = round(pb[i] / avg / 1e9, 1)
                e["frac_hbm"] = round(pb[i] / avg / PEAK_HBM, 4)
            out[name] = e
        return out

    ph32 = phase_report(r32, "fp32")
    for k, v in ph32.items():
        log(f"  fp32 {k:24s} {v}")

    # dominant single-launch kernel (phases 1..7) by total time
    pb = phase_bytes("fp32")
    dom = max(pb, key=lambda i: r32["phase_ms"][i])
    dom_avg = r32["phase_ms"][dom] / r32["phase_cnt"][dom] * 1e-3
    tags = {1: "k_deposit", 2: "k_rows_r2c", 3: "k_cols_fwd", 4: "k_cols_green", 5: "k_cols_bwd", 6: "k_rows_c2r",
            7: "k_gather_kick"}
    achieved = pb[dom] / dom_avg
    roofline = {"bound": "hbm", "kernel": f"{tags[dom]} ({PHASE_NAMES[dom]}), FP32, 512^3",
                "achieved": round(achieved / 1e9, 1), "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM, 4), "traffic": read_traffic(tags[dom]),
                "algorithmic_bytes_per_launch": pb[dom], "avg_launch_ms": round(dom_avg * 1e3, 4)}

    t32 = r32["seconds"] / args.steps
    val32 = NP / t32
    step_roof = {"bytes_per_step": step_bytes("fp32"), "achieved_GBps": round(step_bytes("fp32") / t32 / 1e9, 1),
                 "frac": round(step_bytes("fp32") / t32 / PEAK_HBM, 4)}
    out = {
        "metric": METRIC, "value": round(val32, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t32 * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic: seeded Gaussian white noise -> scale-free GRF (ns=-1.5) -> Zel'dovich start",
        "config": {"workload": "3-D PM N-body, 256^3 particles on a 512^3 mesh, L=100, EdS, "
                               f"a=0.02 + {args.steps} x 0.0049, FP32 form (fixed-point positions, "
                               "FP64 momenta)",
                   "particles": NP, "mesh": NF, "form": "fp32"},
        "roofline": roofline,
        "step_roofline": step_roof,
        "phases_fp32": ph32,
    }
    if "first_ms" in r32:
        out["ms_per_step_first10"] = round(r32["first_ms"], 4)
        out["ms_per_step_last10"] = round(r32["last_ms"], 4)
    if r64 is not None:
        t64 = r64["seconds"] / args.steps
        out["fp64"] = {"value": round(NP / t64, 1), "ms_per_step": round(t64 * 1e3, 4),
                       "step_roofline_frac": round(step_bytes("fp64") / t64 / PEAK_HBM, 4),
                       "phases": phase_report(r64, "fp64")}
        if "first_ms" in r64:
            out["fp64"]["ms_per_step_first10"] = round(r64["first_ms"], 4)
            out["fp64"]["ms_per_step_last10"] = round(r64["last_ms"], 4)
        out["fp32_over_fp64"] = round(t64 / t32, 3)
        # both forms run the reference's operator (the gather truncates the
        # unwrapped coordinate), from the same potential, to the same a
        k32, p32, c32 = r32["pk"]
        k64, p64, c64 = r64["pk"]
        m = (c32 > 0) & (p64 > 0)
        rel = np.abs(p32[m] / p64[m] - 1)
        out["pk_fp32_vs_fp64_plain_max_rel"] = float(rel.max())
        out["pk_fp32_vs_fp64"] = {"a": r64["a"], "bins": int(m.sum()), "max_rel": float(rel.max()),
                                  "median_rel": float(np.median(rel)), "gather": "reference (unwrapped, truncated)"}
    out["copy_probe"] = copy_probe()
    out["grad_fwd_ratio"] = None
    if not args.no_grad:
        runs = []
        for cfg in DIFF_CONFIGS:
            log(f"differentiable run: {cfg[0]}")
            runs.append(diff_bench(*cfg))
            log(f"  {runs[-1]}")
        # (forward + adjoint) / forward of the 512^3 differentiable run with the
        # potentials recomputed (checkpointing only, no stored trajectory)
        out["grad_fwd_ratio"] = runs[1]["grad_fwd_ratio"]
        out["differentiable"] = runs
    if not args.no_fof:
        log(f"friends-of-friends on the final FP32 snapshot (b = {FOF_B})")
        out["fof"] = fof_bench(pm, r32["state"][0], cpu=not args.no_cpu)
        log(f"  {out['fof']}")
    out["final_a"] = r32["a"]
    out["momentum_sum_fp32"] = r32["diag"]["momentum"]

    if not args.no_cpu:
        # all the host threads this process may use (the GPU box gives one
        # GPU's share, 16), and one thread
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                                          else (os.cpu_count() or 1))
        log(f"CPU baseline: oracle, {args.cpu_steps} step(s), {threads} threads")
        X32, P32 = r32["state"]
        out["cpu_baseline"] = cpu_baseline(phi, args.cpu_steps, threads, late=(X32, P32, r32["a"]))
        log(f"  {out['cpu_baseline']['sample']}")
        if threads > 1 and not args.no_cpu1:
            log("CPU baseline: oracle, 1 step, 1 thread")
            one = cpu_baseline(phi, 1, 1)
            out["cpu_baseline"]["one_thread"] = {k: one[k] for k in ("value", "unit", "cores", "kind", "sample")}
            log(f"  {one['sample']}")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
