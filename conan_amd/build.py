"""Builds the native parts of conan_amd in-tree with hipcc/g++.

Outputs (next to this file):
  libconanpm.so   HIP kernels for gfx950 + the C ABI of include/conanpm.h
  bin/conan       the ic / run / cic command-line tool (C++ host code over the C ABI)

Run as ``python -m conan_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libconanpm.so")
BIN = os.path.join(HERE, "bin", "conan")

HIP_SOURCES = ["pm_fft.hip", "pm_sort.hip", "pm_particles.hip", "pm_spectral.hip", "pm_adjoint.hip", "pm_api.hip",
               "pm_diff.hip", "pm_fof.hip"]
HOST_SOURCES = ["conan_cli.cpp", "conan_format.cpp"]

# -ffp-contract=off: no FMA contraction, so FP64 per-particle arithmetic
# rounds exactly like the reference's x86-64 build (make:21).  Packed FP32
# instructions are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-std=c++17", "-O2", "-ffp-contract=off", "-Wall"]
# The FFT is not a bit-matched part of the step (FFTW rounds differently in
# any case), so its multiply-adds may fuse.
FILE_FLAGS = {"pm_fft.hip": ["-ffp-contract=fast"]}


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build conan_amd")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_library(jobs: int = 8, verbose: bool = False) -> str:
    hipcc = _hipcc()
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([hipcc, *HIP_FLAGS, *FILE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> str:
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return ""
    tmp = BIN + ".tmp"
    _run(["g++", *HOST_FLAGS, "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp,
          "-L", HERE, "-lconanpm", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, BIN)
    return BIN


def build_oracle() -> str:
    """Test infrastructure: the CPU FP64 oracle under oracle/."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-B", "-C", odir])
    return os.path.join(odir, "liboracle.so")


def build_all(jobs: int = 8, verbose: bool = False) -> None:
    build_library(jobs, verbose)
    build_cli()
    build_oracle()


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB, BIN)
