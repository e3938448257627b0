"""Checks on the gfx950 assembly of the device code (CPU: hipcc cross-compiles).

The gather fills its LDS ring with direct global->LDS loads issued from inline
assembly that writes M0 (conan_amd/csrc/pm_brick.hpp, lds_dma_dword).  The asm
lists M0 as clobbered, but the compiler treats M0 as a reserved register and
warns that it may not honour that.  The loads are safe as long as nothing else
in a kernel that issues them keeps a value in M0, so this test reads the
assembly of every such kernel and requires each M0 reference to be the asm's
own write or the direct load right after it.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _device_asm(tmp_path, src):
    out = tmp_path / "dev.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-inline-asm",
           "-Wno-unused-command-line-argument",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "conan_amd", "csrc"),
           "--cuda-device-only", "-S", os.path.join(ROOT, "conan_amd", "csrc", src), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    return out.read_text()


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_m0_is_only_written_by_the_direct_loads(tmp_path):
    asm = _device_asm(tmp_path, "pm_particles.hip")
    funcs = re.findall(r"^(\w+):[^\n]*\n(.*?)^\.Lfunc_end", asm, re.M | re.S)
    checked = 0
    for name, body in funcs:
        lines = [l.split(";")[0].strip() for l in body.split("\n")]
        lines = [l for l in lines if l]
        if not any(l.startswith("global_load_lds_dword") for l in lines):
            continue
        checked += 1
        for i, l in enumerate(lines):
            if not re.search(r"\bm0\b", l) and not l.startswith("global_load_lds"):
                continue
            if l.startswith("s_mov_b32 m0,"):
                # the asm block: the write, one wait state, the load
                assert lines[i + 1] == "s_nop 0" and lines[i + 2].startswith("global_load_lds_dword"), (name, l)
            elif l.startswith("global_load_lds_dword"):
                assert lines[i - 1] == "s_nop 0" and lines[i - 2].startswith("s_mov_b32 m0,"), (name, l)
            else:
                pytest.fail(f"{name}: M0 used outside the direct-load asm: {l}")
    # the FP32 and FP64 brick gathers, 2-D and 3-D, full and partial bricks
    assert checked >= 8
