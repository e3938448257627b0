"""CPU tests of the drop-in boundary: the C ABI library's exports and error
paths, the conan CLI's option handling and error behaviour, and the .conan
format.  Nothing here launches a kernel."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "conanpm.h")
CONAN = os.path.join(ROOT, "conan_amd", "bin", "conan")
HAS_GPU = os.path.exists("/dev/kfd")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from conan_amd import _lib

    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    assert sorted(_lib.EXPORTS) == names  # the Python binding covers the whole ABI


def _cfg(**kw):
    from conan_amd import _lib

    c = dict(dim=3, Nm=8, Nf=8, L=10.0, H0=68.5, Omega_m=1.0, Omega_L=0.0, precision=0, flags=0)
    c.update(kw)
    return _lib.pm_config(*[c[k] for k in ("dim", "Nm", "Nf", "L", "H0", "Omega_m", "Omega_L",
                                           "precision", "flags")])


@pytest.mark.parametrize("kw,code", [({"dim": 4}, -2), ({"dim": 1}, -2), ({"Nf": 12}, -2), ({"Nf": 2048}, -2),
                                     ({"L": 0.0}, -1), ({"L": float("nan")}, -1), ({"Nm": 0}, -1),
                                     ({"precision": 7}, -1), ({"flags": 8}, -1)])
def test_create_rejects_bad_configs(kw, code):
    import ctypes

    from conan_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.pm_create(ctypes.byref(_cfg(**kw)), ctypes.byref(h))
    assert rc == code and not h.value
    assert lib.pm_last_error(None)


def test_null_handle_and_null_outputs():
    from conan_amd import _lib

    lib = _lib.load()
    assert lib.pm_drift(None, 0.1, 0.01) == -1
    assert lib.pm_kick(None, 0.1, 0.01) == -1
    assert lib.pm_steps(None, 0.1, 0.01, 1, None) == -1
    assert b"null" in lib.pm_last_error(None)


@pytest.mark.skipif(HAS_GPU, reason="checks the no-device error path")
def test_create_without_device_fails_cleanly():
    import ctypes

    from conan_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.pm_create(ctypes.byref(_cfg()), ctypes.byref(h)) == -3
    assert b"no HIP device" in lib.pm_last_error(None)
    p = _lib.pm_ic_params(3, 4, 100.0, -2.0, 0, 8.0, 0.83)
    noise = np.zeros(16 ** 3)
    assert lib.pm_ic_scale_free(ctypes.byref(p), _lib.ptr(noise), _lib.ptr(noise), None) == -3


def _dparams(**kw):
    from conan_amd import _lib

    c = dict(dim=3, mbits=4, f=2, L=100.0, ns=-2.0, smooth=0, scale=8.0, sigma=0.83, H0=68.5, Omega_m=1.0,
             Omega_L=0.0, a0=0.05, da=0.05, nsteps=4, ckpt_every=2, precision=1, flags=0)
    c.update(kw)
    return _lib.pm_diff_params(*[c[k] for k, _ in _lib.pm_diff_params._fields_])


@pytest.mark.parametrize("kw,code", [({"dim": 4}, -2), ({"mbits": 11}, -2), ({"f": 3}, -2), ({"f": 128}, -2),
                                     ({"precision": 0}, -2), ({"nsteps": 0}, -1), ({"ckpt_every": 0}, -1),
                                     ({"da": 0.0}, -1), ({"sigma": -1.0}, -1), ({"flags": 4}, -1)])
def test_diff_create_rejects_bad_params(kw, code):
    import ctypes

    from conan_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.pm_diff_create(ctypes.byref(_dparams(**kw)), ctypes.byref(h)) == code and not h.value
    assert lib.pm_diff_last_error(None)
    assert lib.pm_diff_forward(None, None, None, 0) == -1
    assert lib.pm_diff_backward(None, None, None, 0) == -1


@pytest.mark.skipif(HAS_GPU, reason="checks the no-device error path")
def test_diff_create_without_device_fails_cleanly():
    import ctypes

    from conan_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.pm_diff_create(ctypes.byref(_dparams()), ctypes.byref(h)) == -3
    assert b"no HIP device" in lib.pm_diff_last_error(None)


@pytest.mark.parametrize("args,code", [((4, 10.0, 0.1), -2), ((1, 10.0, 0.1), -2), ((3, 0.0, 0.1), -1),
                                       ((3, float("inf"), 0.1), -1), ((3, 10.0, 0.0), -1),
                                       ((3, 10.0, float("nan")), -1)])
def test_fof_create_rejects_bad_arguments(args, code):
    import ctypes

    from conan_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.pm_fof_create(*args, ctypes.byref(h)) == code and not h.value
    assert lib.pm_fof_last_error(None)
    assert lib.pm_fof_run(None, 0, None, None, None, 0) == -1
    assert b"null" in lib.pm_fof_last_error(None)
    X = np.zeros((4, 3))
    g = np.zeros(4, dtype=np.int64)
    ng = ctypes.c_longlong(0)
    assert lib.pm_fof(*args[:1], 4, *args[1:], _lib.ptr(X), g.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                      ctypes.byref(ng)) == code


@pytest.mark.skipif(HAS_GPU, reason="checks the no-device error path")
def test_fof_without_device_fails_cleanly():
    import ctypes

    from conan_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.pm_fof_create(3, 10.0, 0.1, ctypes.byref(h)) == -3 and not h.value
    assert b"no HIP device" in lib.pm_fof_last_error(None)


# ------------------------------------------------------------------ CLI
def run_cli(*args, cwd=None):
    return subprocess.run([CONAN, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, cwd=cwd,
                          timeout=60)


def test_cli_usage_and_unknown_command():
    r = run_cli()
    assert r.returncode == 0 and "where [command] is one of: cic, fof, ic, run" in r.stdout
    r = run_cli("nv")
    assert r.returncode == 255
    assert "[error] command [nv] not understood." in r.stderr
    assert "usage:" in r.stdout


def test_cli_option_errors():
    r = run_cli("ic", "--bogus")
    assert r.returncode == 255 and r.stdout.strip() == "ERROR: Got the argument [--bogus], but no such option exists."
    r = run_cli("run", "-n")
    assert r.returncode == 255 and "Got the last argument [-n], but no value given." in r.stdout
    r = run_cli("cic", "stray")
    assert r.returncode == 255 and "Don't know how to handle [stray]" in r.stdout


def test_cli_help_layout():
    r = run_cli("run", "--help")
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    assert "        --help | -h (flag) : print help on the use of this program. " in lines
    assert "        --n | -n (default: 50) : number of iterations. " in lines
    assert any(l.startswith("        --f | -f (default: 2) : fractional force resolution.") for l in lines)
    # wrapped continuation lines are indented by 16 spaces and kept short
    assert all(len(l) < 100 for l in lines)
    assert any(l.startswith(" " * 16) and not l.startswith(" " * 17) for l in lines)


def test_cli_run_file_errors(tmp_path):
    from conan_amd import conanfile

    # missing input
    r = run_cli("run", "-i", "nope", cwd=tmp_path)
    assert r.returncode == 255 and "[error] Could not read from file" in r.stderr
    # corrupted framing
    p = tmp_path / "bad.density.init.conan"
    p.write_bytes(struct.pack("<Q", 5) + b"N=4\nx" + struct.pack("<Q", 6))
    r = run_cli("run", "-i", "bad", cwd=tmp_path)
    assert "[error] The block seems to be corrupted" in r.stderr and r.returncode == 255
    # no potential record: the record search runs off the end of the file
    conanfile.write(str(tmp_path / "nopot.density.init.conan"),
                    {"N": 4, "dim": 3, "size": 10, "f": 1}, [(0, "x")],
                    {"density": ("float64", np.zeros(64))})
    r = run_cli("run", "-i", "nopot", cwd=tmp_path)
    assert r.returncode == 255 and "[error] Could not read from file" in r.stderr


def test_cli_ic_rejects_bad_dim(tmp_path):
    r = run_cli("ic", "-d", "4", "-b", "3", "-i", "x", cwd=tmp_path)
    assert r.returncode == 255 and "[error] only 2 and 3 dimensions supported." in r.stderr


# ------------------------------------------------------------------ .conan format
def test_conanfile_layout_and_round_trip(tmp_path):
    from conan_amd import conanfile

    path = str(tmp_path / "t.conan")
    X = np.arange(24, dtype=np.float64).reshape(8, 3)
    conanfile.write(path, {"dim": 3, "N": 2, "L": 100, "a-pos": 0.1}, [(1700000000, "ran nbody")],
                    {"pos": ("3f8", X), "rho": ("float64", np.ones(8))})
    raw = open(path, "rb").read()
    hdr = b"L=100\nN=2\na-pos=0.1\ndim=3\n"  # std::map order: uppercase first
    assert raw[:8] == struct.pack("<Q", len(hdr)) and raw[8:8 + len(hdr)] == hdr
    off = 16 + len(hdr)
    hist = b"1700000000:ran nbody\n"
    assert raw[off:off + 8] == struct.pack("<Q", len(hist)) and raw[off + 8:off + 8 + len(hist)] == hist
    off += 16 + len(hist)
    sub = b"dtype=3f8\nname=pos\n"
    assert raw[off + 8:off + 8 + len(sub)] == sub
    h, hi, rec = conanfile.read(path)
    assert h["dim"] == "3" and hi == [(1700000000, "ran nbody")]
    assert np.array_equal(rec["pos"][1], X) and np.array_equal(rec["rho"][1], np.ones(8))


def test_conanfile_reads_thulsa_style_records(tmp_path):
    """The reference's Python writer puts type= instead of dtype= in record
    headers (scripts/thulsa.py:102-105); readers only need name=."""
    from conan_amd import conanfile

    path = tmp_path / "p.conan"
    with open(path, "wb") as f:
        conanfile.write_block(f, b"dim=2\nN=4\nsize=1.000000\ncosmology=Planck-nbody\nmbits=2\n")
        conanfile.write_block(f, b"1:python converted thing\n")
        conanfile.write_block(f, b"name=potential\nL=1.000000\nN=4\nrank=2\nshape=(4, 4)\ntype=float64\n")
        conanfile.write_block(f, np.arange(16, dtype=np.float64).tobytes())
    h, _, rec = conanfile.read(str(path))
    assert h["cosmology"] == "Planck-nbody"
    assert np.array_equal(rec["potential"][1], np.arange(16.0))


def test_cli_fof_options_and_input_names(tmp_path):
    r = run_cli("fof", "--help")
    assert r.returncode == 0
    for opt in ("--id | -i", "--time | -t (default: 1.0)", "--linking-length | -l (default: 0.1)",
                "--filename | -f", "--ascii"):
        assert opt in r.stdout
    # timed_filename (src/base/date.cc:34-50): round(t * 10000), 5 digits
    r = run_cli("fof", "-i", "x", "-t", "0.5", cwd=tmp_path)
    assert r.returncode == 255 and "could not open x.pos.05000.conan" in r.stderr
    r = run_cli("fof", "-i", "x", cwd=tmp_path)
    assert "could not open x.pos.10000.conan" in r.stderr
    r = run_cli("fof", "-f", "given.conan", cwd=tmp_path)
    assert "could not open given.conan" in r.stderr
