"""The conan CLI end to end on the GPU: ic -> run -> cic with the reference's
options and .conan files, checked against the oracle run from the same stored
white noise."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONAN = os.path.join(ROOT, "conan_amd", "bin", "conan")


def run(args, cwd):
    print("[case] conan", " ".join(args), flush=True)
    r = subprocess.run([CONAN, *args], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r


@pytest.fixture(scope="module")
def sim_dir(tmp_path_factory, orc):
    from conan_amd import conanfile

    d = tmp_path_factory.mktemp("cli")
    dim, mbits, L = 3, 4, 100.0
    noise = orc.white_noise(5, (1 << mbits) ** dim)
    conanfile.write_noise(str(d / "noise.conan"), noise, dim, mbits)
    run(["ic", "-i", "t", "-d", "3", "-b", str(mbits), "-L", "100", "--scale-free", "-n", "-2", "-p",
         "--noise", "noise.conan"], d)
    return d, noise


def test_ic_file_matches_oracle(sim_dir, orc):
    from conan_amd import conanfile

    d, noise = sim_dir
    h, hist, rec = conanfile.read(str(d / "t.density.init.conan"))
    # every option is in the header (Header << Argv), plus N (ic/main.cc:74)
    for k in ("dim", "mbits", "size", "seed", "ns", "scale-free", "potential", "N", "id", "H0", "Omega0"):
        assert k in h
    assert h["N"] == "16" and h["scale-free"] == "1" and h["smooth"] == "false"
    assert len(hist) == 1 and hist[0][1].startswith("ic ")
    dens = orc.ic_scale_free(3, 4, 100.0, -2.0, noise)
    phi = orc.compute_potential(3, 4, 100.0, dens)
    assert rec["density"][0]["dtype"] == "float64"
    assert np.max(np.abs(rec["density"][1] - dens)) < 1e-12 * np.max(np.abs(dens))
    assert np.max(np.abs(rec["potential"][1] - phi)) < 1e-12 * np.max(np.abs(phi))


def test_ic_seed_matches_oracle_noise(sim_dir, orc):
    """conan ic --seed S draws the white noise with std::mt19937(S) and
    std::normal_distribution<double> in linear order (Gaussian_white_noise,
    src/base/misc.hh:179-188; ic.cc:55-57), as the oracle's white_noise does."""
    from conan_amd import conanfile

    d, _ = sim_dir
    run(["ic", "-i", "s", "-d", "3", "-b", "4", "-L", "100", "--scale-free", "-n", "-2", "--seed", "4242"], d)
    h, _, rec = conanfile.read(str(d / "s.density.init.conan"))
    assert h["seed"] == "4242"
    want = orc.ic_scale_free(3, 4, 100.0, -2.0, orc.white_noise(4242, 16 ** 3))
    assert np.max(np.abs(rec["density"][1] - want)) < 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("cut", [False, True])
def test_ic_cdm_file_matches_oracle(sim_dir, orc, cut):
    """conan ic without --scale-free (the CDM branch, src/ic/ic.cc:84-97),
    with and without --nyquist-cutoff."""
    from conan_amd import conanfile

    d, noise = sim_dir
    run(["ic", "-i", "c", "-d", "3", "-b", "4", "-L", "100", "--noise", "noise.conan"]
        + (["--nyquist-cutoff"] if cut else []), d)
    h, hist, rec = conanfile.read(str(d / "c.density.init.conan"))
    assert h["scale-free"] == "false" and h["nyquist-cutoff"] == ("1" if cut else "false")
    want = orc.ic_cdm(3, 4, 100.0, 0.96, 0.32, 69.0, noise, nyquist_cutoff=cut)
    assert np.max(np.abs(rec["density"][1] - want)) < 1e-12 * np.max(np.abs(want))


def test_run_snapshots_match_oracle(sim_dir, orc):
    from conan_amd import conanfile

    d, noise = sim_dir
    run(["run", "-i", "t", "-f", "2", "--a0", "0.02", "--da", "0.0098", "-n", "20"], d)
    assert (d / "t.pos.1.conan").exists() and (d / "t.pos.2.conan").exists()
    h, hist, rec = conanfile.read(str(d / "t.pos.2.conan"))
    # save_conan's header and records (src/run/nbody.hh:168-183)
    assert sorted(h) == ["L", "N", "a-mom", "a-pos", "dim", "id"]
    assert h["N"] == "16" and h["L"] == "100" and h["dim"] == "3" and h["id"] == "t"
    assert hist[0][1] == "ran nbody"
    assert rec["pos"][0]["dtype"] == "3f8" and rec["pos"][1].shape == (16 ** 3, 3)
    _, _, irec = conanfile.read(str(d / "t.density.init.conan"))
    sim = orc.Sim(3, 16, 32, 100.0)
    sim.init(irec["potential"][1], 0.02, 0.0098)
    sim.steps(20)
    Xo, Po, t = sim.state()
    assert float(h["a-pos"]) == pytest.approx(t[0], rel=1e-5)
    assert float(h["a-mom"]) == pytest.approx(t[1], rel=1e-5)
    assert np.max(np.abs(rec["pos"][1] - Xo)) < 1e-9 * 100
    assert np.max(np.abs(rec["mom"][1] - Po)) < 1e-9 * np.max(np.abs(Po))


def test_cic_tool_matches_oracle(sim_dir, orc):
    from conan_amd import conanfile

    d, _ = sim_dir
    if not (d / "t.pos.2.conan").exists():
        run(["run", "-i", "t", "-f", "2", "--a0", "0.02", "--da", "0.0098", "-n", "20"], d)
    run(["cic", "-i", "t", "-m", "2", "-f", "2"], d)
    h, hist, rec = conanfile.read(str(d / "t.rho.2.conan"))
    assert h["N"] == "32" and h["sub-sample"] == "1"
    assert [c for _, c in hist] == ["ran nbody", "cic -i t -m 2 -f 2"]
    _, _, prec = conanfile.read(str(d / "t.pos.2.conan"))
    ref = orc.cic_tool(3, 16, 32, 100.0, 1, prec["pos"][1])
    assert np.array_equal(rec["density"][1], ref)
    run(["cic", "-i", "t", "-m", "2", "-f", "2", "-s", "2"], d)
    _, _, rec2 = conanfile.read(str(d / "t.rho.2.conan"))
    ref2 = orc.cic_tool(3, 16, 32, 100.0, 2, prec["pos"][1])
    assert np.max(np.abs(rec2["density"][1] - ref2)) < 1e-12 * np.max(ref2)


def test_ifrit_and_ascii_outputs(sim_dir):
    """The IFRIT and ASCII snapshots of a run hold the values of its .conan
    snapshot: IFRIT (src/run/nbody.hh:135-166) uint32-framed blocks of the
    count, the box (0, 0, 0, L, L, L) and the six float32 arrays X_k, P_k;
    ASCII (nbody.hh:204-208) one line per particle, `X P/a` with the
    default stream format and the vectors' space-terminated components."""
    import struct

    from conan_amd import conanfile

    d, _ = sim_dir
    run(["run", "-i", "t", "-f", "1", "--da", "0.0098", "-n", "10"], d)
    _, _, rec = conanfile.read(str(d / "t.pos.1.conan"))
    X, P = rec["pos"][1], rec["mom"][1]
    run(["run", "-i", "t", "-f", "1", "--da", "0.0098", "-n", "10", "--ifrit"], d)
    raw = (d / "t.pos.1.bin").read_bytes()
    assert len(raw) == (4 + 4 + 4) + (4 + 24 + 4) + 6 * (4 + 4096 * 4 + 4)
    off = 0

    def block():
        nonlocal off
        n = struct.unpack_from("<I", raw, off)[0]
        body = raw[off + 4: off + 4 + n]
        assert struct.unpack_from("<I", raw, off + 4 + n)[0] == n  # framed on both sides
        off += 8 + n
        return body

    assert struct.unpack("<I", block())[0] == 4096
    assert np.array_equal(np.frombuffer(block(), "<f4"), np.array([0, 0, 0, 100, 100, 100], np.float32))
    for A in (X, P):
        for k in range(3):
            assert np.array_equal(np.frombuffer(block(), "<f4"), A[:, k].astype(np.float32))
    run(["run", "-i", "t", "-f", "1", "--da", "0.0098", "-n", "10", "--ascii"], d)
    lines = (d / "t.pos.1.txt").read_text().splitlines()
    a = 0.02
    for _ in range(10):  # LeapFrog::run's a after the 10th step (src/run/leapfrog.hh:38-44)
        a += 0.0098
    want = ["".join(f"{v:g} " for v in X[i]) + " " + "".join(f"{v / a:g} " for v in P[i]) for i in range(len(X))]
    assert lines == want


def test_concordance_run_matches_oracle(sim_dir, orc):
    """run --concordance (H0 68.5, Omega_m 0.31, Omega_L 0.69, src/run/main.cc:
    94-101): a LambdaCDM expansion in every drift and kick factor."""
    from conan_amd import conanfile

    d, _ = sim_dir
    run(["run", "-i", "t", "-f", "2", "--a0", "0.05", "--da", "0.02", "-n", "10", "--concordance"], d)
    _, _, rec = conanfile.read(str(d / "t.pos.1.conan"))
    _, _, irec = conanfile.read(str(d / "t.density.init.conan"))
    sim = orc.Sim(3, 16, 32, 100.0, H0=68.5, Om=0.31, OL=0.69)
    sim.init(irec["potential"][1], 0.05, 0.02)
    sim.steps(10)
    Xo, Po, _ = sim.state()
    assert np.max(np.abs(rec["pos"][1] - Xo)) < 1e-9 * 100
    assert np.max(np.abs(rec["mom"][1] - Po)) < 1e-9 * np.max(np.abs(Po))
    # and it differs from the Einstein-de Sitter run
    sim = orc.Sim(3, 16, 32, 100.0)
    sim.init(irec["potential"][1], 0.05, 0.02)
    sim.steps(10)
    assert np.max(np.abs(sim.state()[1] - Po)) > 1e-3 * np.max(np.abs(Po))


def test_fp32_precision_flag(sim_dir):
    from conan_amd import conanfile

    d, _ = sim_dir
    run(["run", "-i", "t", "-f", "2", "--a0", "0.02", "--da", "0.0098", "-n", "10"], d)
    X64 = conanfile.read(str(d / "t.pos.1.conan"))[2]["pos"][1]
    run(["run", "-i", "t", "-f", "2", "--a0", "0.02", "--da", "0.0098", "-n", "10", "--precision", "fp32"], d)
    h, _, rec = conanfile.read(str(d / "t.pos.1.conan"))
    X = rec["pos"][1]
    # the fixed-point form keeps unwrapped coordinates too, and tracks the FP64 run
    assert X.min() < 0 and X64.min() < 0
    assert np.sqrt(np.mean((X - X64) ** 2)) < 1e-4 * 100.0 / 16


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_resume_continues_bit_for_bit(sim_dir, prec):
    """run --resume <id>.pos.1.conan repeats no step and ends in the same
    snapshot as the uninterrupted run (FP32: the momentum grid travels in the
    snapshot header)."""
    import shutil

    from conan_amd import conanfile

    d, _ = sim_dir
    args = ["run", "-i", "t", "-f", "2", "--a0", "0.02", "--da", "0.0098", "-n", "30", "--precision", prec]
    run(args, d)
    full = conanfile.read(str(d / "t.pos.3.conan"))
    shutil.copy(d / "t.pos.1.conan", d / "r.pos.1.conan")
    (d / "t.pos.2.conan").unlink()
    (d / "t.pos.3.conan").unlink()
    out = run(args + ["--resume", "r.pos.1.conan"], d)
    assert out.stdout.count(".") == 20  # steps 11 .. 30 only
    res = conanfile.read(str(d / "t.pos.3.conan"))
    if prec == "fp32":
        assert "momentum-quantum-log2" in full[0]
        assert res[0]["momentum-quantum-log2"] == full[0]["momentum-quantum-log2"]
    for name in ("pos", "mom"):
        assert np.array_equal(res[2][name][1], full[2][name][1])
    assert res[0]["a-pos"] == full[0]["a-pos"] and res[0]["a-mom"] == full[0]["a-mom"]


def test_resume_rejects_a_foreign_snapshot(sim_dir):
    d, _ = sim_dir
    run(["run", "-i", "t", "-f", "2", "--a0", "0.02", "--da", "0.0098", "-n", "10"], d)
    r = subprocess.run([CONAN, "run", "-i", "t", "-f", "2", "--a0", "0.05", "--da", "0.0098", "-n", "20",
                        "--resume", "t.pos.1.conan"], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode != 0 and "a-pos" in (r.stdout + r.stderr)


def test_run_midpoint_matches_oracle(sim_dir, orc):
    from conan_amd import conanfile

    d, _ = sim_dir
    run(["run", "-i", "t", "-f", "2", "--a0", "0.02", "--da", "0.0098", "-n", "10", "--midpoint"], d)
    h, hist, rec = conanfile.read(str(d / "t.pos.1.conan"))
    _, _, irec = conanfile.read(str(d / "t.density.init.conan"))
    sim = orc.Sim(3, 16, 32, 100.0, midpoint=True)
    sim.init(irec["potential"][1], 0.02, 0.0098)
    sim.steps(10)
    Xo, Po, t = sim.state()
    assert np.max(np.abs(rec["pos"][1] - Xo)) < 1e-9 * 100
    assert np.max(np.abs(rec["mom"][1] - Po)) < 1e-9 * np.max(np.abs(Po))
