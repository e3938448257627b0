"""GPU parity of the HIP path against the oracle (CPU FP64 restatement).

Every comparison calls libconanpm.so through its C ABI.  Integer results
(cell keys) are compared bit for bit; FP64 quantities that the GPU computes
with the reference's own operation order (deposit, gather, Zel'dovich start)
are compared bit for bit too; quantities that pass through the FFT are
compared within the stated relative tolerances, because the reference's
FFTW and this library's Stockham FFT round differently.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(12345)


def say(*a):
    print("[case]", *a, flush=True)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("dim,N", [(2, 8), (2, 16), (2, 32), (2, 64), (2, 256), (3, 4), (3, 8), (3, 16), (3, 32),
                                   (3, 64)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
def test_fft_r2c_matches_numpy(pm, dim, N, prec, tol):
    say("fft r2c", dim, N, prec)
    x = RNG.standard_normal(N ** dim)
    got = pm.fft_r2c(x, dim, prec)
    # numpy with axis 0 fastest: reshape reversed, rfft along the last numpy axis
    ref = np.fft.rfftn(x.reshape((N,) * dim))  # numpy axes (z,y,x) -> last axis is x
    ref = ref.reshape(-1, N // 2 + 1)
    assert rel(got, ref) < tol
    back = pm.fft_c2r(got, dim, N, prec) / N ** dim
    assert rel(back, x) < tol


# ------------------------------------------------------------------- keys / deposit
def _positions(dim, Nm, L, disp, special=True):
    q = np.stack(np.meshgrid(*([np.arange(Nm) * (L / Nm)] * dim), indexing="ij"), -1)
    q = q.reshape(-1, dim)[:, ::-1].copy()  # axis 0 fastest
    X = q + disp * RNG.standard_normal(q.shape)
    if special:
        n = min(64, X.shape[0])
        cases = np.array([0.0, -0.0, L, -L, 2 * L, L - 1e-13, -1e-300, 1e-300, L * (1 - 2 ** -52),
                          -L / 3, 7 * L / 3, L / 2, -2 * L - 1e-9, 0.5 * L / Nm, 3.0, -3.0])
        X[:n] = RNG.choice(cases, size=(n, dim))
    return X


def _check_fp64_deposit(rho, ro, serial):
    if serial:
        assert np.array_equal(rho, ro)  # same weights, summed in the same particle order
    else:
        # exact sum of the weights rounded to 2^-96, against the serial double sum
        err = np.abs(rho - ro) / np.maximum(np.abs(ro), 1e-300)
        say(f"  fixed-point deposit vs serial sum: max rel {float(err[ro > 1e-12].max()):.3g}")
        assert np.all(np.abs(rho - ro) <= 1e-13 * np.abs(ro) + 1e-24)


@pytest.mark.parametrize("serial", [True, False])
@pytest.mark.parametrize("dim,Nm,Nf", [(2, 16, 16), (2, 32, 64), (3, 8, 8), (3, 16, 32), (3, 32, 32)])
def test_keys_and_deposit_fp64(pm, orc, dim, Nm, Nf, serial):
    """Cell keys bit-exact; the deposit bit-exact with PM_FLAG_SERIAL_DEPOSIT
    (the reference's serial summation order), else the fixed-point sum
    within 1e-13 of it."""
    L = 100.0
    X = _positions(dim, Nm, L, 3.0 * L / Nm)
    say("deposit fp64", dim, Nm, Nf, serial)
    with pm.Gravity(dim, Nm, Nf, L, serial_deposit=serial) as g:
        g.set_state(X)
        rho, keys = g.deposit(keys=True)
    m = (Nf / Nm) ** dim
    ko = orc.cic_keys(dim, Nf, L, X)
    ro = orc.md_cic(dim, Nf, L, X, m)
    assert np.array_equal(keys, ko)
    _check_fp64_deposit(rho, ro, serial)


@pytest.mark.parametrize("serial", [True, False])
@pytest.mark.parametrize("dim,Nm,Nf", [(2, 32, 32), (3, 16, 32), (3, 32, 32)])
def test_deposit_clustered_fp64(pm, orc, dim, Nm, Nf, serial):
    """Many particles per cell (late-time clustering): the merge order still
    reproduces the serial sum exactly; the fixed-point sum stays within
    1e-13 of it."""
    L = 10.0
    X = L * 0.5 + 0.02 * L * RNG.standard_normal((Nm ** dim, dim))
    say("deposit clustered", dim, Nm, Nf, serial)
    with pm.Gravity(dim, Nm, Nf, L, serial_deposit=serial) as g:
        g.set_state(X)
        rho = g.deposit()
    ro = orc.md_cic(dim, Nf, L, X, (Nf / Nm) ** dim)
    _check_fp64_deposit(rho, ro, serial)
    assert abs(rho.sum() - Nm ** dim * (Nf / Nm) ** dim) < 1e-9 * Nm ** dim


def test_deposit_fp32_mass_and_density(pm, orc):
    dim, Nm, Nf, L = 3, 32, 64, 100.0
    X = _positions(dim, Nm, L, 2.0, special=False)
    say("deposit fp32")
    with pm.Gravity(dim, Nm, Nf, L, precision="fp32") as g:
        g.set_state(X)
        rho, keys = g.deposit(keys=True)
    Xw = np.mod(X, L)
    ro = orc.md_cic(dim, Nf, L, Xw, (Nf / Nm) ** dim)
    assert abs(rho.sum() / ro.sum() - 1) < 1e-6
    assert np.max(np.abs(rho - ro)) < 1e-5 * max(1.0, ro.max())


# ------------------------------------------------------------------- gather
@pytest.mark.parametrize("wrap", [0, 1])
@pytest.mark.parametrize("dim,Nm,Nf", [(2, 32, 32), (3, 16, 32)])
def test_gather_bit_exact_fp64(pm, orc, dim, Nm, Nf, wrap):
    L = 50.0
    X = _positions(dim, Nm, L, 2 * L / Nm)
    phi = RNG.standard_normal(Nf ** dim)
    say("gather", dim, Nm, Nf, wrap)
    with pm.Gravity(dim, Nm, Nf, L, wrap_gather=bool(wrap)) as g:
        g.set_state(X)
        F = g.gather(phi)
    Fo = orc.gather(dim, Nf, L, phi, X, wrap)
    assert np.array_equal(F, Fo)


# ------------------------------------------------------------------- Poisson
@pytest.mark.parametrize("dim,N", [(2, 16), (2, 64), (3, 32), (3, 64)])
def test_poisson_fp64(pm, orc, dim, N):
    L, a = 100.0, 0.37
    rho = 1.0 + 0.3 * RNG.standard_normal(N ** dim)
    say("poisson", dim, N)
    with pm.Gravity(dim, N, N, L) as g:
        phi = g.poisson(rho, a)
    phio = orc.poisson(dim, N, L, rho, a)
    assert rel(phi, phio) < 1e-12


def test_poisson_fp32(pm, orc):
    dim, N, L, a = 3, 64, 100.0, 0.5
    rho = 1.0 + 0.3 * RNG.standard_normal(N ** dim)
    say("poisson fp32")
    with pm.Gravity(dim, N, N, L, precision="fp32") as g:
        phi = g.poisson(rho, a)
    phio = orc.poisson(dim, N, L, rho, a)
    assert rel(phi, phio) < 1e-5


# ------------------------------------------------------------------- Zel'dovich + IC
@pytest.mark.parametrize("dim,mbits", [(2, 6), (3, 4), (3, 5)])
def test_ic_and_zeldovich(pm, orc, dim, mbits):
    L, ns = 100.0, -2.0
    N = 1 << mbits
    noise = orc.white_noise(7, N ** dim)
    say("ic", dim, mbits)
    dens, phi = pm.ic_scale_free(dim, mbits, L, ns, noise)
    dref = orc.ic_scale_free(dim, mbits, L, ns, noise)
    pref = orc.compute_potential(dim, mbits, L, dref)
    assert rel(dens, dref) < 1e-12
    assert rel(phi, pref) < 1e-12
    # Zel'dovich start from the oracle's potential: bit-exact
    a0, da = 0.02, 0.0098
    with pm.Gravity(dim, N, N, L) as g:
        g.init(pref, a0, a0 + da / 2)
        X, P = g.get_state()
    Xo, Po = orc.zeldovich(dim, N, L, pref, a0, a0 + da / 2)
    assert np.array_equal(X, Xo)
    assert np.array_equal(P, Po)


def test_ic_displacement(pm, orc):
    dim, mbits, L = 3, 4, 100.0
    N = 1 << mbits
    phi = RNG.standard_normal(N ** dim)
    say("ic displacement")
    psi = pm.ic_displacement(dim, mbits, L, phi)
    ref = orc.compute_displacement(dim, mbits, L, phi)
    assert rel(psi, ref) < 1e-12


# ------------------------------------------------------------------- leapfrog
@pytest.mark.parametrize("cosmo", [(68.5, 1.0, 0.0), (68.5, 0.31, 0.69), (70.0, 0.3, 0.0)])
@pytest.mark.parametrize("dim,Nm,Nf,nsteps", [(2, 32, 32, 10), (3, 16, 16, 10), (3, 16, 32, 5)])
def test_leapfrog_fp64_vs_oracle(pm, orc, dim, Nm, Nf, nsteps, cosmo):
    """Einstein-de Sitter, LambdaCDM (--concordance) and an open universe
    (Omega_k = 0.7): the expansion enters every factor (src/run/cosmology.hh:
    16-22)."""
    L, a0, da = 100.0, 0.02, 0.0098
    H0, Om, OL = cosmo
    mbits = int(np.log2(Nm))
    noise = orc.white_noise(3, Nm ** dim)
    dens = orc.ic_scale_free(dim, mbits, L, -2.0, noise)
    phi = orc.compute_potential(dim, mbits, L, dens)
    say("leapfrog", dim, Nm, Nf, nsteps, cosmo)
    sim = orc.Sim(dim, Nm, Nf, L, H0=H0, Om=Om, OL=OL)
    sim.init(phi, a0, da)
    sim.steps(nsteps)
    Xo, Po, t = sim.state()
    with pm.Gravity(dim, Nm, Nf, L, H0=H0, Omega_m=Om, Omega_L=OL) as g:
        g.init(phi, a0, a0 + da / 2)
        a = g.steps(a0, da, nsteps)
        X, P = g.get_state()
        apos, amom = g.times()
    assert a == t[2]
    assert (apos, amom) == (t[0], t[1])
    # the potential passes through the FFT each step: rounding-level drift
    assert np.max(np.abs(X - Xo)) < 1e-9 * L
    assert rel(P, Po) < 1e-9


@pytest.mark.parametrize("dim,Nm,Nf,nsteps", [(2, 32, 32, 10), (3, 16, 32, 5)])
def test_leapfrog_midpoint_fp64_vs_oracle(pm, orc, dim, Nm, Nf, nsteps):
    """PM_FLAG_MIDPOINT: the same steps with the factors at the middle of each
    interval; against the oracle run the same way."""
    L, a0, da = 100.0, 0.02, 0.0098
    mbits = int(np.log2(Nm))
    noise = orc.white_noise(3, Nm ** dim)
    phi = orc.compute_potential(dim, mbits, L, orc.ic_scale_free(dim, mbits, L, -2.0, noise))
    say("leapfrog midpoint", dim, Nm, Nf, nsteps)
    sim = orc.Sim(dim, Nm, Nf, L, midpoint=True)
    sim.init(phi, a0, da)
    sim.steps(nsteps)
    Xo, Po, t = sim.state()
    out = {}
    for mid in (True, False):
        with pm.Gravity(dim, Nm, Nf, L, midpoint=mid) as g:
            g.init(phi, a0, a0 + da / 2)
            g.steps(a0, da, nsteps)
            out[mid] = g.get_state()
    X, P = out[True]
    assert np.max(np.abs(X - Xo)) < 1e-9 * L
    assert rel(P, Po) < 1e-9
    assert np.max(np.abs(X - out[False][0])) > 1e-6 * L  # the flag changes the run


def _gpu_plane_wave_amplitude(pm, da, midpoint, prec="fp64"):
    dim, N, L = 2, 32, 100.0
    q = np.stack(np.meshgrid(*([np.arange(N) * (L / N)] * dim), indexing="ij"), -1).reshape(-1, dim)[:, ::-1].copy()
    k = 2 * np.pi / L
    phi = 1e-4 * np.cos(k * q[:, 0])
    a0 = 0.02
    n = int(round((0.5 - a0) / da))
    with pm.Gravity(dim, N, N, L, precision=prec, midpoint=midpoint) as g:
        g.init(phi, a0, a0)  # synchronised start: P at a0
        g.kick(a0, da / 2)   # to the staggered start
        g.steps(a0, da, n)
        X, _ = g.get_state()
    sn = np.sin(k * q[:, 0])
    dx = (X[:, 0] - q[:, 0] + L / 2) % L - L / 2
    return float(np.sum(dx * sn) / np.sum(sn * sn))


@pytest.mark.parametrize("midpoint,order", [(False, 1), (True, 2)])
def test_gpu_leapfrog_convergence_order(pm, midpoint, order):
    """The Richardson ratio of three step sizes on a linear plane wave: first
    order with the reference's factors, second order with PM_FLAG_MIDPOINT."""
    say("convergence order, midpoint =", midpoint)
    s = [_gpu_plane_wave_amplitude(pm, da, midpoint) for da in (0.004, 0.002, 0.001)]
    p = np.log2(abs(s[0] - s[1]) / abs(s[1] - s[2]))
    assert abs(p - order) < 0.25


@pytest.mark.parametrize("dim,Nm,Nf,wrap", [(3, 32, 64, False), (3, 32, 64, True), (2, 64, 64, False),
                                            (3, 16, 8, False)])
def test_kick_is_the_gather_of_its_potential_fp64(pm, orc, dim, Nm, Nf, wrap):
    """pm_kick takes its forces by bricks with the potential in LDS and the
    particles in the deposit's cell order; they must equal the direct gather
    of the same potential (pm_gather_mesh) bit for bit.  Positions include
    negative and out-of-box coordinates (a gather origin outside the brick)
    and a dense blob (long row parts)."""
    L, a, da = 100.0, 0.5, 0.02
    X = _positions(dim, Nm, L, 3.0 * L / Nm)
    X[1::3] = 37.0 + 0.05 * RNG.standard_normal(X[1::3].shape)
    P = RNG.standard_normal(X.shape)
    say("kick vs gather of its potential", dim, Nm, Nf, wrap)
    with pm.Gravity(dim, Nm, Nf, L, wrap_gather=wrap) as g:
        g.set_state(X, P)
        phi = g.potential(a)
        F = g.gather(phi)
        g.kick(a, da)
        X2, P2 = g.get_state()
    ck = da / orc.adot(68.5, 1.0, 0.0, a)
    assert np.array_equal(X2, X)
    assert np.array_equal(P2, P - F * ck)


@pytest.mark.parametrize("dim,Nm,Nf,wrap,corner", [(3, 32, 64, False, False), (3, 32, 64, True, False),
                                                   (2, 64, 64, False, False), (3, 16, 8, False, False),
                                                   (3, 64, 256, False, False), (2, 128, 512, False, False),
                                                   (3, 32, 64, False, True), (2, 64, 128, False, True)])
def test_kick_is_the_gather_of_its_potential_fp32(pm, orc, dim, Nm, Nf, wrap, corner):
    """The FP32 kick reads the potential from an LDS tile per brick; a
    particle with a negative unwrapped coordinate on a brick's upper face
    (its truncated origin is the next cell up, outside the tile) reads the
    mesh instead.  Either way its kick must be the direct gather of the same
    potential (pm_gather_mesh) rounded to the momentum grid, bit for bit.  A
    third of the particles sit one box below [0, L), so every brick holds
    such particles.  corner: every particle just below the origin of the box
    on every axis, all in the one corner cell of one brick, which overflows
    the workgroup's queue of them."""
    L, a, da = 100.0, 0.5, 0.02
    X = _positions(dim, Nm, L, 3.0 * L / Nm, special=False)
    X[::3] -= L
    X[1::7] = 37.0 + 0.05 * RNG.standard_normal(X[1::7].shape)
    if corner:
        X = -(L / Nf) * RNG.uniform(0.05, 0.95, X.shape)
    P = RNG.standard_normal(X.shape)
    say("fp32 kick vs gather of its potential", dim, Nm, Nf, wrap, corner)
    with pm.Gravity(dim, Nm, Nf, L, precision="fp32", wrap_gather=wrap) as g:
        g.set_state(X, P)
        X0, P0 = g.get_state()
        pq = g.momentum_quantum()
        phi = g.potential(a)
        F = g.gather(phi)
        g.kick(a, da)
        X2, P2 = g.get_state()
    ck = da / orc.adot(68.5, 1.0, 0.0, a)
    assert (X0 < 0).any() and np.array_equal(X0, np.round(X0 / (L / 2 ** 30)) * (L / 2 ** 30))
    assert np.array_equal(X2, X0)
    assert np.array_equal(P2, P0 - np.rint(F * (ck * (1.0 / pq))) * pq)


@pytest.mark.parametrize("dim,Nm,Nf", [(2, 32, 64), (3, 16, 32), (3, 32, 64), (3, 16, 8)])
def test_fp32_gather_truncates_like_the_reference(pm, orc, dim, Nm, Nf):
    """The FP32 gather of a given potential at unwrapped positions in
    [-2L, 2L) against the oracle's FP64 gather (src/run/numeric.hh:17-44,
    which truncates the unwrapped coordinate towards zero and extrapolates
    for a negative one), on the same fixed-point positions and the same
    FP32-rounded potential: FP32 rounding only.  The folded gather differs
    from it by O(1) on these positions, so a gather that ignored the sign
    would fail."""
    L = 100.0
    X = _positions(dim, Nm, L, 3.0 * L / Nm, special=False)
    X[::2] -= L
    X[1::10] += 0.5 * L
    X = np.clip(X, -1.9 * L, 1.9 * L)  # all within [-2L, 2L)
    phi = RNG.standard_normal(Nf ** dim).astype(np.float32).astype(np.float64)
    say("fp32 gather, truncation", dim, Nm, Nf)
    with pm.Gravity(dim, Nm, Nf, L, precision="fp32") as g:
        g.set_state(X)
        Xq, _ = g.get_state()
        F32 = g.gather(phi)
    assert np.max(np.abs(Xq - X)) < 1e-6 and (Xq < 0).mean() > 0.3
    F64 = orc.gather(dim, Nf, L, phi, Xq, 0)
    Ffold = orc.gather(dim, Nf, L, phi, Xq, 1)
    say(f"  rel {rel(F32, F64):.3g}, folded vs truncated {rel(Ffold, F64):.3g}")
    assert rel(F32, F64) < 1e-5
    assert rel(Ffold, F64) > 1e-1


def test_fp32_set_positions_keeps_momenta(pm, orc):
    """pm_set_state(X, NULL) on an FP32 handle that has stepped (particles
    stored in brick order) keeps every particle's momentum."""
    dim, Nm, Nf, L, a0, da = 3, 32, 64, 100.0, 0.02, 0.0098
    noise = orc.white_noise(12, Nm ** dim)
    phi = orc.compute_potential(dim, 5, L, orc.ic_scale_free(dim, 5, L, -2.0, noise))
    say("fp32 set_state(X) after steps")
    with pm.Gravity(dim, Nm, Nf, L, precision="fp32") as g:
        g.init(phi, a0, a0 + da / 2)
        g.steps(a0, da, 5)
        X, P = g.get_state()
        g.set_state(X)  # the same positions, uploaded in mass-grid order; momenta kept
        X2, P2 = g.get_state()
        g.steps(a0 + 5 * da, da, 2)  # and the run goes on
    assert np.array_equal(X2, X) and np.array_equal(P2, P)


def test_steps_fused_equals_unfused(pm, orc):
    dim, Nm, Nf, L, a0, da = 3, 16, 16, 100.0, 0.02, 0.0098
    noise = orc.white_noise(5, Nm ** dim)
    phi = orc.compute_potential(dim, 4, L, orc.ic_scale_free(dim, 4, L, -2.0, noise))
    say("fused vs unfused")
    with pm.Gravity(dim, Nm, Nf, L) as g1, pm.Gravity(dim, Nm, Nf, L) as g2:
        g1.init(phi, a0, a0 + da / 2)
        g2.init(phi, a0, a0 + da / 2)
        g1.steps(a0, da, 4)
        a = a0
        for _ in range(4):
            g2.drift(a, da)
            g2.kick(a + da / 2, da)
            a += da
        X1, P1 = g1.get_state()
        X2, P2 = g2.get_state()
    assert np.array_equal(X1, X2) and np.array_equal(P1, P2)


@pytest.mark.parametrize("wrap", [False, True])
def test_fp32_form_tracks_fp64(pm, orc, wrap):
    dim, Nm, Nf, L, a0, da, n = 3, 32, 32, 100.0, 0.02, 0.0098, 20
    noise = orc.white_noise(11, Nm ** dim)
    phi = orc.compute_potential(dim, 5, L, orc.ic_scale_free(dim, 5, L, -2.0, noise))
    say("fp32 vs fp64, wrap", wrap)
    out = {}
    for prec in ("fp64", "fp32"):
        with pm.Gravity(dim, Nm, Nf, L, precision=prec, wrap_gather=wrap) as g:
            g.init(phi, a0, a0 + da / 2)
            g.steps(a0, da, n)
            X, P = g.get_state()
            out[prec] = (np.mod(X, L), P, g.deposit())
    X64, P64, r64 = out["fp64"]
    X32, P32, r32 = out["fp32"]
    dx = np.abs(X32 - X64)
    dx = np.minimum(dx, L - dx)
    assert np.sqrt(np.mean(dx ** 2)) < 1e-4 * L / Nm
    assert rel(P32, P64) < 1e-3
    assert np.sqrt(np.mean((r32 - r64) ** 2)) < 1e-3


@pytest.mark.parametrize("wrap", [False, True])
@pytest.mark.parametrize("dim,Nm,Nf", [(3, 128, 256), (3, 64, 256), (3, 128, 512)])
def test_fp32_kick_matches_fp64_on_large_meshes(pm, orc, dim, Nm, Nf, wrap):
    """On meshes with enough bricks, a gather workgroup walks several bricks
    along the last axis and keeps their potential planes in an LDS ring; the
    FP32 kicks must then still match the FP64 form's (same gather mode) to FP32
    accuracy: a misplaced plane would show as an O(1) error."""
    L, a0, da, n = 100.0, 0.02, 0.0098, 3
    mbits = int(np.log2(Nm))
    noise = orc.white_noise(41, Nm ** dim)
    phi = orc.compute_potential(dim, mbits, L, orc.ic_scale_free(dim, mbits, L, -2.0, noise))
    say("fp32 vs fp64 kicks", dim, Nm, Nf)
    out = {}
    for prec in ("fp64", "fp32"):
        with pm.Gravity(dim, Nm, Nf, L, precision=prec, wrap_gather=wrap) as g:
            g.init(phi, a0, a0 + da / 2)
            _, P0 = g.get_state()
            g.steps(a0, da, n)
            _, P1 = g.get_state()
            out[prec] = P1 - P0
    assert rel(out["fp32"], out["fp64"]) < 1e-4


def test_cic_tool(pm, orc):
    dim, Nm, Nf, L = 3, 16, 32, 100.0
    X = _positions(dim, Nm, L, 1.0, special=False)
    say("cic tool")
    r1 = pm.cic(dim, Nm, Nf, L, X, 1)
    assert np.array_equal(r1, orc.cic_tool(dim, Nm, Nf, L, 1, X))
    r2 = pm.cic(dim, Nm, Nf, L, X, 2)
    ref2 = orc.cic_tool(dim, Nm, Nf, L, 2, X)
    assert rel(r2, ref2) < 1e-12


def test_power_spectrum(pm, orc):
    dim, Nm, Nf, L = 3, 32, 32, 100.0
    X = _positions(dim, Nm, L, 2.0, special=False)
    edges = orc.log_edges(Nf, L, 16)
    say("power spectrum")
    with pm.Gravity(dim, Nm, Nf, L) as g:
        g.set_state(X)
        rho = g.deposit()
        k, p, c = g.power_spectrum(16, edges)
    ko, po, co = orc.power_spectrum(dim, Nf, L, rho - 1.0, edges)
    assert np.array_equal(c.astype(np.int64), co)
    assert rel(k, ko) < 1e-12
    assert rel(p, po) < 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_runs_are_bit_reproducible(pm, orc, prec):
    """The deposit has no order-dependent floating-point sums: two runs from
    the same state agree bit for bit (the reference's omp critical deposit
    does not have this property), and so do the FP64 diagnostics sums (in
    the FX form they run in particle order, not in the storage order the
    re-sort's atomics leave)."""
    dim, Nm, Nf, L, a0, da = 3, 32, 64, 100.0, 0.02, 0.0098
    noise = orc.white_noise(21, Nm ** dim)
    phi = orc.compute_potential(dim, 5, L, orc.ic_scale_free(dim, 5, L, -2.0, noise))
    say("reproducible", prec)
    outs = []
    for _ in range(2):
        with pm.Gravity(dim, Nm, Nf, L, precision=prec) as g:
            g.init(phi, a0, a0 + da / 2)
            g.steps(a0, da, 8)
            X, P = g.get_state()
            rho = g.deposit()
            d = g.diagnostics()
            outs.append((X, P, rho, np.array(d["momentum"] + [d["p2"], d["mass"], d["potential"]])))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("dim,Nm,Nf", [(3, 32, 64), (2, 128, 256), (3, 64, 64)])
def test_incremental_resort_equals_full_sort(pm, orc, dim, Nm, Nf, precision):
    """Both forms keep their particles in brick order and re-sort
    incrementally (only particles that changed brick or face class are placed
    anew).  A run that forces a full radix sort before every kick (a deposit
    between steps breaks the sorted storage order) ends in the same state bit
    for bit: the deposit sums are exact integers, so no result depends on the
    order within a brick.  64^3 particles on a 64^3 mesh put 4096 particles in
    each FP32 brick, whose movers reserve their slots by new group."""
    L, a0, da, n = 100.0, 0.02, 0.0196, 12
    mbits = int(np.log2(Nm))
    noise = orc.white_noise(31, Nm ** dim)
    phi = orc.compute_potential(dim, mbits, L, orc.ic_scale_free(dim, mbits, L, -2.0, noise))
    say("resort", dim, Nm, Nf, precision)
    with pm.Gravity(dim, Nm, Nf, L, precision=precision) as g:
        g.init(phi, a0, a0 + da / 2)
        g.steps(a0, da, n)
        Xa, Pa = g.get_state()
        ra = g.deposit()
    with pm.Gravity(dim, Nm, Nf, L, precision=precision) as g:
        g.init(phi, a0, a0 + da / 2)
        a = a0
        for i in range(n):
            g.drift(a, da)
            g.deposit()
            g.kick(a + da / 2, da)
            a += da
        Xb, Pb = g.get_state()
        rb = g.deposit()
    assert np.array_equal(Xa, Xb) and np.array_equal(Pa, Pb) and np.array_equal(ra, rb)


def test_full_size_run(pm, orc):
    """The 3-D mixed-precision configuration at its real size: 256^3 particles
    on a 512^3 mesh, L = 100, P(k) ~ k^-1.5 from seeded white noise, a = 0.02
    to 1 in 200 steps.  The oracle cannot run it in test time; parity is shown
    through what does not depend on the size: the FP64 form's cell keys of its
    final, clustered state are bit-exact against the oracle's, and so is its
    deposit with PM_FLAG_SERIAL_DEPOSIT (the fixed-point one within 1e-13);
    the FP32 form, against the FP64 form with the reference's own gather (the
    unwrapped coordinate truncated, src/run/numeric.hh:19,26-28): P(k) within
    1e-3 per bin over 64 log bins, displacement rms below 1e-3 mesh cell,
    density rms below 2e-3 (mean 1), force rms below 5e-3 of the force rms
    (measured 1.8e-3: the trajectories of the clustered end state diverge at
    the rounding level and Q1 makes the force jump at negative cell
    boundaries), and, at the same positions (the FP32 operator alone), below
    1e-4; mass is conserved; a second FP32 run ends bit-identical."""
    dim, Nm, Nf, L, ns, a0, n = 3, 256, 512, 100.0, -1.5, 0.02, 200
    da = (1.0 - a0) / n
    say("full size: IC", Nm, Nf)
    noise = np.random.default_rng(1).standard_normal(Nm ** dim)
    _, phi = pm.ic_scale_free(dim, 8, L, ns, noise)
    del noise
    runs = {}
    for name, prec in (("fp32", "fp32"), ("fp32b", "fp32"), ("fp64", "fp64")):
        say("full size:", name, n, "steps")
        with pm.Gravity(dim, Nm, Nf, L, precision=prec) as g:
            g.init(phi, a0, a0 + da / 2)
            a = g.steps(a0, da, n)
            assert abs(a - 1.0) < 1e-12
            X, P = g.get_state()
            if name == "fp32b":
                runs[name] = (X, P)
                continue
            pk = g.power_spectrum(64)
            F = g.forces(a)
            if prec == "fp64":
                rho, keys = g.deposit(keys=True)
                runs[name] = (X, P, pk, rho, keys, F)
            else:
                runs[name] = (X, P, pk, g.deposit(), None, F)
    X32, P32, (k32, p32, c32), r32, _, F32 = runs["fp32"]
    X32b, P32b = runs.pop("fp32b")
    assert np.array_equal(X32, X32b) and np.array_equal(P32, P32b)
    del X32b, P32b
    X64, P64, (k64, p64, c64), r64, keys, F64 = runs["fp64"]

    say("full size: oracle keys and deposit")
    assert np.array_equal(keys, orc.cic_keys(dim, Nf, L, X64))
    ro = orc.md_cic(dim, Nf, L, X64, (Nf / Nm) ** dim)
    _check_fp64_deposit(r64, ro, False)
    with pm.Gravity(dim, Nm, Nf, L, serial_deposit=True) as g:
        g.set_state(X64)
        assert np.array_equal(g.deposit(), ro)
    del ro
    assert abs(r64.sum() / Nf ** dim - 1) < 1e-12
    assert abs(r32.sum() / Nf ** dim - 1) < 1e-6
    assert r64.max() > 100  # the end state is clustered
    assert (X64 < 0).any()  # some coordinates are negative: the gather's truncation matters

    m = (c32 > 0) & (p64 > 0)
    assert m.sum() > 50 and np.array_equal(c32, c64) and np.array_equal(k32, k64)
    pk_rel = np.max(np.abs(p32[m] / p64[m] - 1))
    dx = np.abs(X32 - X64)  # both unwrapped
    dx = np.minimum(dx, np.abs(dx - L))
    dx_rms = np.sqrt(np.mean(dx ** 2)) / (L / Nf)
    rho_rms = np.sqrt(np.mean((r32 - r64) ** 2))
    f_scale = np.sqrt(np.mean(F64 ** 2))
    f_rms = np.sqrt(np.mean((F32 - F64) ** 2)) / f_scale
    p_rel = rel(P32, P64)
    del X32, P32, F32
    # the FP32 operator alone: its force at the FP64 run's final positions
    with pm.Gravity(dim, Nm, Nf, L, precision="fp32") as g:
        g.set_state(X64)
        f_op = np.sqrt(np.mean((g.forces(1.0) - F64) ** 2)) / f_scale
    say(f"full size: P(k) max rel {pk_rel:.3g}, displacement rms {dx_rms:.3g} mesh cells, "
        f"density rms {rho_rms:.3g}, force rms rel {f_rms:.3g} (same positions: {f_op:.3g}), "
        f"momentum max rel {p_rel:.3g}")
    assert pk_rel < 1e-3
    assert dx_rms < 1e-3
    assert rho_rms < 2e-3
    assert f_rms < 5e-3
    assert f_op < 1e-4
