"""Host checks of the FX deposit's fixed-point rounding (pm_particles.hip,
fx_quantum): round(w 2^40), ties to even, computed with FP32 operations
(a = w 2^8, hi = floor(a), lo = rint((a - hi) 2^32), hi 2^32 + lo) equals the
FP64 rounding of the exact product w 2^40 that the accumulators were
specified with.  numpy's float32 floor / rint / multiply are the IEEE
operations the kernel compiles to (round-to-nearest-even products, exact
here).  Also the FP64 force's division by 3 (div3, below)."""
import numpy as np


def fx_quantum_f32(w):
    w = np.asarray(w, dtype=np.float32)
    a = w * np.float32(256.0)
    hi = np.floor(a)
    lo = np.rint((a - hi) * np.float32(4294967296.0))
    return (hi.astype(np.uint64) << np.uint64(32)) + lo.astype(np.uint64)


def fx_quantum_exact(w):
    # w * 2^40 is exact in float64 (a 24-bit significand scaled by a power of
    # two); np.rint rounds half to even
    return np.rint(np.asarray(w, dtype=np.float32).astype(np.float64) * 2.0 ** 40).astype(np.uint64)


def test_fp32_rounding_equals_fp64_rounding():
    r = np.random.default_rng(7)
    # log-uniform over [2^-60, 2^23), the weights' whole range, and uniform
    # weights of the usual size (mass 8 times products of fractions)
    w = np.concatenate([2.0 ** r.uniform(-60, 23, 400000), r.uniform(0, 8, 400000)]).astype(np.float32)
    w = w[w < 2.0 ** 23]
    assert np.array_equal(fx_quantum_f32(w), fx_quantum_exact(w))


def test_fp32_rounding_special_values():
    # zero, ties at the 2^-40 quantum (w 2^40 = k + 1/2 for k even and odd),
    # values just below and above a tie, the smallest normal and subnormal
    # floats, and the largest weight below 2^23
    ties = np.array([(k + 0.5) * 2.0 ** -40 for k in range(0, 64)], dtype=np.float64)
    w = np.concatenate([
        [0.0, 2.0 ** -149, 2.0 ** -126, 2.0 ** -41, 2.0 ** -40, 3 * 2.0 ** -41, 1.0, 8.0,
         np.nextafter(np.float32(2.0 ** 23), np.float32(0))],
        ties, np.nextafter(ties.astype(np.float32), np.float32(1)), np.nextafter(ties.astype(np.float32), np.float32(0)),
    ]).astype(np.float32)
    got, ref = fx_quantum_f32(w), fx_quantum_exact(w)
    assert np.array_equal(got, ref)
    # ties round to even
    assert fx_quantum_f32(np.float32(0.5 * 2.0 ** -40)) == 0
    assert fx_quantum_f32(np.float32(1.5 * 2.0 ** -40)) == 2
    assert fx_quantum_f32(np.float32(2.5 * 2.0 ** -40)) == 2


# ---------------------------------------------------------------- div3
# The FP64 force (pm_particles.hip, force_f64) forms the reference's f / 12
# and 2 f / 3 (src/run/numeric.hh:98,101) from one third per node,
# t = RN(x / 3) of x = f / 4, with 2 f / 3 rounded = 8 t; RN(x / 3) is
# q0 = RN(x RN(1/3)), r = fma(-q0, 3, x), RN(q0 + r RN(1/3)), with the sign
# of x.  Here fma is evaluated exactly with fractions and rounded once
# (float(Fraction) rounds to nearest even).
from fractions import Fraction

_C3 = float.fromhex("0x1.5555555555555p-2")  # RN(1/3)


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _div3(x, corrected=True):
    q0 = x * _C3
    if not corrected:
        return q0
    r = _fma(-q0, 3.0, x)
    return float(np.copysign(_fma(r, _C3, q0), x))


def _div3_cases():
    r = np.random.default_rng(11)
    xs = list(r.standard_normal(3000) * 2.0 ** r.integers(-300, 300, 3000))
    # x = (3k + d) ulp: quotients just off a representable number, both signs
    for e in (-3, 0, 7, 40):
        for k in r.integers(2 ** 50, 2 ** 51, 300):
            for d in (-1, 0, 1, 2):
                m = int(3 * k + d)
                if 2 ** 52 <= m < 2 ** 53:
                    xs.append(float(m) * 2.0 ** (e - 52))
                    xs.append(-float(m) * 2.0 ** (e - 52))
    return [x for x in xs if x != 0.0 and abs(x) >= 2.0 ** -1000] + [0.0, -0.0, 1.0, 3.0, 2.0 ** -1000]


def test_div3_is_the_rounded_quotient():
    for x in _div3_cases():
        for y, ref in ((x, x / 3.0), (x * 0.25, x / 12.0), (2.0 * x, 2.0 * x / 3.0)):
            got = _div3(y)
            assert got == ref and np.signbit(got) == np.signbit(ref), (x, y, got, ref)


def test_div3_without_its_correction_fails():
    # planted mistake: the bare product x RN(1/3) is off by an ulp for some x
    assert any(_div3(x, corrected=False) != x / 3.0 for x in _div3_cases())


def test_one_third_per_node_gives_both_terms():
    # RN(f / 12) = div3(f / 4), and RN(2 f / 3) = 8 of it exactly
    for f in _div3_cases():
        t = _div3(f * 0.25)
        for got, ref in ((t, f / 12.0), (8.0 * t, 2.0 * f / 3.0)):
            assert got == ref and np.signbit(got) == np.signbit(ref), (f, got, ref)


def _fd4_reference(fm2, fm1, fp1, fp2):
    # src/run/numeric.hh:98,101, left to right
    return fm2 / 12. - 2 * fm1 / 3. + 2 * fp1 / 3. - fp2 / 12.


def _fd4_thirds(fm2, fm1, fp1, fp2, scale=8.0):
    t = [_div3(f * 0.25) for f in (fm2, fm1, fp1, fp2)]
    return t[0] - scale * t[1] + scale * t[2] - t[3]


def test_fd4_from_shared_thirds_is_the_reference_derivative():
    r = np.random.default_rng(19)
    # potentials of one magnitude (neighbouring nodes) and of mixed ones
    for _ in range(4000):
        e = int(r.integers(-60, 60))
        f = r.standard_normal(4) * 2.0 ** e
        if r.random() < 0.3:
            f = f * 2.0 ** r.integers(-30, 30, 4)
        f = [float(v) for v in f]
        assert _fd4_thirds(*f) == _fd4_reference(*f), f
    for f in ([0.0, -0.0, 0.0, -0.0], [1.0, 1.0, 1.0, 1.0], [3.0, -3.0, 3.0, -3.0]):
        got, ref = _fd4_thirds(*f), _fd4_reference(*f)
        assert got == ref and np.signbit(got) == np.signbit(ref)


def test_fd4_from_thirds_with_a_wrong_scale_fails():
    # planted mistake: 2 f / 3 taken as 4 thirds instead of 8
    r = np.random.default_rng(23)
    cases = [[float(v) for v in r.standard_normal(4)] for _ in range(100)]
    assert any(_fd4_thirds(*f, scale=4.0) != _fd4_reference(*f) for f in cases)


# ---------------------------------------------------------------- div_fma
# v / res in the FP64 force (pm_particles.hpp, div_fma), taken when the odd
# part of res's significand is below 2^50 (divides_by_fma): q0 = RN(v
# RN(1/res)), r = fma(-q0, res, v), RN(q0 + r RN(1/res)) with v's sign.
def _odd_significand(b):
    m, e = np.frexp(b)
    m = int(float(m) * 2.0 ** 53)
    while m and m % 2 == 0:
        m //= 2
    return m


def _div_fma(v, b):
    y = 1.0 / b
    q0 = v * y
    r = _fma(-q0, b, v)
    return float(np.copysign(_fma(r, y, q0), v))


def test_div_fma_is_the_rounded_quotient():
    r = np.random.default_rng(5)
    vs = list(r.standard_normal(1500) * 2.0 ** r.integers(-40, 40, 1500)) + [0.0, -0.0, 1.0]
    for L in (100.0, 50.0, 1000.0, 250.0, 1.0, 2.5, 0.75, 3.0):
        for N in (16, 128, 512, 1024):
            b = L / N
            assert _odd_significand(b) < 2 ** 50
            for v in vs:
                # values just off the midpoints of quotients near v / b
                for w in (v, b * (float(np.nextafter(v / b, np.inf)) + 0.0)):
                    got, ref = _div_fma(w, b), w / b
                    assert got == ref and np.signbit(got) == np.signbit(ref), (L, N, w, got, ref)


def test_div_fma_is_not_taken_for_long_odd_significands():
    # a spacing like 137.3 / 512 has an odd significand far above 2^50: the
    # force divides there instead
    assert _odd_significand(137.3 / 512) >= 2 ** 50
    assert _odd_significand(100.0 / 512) == 25
