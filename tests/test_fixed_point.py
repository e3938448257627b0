"""Host checks of the FX deposit's fixed-point rounding (pm_particles.hip,
fx_quantum): round(w 2^40), ties to even, computed with FP32 operations
(a = w 2^8, hi = floor(a), lo = rint((a - hi) 2^32), hi 2^32 + lo) equals the
FP64 rounding of the exact product w 2^40 that the accumulators were
specified with.  numpy's float32 floor / rint / multiply are the IEEE
operations the kernel compiles to (round-to-nearest-even products, exact
here)."""
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
