// pm_fft.hpp -- Stockham mixed-radix FFT building blocks for gfx950.
//
// The transforms replace the reference's FFTW c2c plans (src/base/fft.cc:15-38)
// with real-to-complex / complex-to-real transforms on LDS-staged pencils.
// Conventions are FFTW's: unnormalised, forward sign -1, backward sign +1.
//
// Stockham autosort, radix r, sub-transform length Ns (Ns = 1, r, r*r', ...):
// butterfly j in [0, N/r) reads x[j + t*N/r] (t < r), multiplies input t by
// W^(DIR * t * (j mod Ns) / (Ns*r)), does an r-point DFT and writes
// y[(j - j mod Ns)*r + (j mod Ns) + t*Ns].  Radices are 8 where possible,
// with one leading radix-2 or radix-4 stage when log2 N is not a multiple of 3.
#pragma once
#include "pm_common.hpp"

namespace cpm {

// Sign of the exponent: -1 forward, +1 backward.
template <typename T, int DIR>
__host__ __device__ __forceinline__ cx<T> rot_quarter(cx<T> a) {
    // multiply by -i (DIR < 0) or +i (DIR > 0)
    return DIR < 0 ? mk(a.im, -a.re) : mk(-a.im, a.re);
}

template <typename T, int DIR>
__host__ __device__ __forceinline__ void dft2(cx<T>* v) {
    cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <typename T, int DIR>
__host__ __device__ __forceinline__ void dft4(cx<T>* v) {
    cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
    cx<T> b0 = v[1] + v[3], b1 = rot_quarter<T, DIR>(v[1] - v[3]);
    v[0] = a0 + b0;
    v[2] = a0 - b0;
    v[1] = a1 + b1;
    v[3] = a1 - b1;
}

template <typename T, int DIR>
__host__ __device__ __forceinline__ void dft8(cx<T>* v) {
    const T h = (T)0.70710678118654752440;
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft4<T, DIR>(e);
    dft4<T, DIR>(o);
    // o[k] *= W8^(DIR k): W8^1 = (1 - i)/sqrt2 forward, (1 + i)/sqrt2 backward
    cx<T> o1, o3;
    if (DIR < 0) {
        o1 = mk((o[1].re + o[1].im) * h, (o[1].im - o[1].re) * h);
        o3 = mk((o[3].im - o[3].re) * h, -(o[3].re + o[3].im) * h);
    } else {
        o1 = mk((o[1].re - o[1].im) * h, (o[1].re + o[1].im) * h);
        o3 = mk(-(o[3].re + o[3].im) * h, (o[3].re - o[3].im) * h);
    }
    cx<T> o2 = rot_quarter<T, DIR>(o[2]);
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + o1;
    v[5] = e[1] - o1;
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = e[3] + o3;
    v[7] = e[3] - o3;
}

template <typename T, int R, int DIR>
__host__ __device__ __forceinline__ void dft(cx<T>* v) {
    if constexpr (R == 2) dft2<T, DIR>(v);
    else if constexpr (R == 4) dft4<T, DIR>(v);
    else if constexpr (R == 8) dft8<T, DIR>(v);
}

// Twiddle + r-point DFT for butterfly j of a length-N transform at
// sub-length Ns.  tw[m * tws] = exp(-2 pi i m / N) (conjugated for DIR > 0).
template <typename T, int N, int R, int DIR>
__host__ __device__ __forceinline__ void butterfly(cx<T>* v, int j, int Ns, const cx<T>* tw,
                                                   int tws) {
    // Every lane multiplies, also at k = 0 (twiddles exactly 1): a branch on
    // k would diverge within the wave and cost register copies at the join.
    // The first stage (Ns = 1, all k = 0) skips the multiplications.
    int k = j & (Ns - 1);
    if (Ns > 1) {
        const int mstep = k * (N / (Ns * R));
#pragma unroll
        for (int t = 1; t < R; ++t) {
            cx<T> w = tw[mstep * t * tws];
            if (DIR > 0) w.im = -w.im;
            v[t] = v[t] * w;
        }
    }
    dft<T, R, DIR>(v);
}

template <int N, int R>
__host__ __device__ __forceinline__ int stock_out(int j, int Ns, int t) {
    int k = j & (Ns - 1);
    return (j - k) * R + k + t * Ns;
}

// Radix schedule for N = 2^L: one leading stage of radix 2^(L mod 3) when
// L mod 3 != 0, then radix-8 stages.
template <int N>
struct Sched {
    static constexpr int L = ilog2c(N);
    static constexpr int lead = (L % 3 == 0) ? 0 : (1 << (L % 3));
    static constexpr int n8 = L / 3;
};

}  // namespace cpm
