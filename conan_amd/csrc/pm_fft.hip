// pm_fft.hip -- real-to-complex / complex-to-real 2-D and 3-D FFTs on LDS pencils,
// and the fused k-space passes of the Poisson solve.
//
// Replaces the FFTW c2c plans of the reference (src/base/fft.cc:15-38, used at
// src/run/nbody.hh:105,109 and src/ic/ic.cc:66-115).  Data layout in HBM:
//   real mesh   T[N^R], axis 0 fastest (src/base/box.hh:51-53);
//   spectrum    main  cx<T>[H * N^(R-1)], H = N/2, modes kx in [0, H), kx fastest;
//               nyq   cx<T>[N^(R-1)], the kx = N/2 plane, ky fastest.
// Keeping the Nyquist plane apart leaves every pencil of the main block a
// power of two long and aligned.
//
// Passes of one Poisson solve (2R-1 read+write sweeps of the half spectrum):
//   rows r2c (fused: delta = rho - mean) -> [3-D: y forward] ->
//   last axis forward * Green's function * last axis backward (one pass) ->
//   [3-D: y backward] -> rows c2r (fused: phi = x / s).
#include "pm_fft.hpp"
#include "pm_fft_api.hpp"

#include <cmath>
#include <vector>

namespace cpm {

// Workgroup barrier that orders LDS accesses only.  Every barrier of these
// kernels separates LDS writes from LDS reads of other threads; global memory
// is never exchanged between the threads of a workgroup (a tile's HBM loads
// are consumed in registers before its first barrier, and its stores go to
// addresses no thread reads again), so the barrier need not wait for
// outstanding global stores or prefetch loads, as __syncthreads() does.
__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS slot of element n of pencil p: pencils are N apart, and the low bits
// of n are XOR-swizzled with the pencil number so that both the column-wise
// global<->LDS copies and the in-pencil Stockham accesses spread over banks.
template <int N>
__device__ __forceinline__ int lidx(int p, int n) {
    constexpr int SW = (N < 32 ? N : 32) - 1;
    return p * N + (n ^ (p & SW));
}

template <typename T, int N, int E, int NT, int DIR, int R>
__device__ __forceinline__ void lds_stage(cx<T>* sh, int Ns, const cx<T>* __restrict__ tw, int tws,
                                          int tid) {
    constexpr int NB = E / R;
    constexpr int BPP = N / R;
    cx<T> v[E];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int b = tid + i * NT;
        const int p = b / BPP, j = b % BPP;
#pragma unroll
        for (int t = 0; t < R; ++t) v[i * R + t] = sh[lidx<N>(p, j + t * BPP)];
    }
    lds_sync();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int b = tid + i * NT;
        const int p = b / BPP, j = b % BPP;
        butterfly<T, N, R, DIR>(&v[i * R], j, Ns, tw, tws);
#pragma unroll
        for (int t = 0; t < R; ++t) sh[lidx<N>(p, stock_out<N, R>(j, Ns, t))] = v[i * R + t];
    }
    lds_sync();
}

// In-LDS transform of all pencils of the tile (E*NT elements, pencils of N).
// The caller has synchronised after filling LDS; on return LDS holds the
// result and all threads have synchronised.
template <typename T, int N, int E, int NT, int DIR>
__device__ __forceinline__ void lds_fft(cx<T>* sh, const cx<T>* __restrict__ tw, int tws, int tid) {
    using S = Sched<N>;
    int Ns = 1;
    if constexpr (S::lead != 0) {
        lds_stage<T, N, E, NT, DIR, S::lead>(sh, Ns, tw, tws, tid);
        Ns *= S::lead;
    }
    if constexpr (S::n8 > 0) {
#pragma unroll
        for (int s = 0; s < S::n8; ++s) {
            lds_stage<T, N, E, NT, DIR, 8>(sh, Ns, tw, tws, tid);
            Ns *= 8;
        }
    }
}

// ------------------------------------------------------------------ rows
template <typename T, int N>
struct RowCfg {
    static constexpr int H = N / 2;
    static constexpr int E = H >= 8 ? 8 : H;
    static constexpr int P = (2048 / H) > 0 ? 2048 / H : 1;
    static constexpr int NT = P * H / E;
};

// rows of N reals (minus the shift) -> H complex modes + the Nyquist mode.
template <typename T, int N>
__global__ __launch_bounds__((RowCfg<T, N>::NT)) void k_rows_r2c(
    const T* __restrict__ in, cx<T>* __restrict__ spec, cx<T>* __restrict__ nyq, long nrows,
    T shift, const double* __restrict__ mean_ptr, const cx<T>* __restrict__ tw) {
    using C = RowCfg<T, N>;
    constexpr int H = C::H, P = C::P, NT = C::NT;
    __shared__ cx<T> sh[P * H];
    __shared__ cx<T> stw[N];
    const int tid = threadIdx.x;
    const long row0 = (long)blockIdx.x * P;
    const T sft = mean_ptr ? (T)(*mean_ptr) : shift;
    for (int m = tid; m < N; m += NT) stw[m] = tw[m];
    {
        // all loads of the thread first, then the LDS stores
        cx<T> z[C::E];
#pragma unroll
        for (int i = 0; i < C::E; ++i) {
            const int e = tid + i * NT;
            const long row = row0 + e / H;
            z[i] = row < nrows ? *reinterpret_cast<const cx<T>*>(in + row * N + 2 * (e % H)) : mk<T>(sft, sft);
        }
#pragma unroll
        for (int i = 0; i < C::E; ++i) {
            const int e = tid + i * NT;
            sh[lidx<H>(e / H, e % H)] = mk<T>(z[i].re - sft, z[i].im - sft);
        }
    }
    lds_sync();
    lds_fft<T, H, C::E, NT, -1>(sh, stw, 2, tid);
    for (int e = tid; e < P * H; e += NT) {
        const int p = e / H, k = e % H;
        const long row = row0 + p;
        if (row >= nrows) continue;
        const cx<T> Zk = sh[lidx<H>(p, k)];
        const cx<T> Zm = conj(sh[lidx<H>(p, (H - k) & (H - 1))]);
        const cx<T> A = scale(Zk + Zm, (T)0.5);
        const cx<T> WB = stw[k] * (Zk - Zm);
        spec[row * H + k] = mk<T>(A.re + (T)0.5 * WB.im, A.im - (T)0.5 * WB.re);
        if (k == 0) nyq[row] = mk<T>(Zk.re - Zk.im, (T)0);
    }
}

// H complex modes + Nyquist mode per row -> N reals, divided by s.
template <typename T, int N>
__global__ __launch_bounds__((RowCfg<T, N>::NT)) void k_rows_c2r(
    const cx<T>* __restrict__ spec, const cx<T>* __restrict__ nyq, T* __restrict__ out, long nrows,
    double s, const cx<T>* __restrict__ tw) {
    using C = RowCfg<T, N>;
    constexpr int H = C::H, P = C::P, NT = C::NT, E = C::E;
    __shared__ cx<T> sh[P * H];
    __shared__ cx<T> stw[N];
    const int tid = threadIdx.x;
    const long row0 = (long)blockIdx.x * P;
    using S = Sched<H>;
    constexpr int R0 = S::lead ? S::lead : 8;                // radix of the first stage
    constexpr int NMID = S::n8 - (S::lead ? 1 : 2);          // radix-8 stages between first and last
    // merge and first stage fused, last stage stored from registers: FP32
    // only (FP64 at 512: 0.414 ms fused against 0.397 through LDS; FP32
    // 0.205 against 0.212)
    constexpr bool fused = sizeof(T) == 4 && NMID >= 0 && E % R0 == 0 && E % 8 == 0;
    for (int m = tid; m < N; m += NT) stw[m] = tw[m];
    // The Nyquist mode of row p (the X*_{H-k} of k = 0 in the merge): in the
    // fused path held in a register by the thread whose first-stage butterfly
    // i starts row p (j = 0); otherwise staged in LDS.  (A staged copy in the
    // fused path would put the kernel's LDS just above a fraction of the
    // CU's, one resident workgroup fewer.)
    constexpr int NQ = fused ? E / R0 : 1;
    cx<T> nq[NQ];
    cx<T>* shn = nullptr;
    if constexpr (!fused) {
        __shared__ cx<T> shn_lds[P];
        shn = shn_lds;
    }
    {
        // all loads of the thread first, then the LDS stores
        cx<T> z[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int e = tid + i * NT;
            const long row = row0 + e / H;
            z[i] = row < nrows ? spec[row * H + e % H] : mk<T>(0, 0);
        }
        if constexpr (fused) {
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                const int b = tid + i * NT, p = b / (H / R0), j = b % (H / R0);
                nq[i] = (j == 0 && row0 + p < nrows) ? nyq[row0 + p] : mk<T>(0, 0);
            }
        } else {
            for (int p = tid; p < P; p += NT) shn[p] = (row0 + p < nrows) ? nyq[row0 + p] : mk<T>(0, 0);
        }
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int e = tid + i * NT;
            sh[lidx<H>(e / H, e % H)] = z[i];
        }
    }
    lds_sync();
    // Merge of the half spectrum into the H-point complex input of the
    // inverse transform: z_k = (X_k + X*_{H-k}) + i conj(W_N^k) (X_k - X*_{H-k}),
    // with X_H = xn the Nyquist mode at k = 0.
    auto merged_with = [&](int p, int k, const cx<T>& xn) -> cx<T> {
        const cx<T> Xk = sh[lidx<H>(p, k)];
        const cx<T> Xm = conj(k == 0 ? xn : sh[lidx<H>(p, H - k)]);
        const cx<T> Ep = Xk + Xm;
        const cx<T> wO = conj(stw[k]) * (Xk - Xm);
        return mk<T>(Ep.re - wO.im, Ep.im + wO.re);
    };
    const T inv = (T)(1.0 / s);
    auto scaled = [&](const cx<T>& v) -> cx<T> {
        if constexpr (sizeof(T) == 8) {
            return mk<T>(v.re / (T)s, v.im / (T)s);  // real_part(s): Re(z) / s (fourier.hh:31-37)
        } else {
            return mk<T>(v.re * inv, v.im * inv);
        }
    };
    if constexpr (fused) {
        // the merge fused into the first stage (its inputs read from the
        // loaded spectrum), the last stage stored straight to HBM
        cx<T> v[E];
#pragma unroll
        for (int i = 0; i < E / R0; ++i) {
            const int b = tid + i * NT, p = b / (H / R0), j = b % (H / R0);
#pragma unroll
            for (int t = 0; t < R0; ++t) v[i * R0 + t] = merged_with(p, j + t * (H / R0), nq[i]);
        }
        lds_sync();
#pragma unroll
        for (int i = 0; i < E / R0; ++i) {
            const int b = tid + i * NT, p = b / (H / R0), j = b % (H / R0);
            butterfly<T, H, R0, +1>(&v[i * R0], j, 1, stw, 2);
#pragma unroll
            for (int t = 0; t < R0; ++t) sh[lidx<H>(p, stock_out<H, R0>(j, 1, t))] = v[i * R0 + t];
        }
        lds_sync();
        int Ns = R0;
#pragma unroll
        for (int m = 0; m < NMID; ++m) {
            lds_stage<T, H, E, NT, +1, 8>(sh, Ns, stw, 2, tid);
            Ns *= 8;
        }
#pragma unroll
        for (int i = 0; i < E / 8; ++i) {
            const int b = tid + i * NT, p = b / (H / 8), j = b % (H / 8);
            cx<T>* w = &v[i * 8];
#pragma unroll
            for (int t = 0; t < 8; ++t) w[t] = sh[lidx<H>(p, j + t * (H / 8))];
            butterfly<T, H, 8, +1>(w, j, H / 8, stw, 2);
            const long row = row0 + p;
            if (row < nrows) {
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    *reinterpret_cast<cx<T>*>(out + row * N + 2 * stock_out<H, 8>(j, H / 8, t)) = scaled(w[t]);
            }
        }
    } else {
        cx<T> z[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int e = tid + i * NT;
            z[i] = merged_with(e / H, e % H, shn[e / H]);
        }
        lds_sync();
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int e = tid + i * NT;
            sh[lidx<H>(e / H, e % H)] = z[i];
        }
        lds_sync();
        lds_fft<T, H, E, NT, +1>(sh, stw, 2, tid);
        for (int e = tid; e < P * H; e += NT) {
            const int p = e / H, n = e % H;
            const long row = row0 + p;
            if (row >= nrows) continue;
            *reinterpret_cast<cx<T>*>(out + row * N + 2 * n) = scaled(sh[lidx<H>(p, n)]);
        }
    }
}

// --------------------------------------------------------------- columns
// Column tiles: C adjacent pencils (C * sizeof(cx<T>) contiguous bytes per
// row of the tile), NT threads; LDS holds the C x N tile.
#ifndef CPM_COL_F32
#define CPM_COL_F32 16
#endif
#ifndef CPM_COL_F64
#define CPM_COL_F64 8
#endif
#ifndef CPM_COL_NT
#define CPM_COL_NT 1024
#endif
#ifndef CPM_COL_LDS
#define CPM_COL_LDS 65536
#endif
#ifndef CPM_GREEN_C
#define CPM_GREEN_C 16
#endif
// CP: pencils per tile (0: the default for T)
template <typename T, int N, int CP = 0>
struct ColCfg {
    static constexpr int CS = (int)sizeof(cx<T>);
    static constexpr int Ct = CP ? CP : (sizeof(T) == 4 ? CPM_COL_F32 : CPM_COL_F64);
    static constexpr int C = (Ct * N * CS > CPM_COL_LDS) ? (CPM_COL_LDS / (N * CS)) : Ct;
    static constexpr int NT = N >= 8 ? ((C * N / 8 < CPM_COL_NT) ? C * N / 8 : CPM_COL_NT) : C;
    static constexpr int E = C * N / NT;
};

// Global <-> LDS copies of a column tile: each thread issues all its E loads
// before the first LDS store, so their latencies overlap.
template <typename T, int N>
__device__ __forceinline__ void col_load(cx<T>* sh, const cx<T>* __restrict__ data, const ColGeom& g, long bbase,
                                         long c0, int tid) {
    using C = ColCfg<T, N>;
    cx<T> v[C::E];
#pragma unroll
    for (int i = 0; i < C::E; ++i) {
        const int e = tid + i * C::NT;
        const int c = e % C::C, n = e / C::C;
        const long col = c0 + c;
        v[i] = col < g.nc ? data[bbase + col * g.cs + (long)n * g.es] : mk<T>(0, 0);
    }
#pragma unroll
    for (int i = 0; i < C::E; ++i) {
        const int e = tid + i * C::NT;
        sh[lidx<N>(e % C::C, e / C::C)] = v[i];
    }
}

template <typename T, int N>
__device__ __forceinline__ void col_store(const cx<T>* sh, cx<T>* __restrict__ data, const ColGeom& g, long bbase,
                                          long c0, int tid) {
    using C = ColCfg<T, N>;
    for (int e = tid; e < C::C * N; e += C::NT) {
        const int c = e % C::C, n = e / C::C;
        const long col = c0 + c;
        if (col < g.nc) data[bbase + col * g.cs + (long)n * g.es] = sh[lidx<N>(c, n)];
    }
}

// The tile of workgroup blockIdx.x of a launch over two parts (the main
// block of the half spectrum and the kx = N/2 plane): part, batch base and
// first pencil.
template <typename T, int N, int CP = 0>
struct ColTile {
    const ColPart<T>* part;
    long bbase, c0;
};
template <typename T, int N, int CP = 0>
__device__ __forceinline__ ColTile<T, N, CP> col_tile_at(const ColPart<T>& pa, const ColPart<T>& pb, uint32_t w) {
    const bool second = w >= (uint32_t)pa.nblk;
    const ColPart<T>& p = second ? pb : pa;
    if (second) w -= (uint32_t)pa.nblk;
    const uint32_t gx = (uint32_t)p.gx;
    return ColTile<T, N, CP>{&p, (long)(w / gx) * p.g.bs, (long)(w % gx) * ColCfg<T, N, CP>::C};
}
template <typename T, int N>
__device__ __forceinline__ ColTile<T, N> col_tile(const ColPart<T>& pa, const ColPart<T>& pb) {
    return col_tile_at<T, N>(pa, pb, blockIdx.x);
}

// Register-fused first and last Stockham stages (N >= 16, where every
// thread owns one radix-8 butterfly of the last stage: NT = C N / 8).
// Thread (pencil p = tid mod C, q = tid / C) loads the inputs of its
// first-stage butterflies straight from HBM, transforms them and writes them
// to LDS; the last stage writes the positions it reads (j + t N/8), so it
// runs on values read from LDS and stores them straight to HBM.  Only the
// middle stages exchange through LDS.
template <int N>
struct Fused {
    using S = Sched<N>;
    static constexpr bool ok = N >= 16;
    static constexpr int R0 = S::lead ? S::lead : 8;  // radix of the first stage
    static constexpr int mid = S::n8 - (S::lead ? 1 : 2);  // radix-8 stages between first and last
};

// Pencil of lane p of a tile starting at pencil c0.  A tile wider than the
// block (the last axis of a small 2-D mesh) repeats the block's last pencil:
// such a lane transforms the same values as that pencil's own lane, in the
// same operations, and stores the same results to the same addresses, so
// the fused path needs no per-lane branch around its HBM loads and stores.
template <typename T>
__device__ __forceinline__ long col_of(long c0, int p, const ColGeom& g) {
    const long c = c0 + p;
    return c < g.nc ? c : g.nc - 1;
}

// HBM loads of the first-stage inputs of thread (p, q) of tile t.
template <typename T, int N, int CP = 0>
__device__ __forceinline__ void col_first_load(const ColTile<T, N, CP>& t, int tid, cx<T>* v) {
    using C = ColCfg<T, N, CP>;
    static_assert(C::NT * 8 == C::C * N, "one radix-8 butterfly per thread and stage");
    constexpr int R0 = Fused<N>::R0, NB = 8 / R0;
    const ColPart<T>& P = *t.part;
    const int p = tid % C::C, q = tid / C::C;
    const long es = P.g.es;
    const cx<T>* __restrict__ src = P.data + t.bbase + col_of<T>(t.c0, p, P.g) * P.g.cs + (long)q * es;
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int k = 0; k < R0; ++k) v[u * R0 + k] = src[(long)(u * (N / 8) + k * (N / R0)) * es];
}

// The first stage (no twiddles) on those values, written to LDS.
template <typename T, int N, int DIR, int CP = 0>
__device__ __forceinline__ void col_first_store(cx<T>* sh, cx<T>* v, int tid) {
    using C = ColCfg<T, N, CP>;
    constexpr int R0 = Fused<N>::R0, NB = 8 / R0;
    const int p = tid % C::C, q = tid / C::C;
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        dft<T, R0, DIR>(&v[u * R0]);
        const int j = q + u * (N / 8);
#pragma unroll
        for (int k = 0; k < R0; ++k) sh[lidx<N>(p, j * R0 + k)] = v[u * R0 + k];
    }
}

// Persistent loop over the tiles of a column launch (fused path): while a
// tile is transformed, the HBM loads of the workgroup's next tile are in
// flight.  body(tile, v) gets the first-stage inputs in v.
template <typename T, int N, int CP, typename Body>
__device__ __forceinline__ void col_pipeline(const ColPart<T>& pa, const ColPart<T>& pb, int tid, Body body) {
    const uint32_t ntiles = (uint32_t)(pa.nblk + pb.nblk), G = gridDim.x;
    uint32_t w = blockIdx.x;
    if (w >= ntiles) return;
    // Two register sets in turn: no copy of a prefetched set waits for its
    // loads.  The prefetch is unconditional (past the last tile it reads the
    // current tile again, from the cache, and the values are dropped), so no
    // branch merges a path without it and the waits before a body count only
    // that body's own loads.
    cx<T> va[8], vb[8];
    col_first_load<T, N, CP>(col_tile_at<T, N, CP>(pa, pb, w), tid, va);
    for (;;) {
        col_first_load<T, N, CP>(col_tile_at<T, N, CP>(pa, pb, w + G < ntiles ? w + G : w), tid, vb);
        body(col_tile_at<T, N, CP>(pa, pb, w), va);
        lds_sync();  // LDS is refilled by the next tile
        w += G;
        if (w >= ntiles) break;
        col_first_load<T, N, CP>(col_tile_at<T, N, CP>(pa, pb, w + G < ntiles ? w + G : w), tid, va);
        body(col_tile_at<T, N, CP>(pa, pb, w), vb);
        lds_sync();
        w += G;
        if (w >= ntiles) break;
    }
}

// Twiddles of the radix-8 middle stages of a fused column transform, stage
// after stage: entry (t - 1) Ns + k of the stage of sub-length Ns holds
// W_N^(k t N / (8 Ns)) (k < Ns, 1 <= t < 8), the value the full table holds
// at k t N / (8 Ns).  The lanes of a middle stage read up to eight different
// k at once: in the full table those lie N / (8 Ns) entries apart, in a few
// LDS banks; here they are consecutive.
template <int N>
constexpr int mid_tw_size() {
    int n = 0, ns = Fused<N>::R0;
    for (int s = 0; s < Fused<N>::mid; ++s) {
        n += 7 * ns;
        ns *= 8;
    }
    return n > 0 ? n : 1;
}

template <typename T, int N>
__device__ __forceinline__ void fill_mid_tw(cx<T>* mt, const cx<T>* __restrict__ tw, int tid, int nt) {
    int off = 0, ns = Fused<N>::R0;
#pragma unroll
    for (int s = 0; s < Fused<N>::mid; ++s) {
        for (int e = tid; e < 7 * ns; e += nt) {
            const int t = e / ns + 1, k = e % ns;
            mt[off + e] = tw[k * t * (N / (8 * ns))];
        }
        off += 7 * ns;
        ns *= 8;
    }
}

// A radix-8 Stockham stage in LDS (lds_stage with R = 8, tws = 1) whose
// twiddles come from that stage's part of the middle-stage table.
template <typename T, int N, int E, int NT, int DIR>
__device__ __forceinline__ void lds_stage8_mt(cx<T>* sh, int Ns, const cx<T>* mt, int tid) {
    constexpr int NB = E / 8;
    constexpr int BPP = N / 8;
    cx<T> v[E];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int b = tid + i * NT;
        const int p = b / BPP, j = b % BPP;
#pragma unroll
        for (int t = 0; t < 8; ++t) v[i * 8 + t] = sh[lidx<N>(p, j + t * BPP)];
    }
    lds_sync();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int b = tid + i * NT;
        const int p = b / BPP, j = b % BPP;
        const int k = j & (Ns - 1);  // (all lanes multiply: see butterfly)
        if (Ns > 1) {
#pragma unroll
            for (int t = 1; t < 8; ++t) {
                cx<T> w = mt[(t - 1) * Ns + k];
                if (DIR > 0) w.im = -w.im;
                v[i * 8 + t] = v[i * 8 + t] * w;
            }
        }
        dft<T, 8, DIR>(&v[i * 8]);
#pragma unroll
        for (int t = 0; t < 8; ++t) sh[lidx<N>(p, stock_out<N, 8>(j, Ns, t))] = v[i * 8 + t];
    }
    lds_sync();
}

template <typename T, int N, int DIR, int CP = 0>
__device__ __forceinline__ void col_middle(cx<T>* sh, const cx<T>* mt, int tid) {
    using C = ColCfg<T, N, CP>;
    int Ns = Fused<N>::R0, off = 0;
#pragma unroll
    for (int s = 0; s < Fused<N>::mid; ++s) {
        lds_stage8_mt<T, N, C::E, C::NT, DIR>(sh, Ns, mt + off, tid);
        off += 7 * Ns;
        Ns *= 8;
    }
}

// The last stage of thread (p, j) in registers: v[t] = element j + t N/8.
template <typename T, int N, int DIR, int CP = 0>
__device__ __forceinline__ void col_last_regs(const cx<T>* sh, const cx<T>* __restrict__ tw, int tid, cx<T>* v) {
    using C = ColCfg<T, N, CP>;
    const int p = tid % C::C, j = tid / C::C;
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = sh[lidx<N>(p, j + t * (N / 8))];
    butterfly<T, N, 8, DIR>(v, j, N / 8, tw, 1);
}

template <typename T, int N, int CP = 0>
__device__ __forceinline__ void col_store_regs(const cx<T>* v, cx<T>* __restrict__ data, const ColGeom& g, long bbase,
                                               long c0, int tid) {
    using C = ColCfg<T, N, CP>;
    const int p = tid % C::C, j = tid / C::C;
    const long es = g.es;
    cx<T>* __restrict__ dst = data + bbase + col_of<T>(c0, p, g) * g.cs + (long)j * es;
#pragma unroll
    for (int t = 0; t < 8; ++t) dst[(long)(t * (N / 8)) * es] = v[t];
}

template <typename T, int N, int DIR>
__global__ __launch_bounds__((ColCfg<T, N>::NT)) void k_cols(ColPart<T> pa, ColPart<T> pb,
                                                           const cx<T>* __restrict__ tw) {
    using C = ColCfg<T, N>;
    __shared__ cx<T> sh[C::C * N];
    __shared__ cx<T> stw[N];
    __shared__ cx<T> smt[mid_tw_size<N>()];
    const int tid = threadIdx.x;
    for (int m = tid; m < N; m += C::NT) stw[m] = tw[m];
    if constexpr (Fused<N>::ok) {
        fill_mid_tw<T, N>(smt, tw, tid, C::NT);
        // One tile per workgroup: with a single FFT per tile, prefetching
        // the next tile costs more in registers than it hides.  The tiles are
        // taken from the last slab backwards, against the row passes, which
        // walk the slabs upwards: the forward sweep first reads what the row
        // pass wrote last, still in the Infinity Cache (256 MB, half of a
        // 512^3 spectrum), and the backward sweep writes last the slabs the
        // inverse row pass reads first.
        const ColTile<T, N> t = col_tile_at<T, N>(pa, pb, gridDim.x - 1 - blockIdx.x);
        cx<T> v[8];
        col_first_load<T, N>(t, tid, v);
        col_first_store<T, N, DIR>(sh, v, tid);
        lds_sync();
        col_middle<T, N, DIR>(sh, smt, tid);
        col_last_regs<T, N, DIR>(sh, stw, tid, v);
        col_store_regs<T, N>(v, t.part->data, t.part->g, t.bbase, t.c0, tid);
    } else {
        const ColTile<T, N> t = col_tile<T, N>(pa, pb);
        const ColPart<T>& p = *t.part;
        col_load<T, N>(sh, p.data, p.g, t.bbase, t.c0, tid);
        lds_sync();
        lds_fft<T, N, C::E, C::NT, DIR>(sh, stw, 1, tid);
        col_store<T, N>(sh, p.data, p.g, t.bbase, t.c0, tid);
    }
}

// Signed mode number of index n on an N mesh: n > N/2 maps to n - N, and
// N/2 stays +N/2 (src/base/fourier.hh:52).
__device__ __forceinline__ int signed_mode(int n, int N) { return n > N / 2 ? n - N : n; }

// Green's function -1/|k|^2 with physical k = n * 2 pi / L (src/base/fourier.hh:102-109;
// applied at src/run/nbody.hh:106-108), zero at k = 0.
template <typename T>
__device__ __forceinline__ T green_value(int ix, int iy, int iz, int dim, int N, T kf) {
    if (ix == 0 && iy == 0 && iz == 0) return (T)0;
    T kx = (T)signed_mode(ix, N) * kf;
    T ky = (T)signed_mode(iy, N) * kf;
    T v = (T)0;
    v += kx * kx;
    v += ky * ky;
    if (dim == 3) {
        T kz = (T)signed_mode(iz, N) * kf;
        v += kz * kz;
    }
    if constexpr (sizeof(T) == 4) {
        return -__builtin_amdgcn_rcpf(v);  // 1 ulp reciprocal: below the FP32 FFT's own rounding
    } else {
        return (T)-1. / v;
    }
}

// Last-axis forward transform, multiplication by the Green's function and
// last-axis backward transform, in one read and one write of the spectrum.
template <typename T, int N>
__global__ __launch_bounds__((ColCfg<T, N, CPM_GREEN_C>::NT)) void k_cols_green(ColPart<T> pa, ColPart<T> pb,
                                                                 const cx<T>* __restrict__ tw) {
    constexpr int CP = CPM_GREEN_C;
    using C = ColCfg<T, N, CP>;
    __shared__ cx<T> sh[C::C * N];
    __shared__ cx<T> stw[N];
    __shared__ cx<T> smt[mid_tw_size<N>()];
    const int tid = threadIdx.x;
    const int H = N / 2;
    for (int m = tid; m < N; m += C::NT) stw[m] = tw[m];
    if constexpr (Fused<N>::ok) {
        fill_mid_tw<T, N>(smt, tw, tid, C::NT);
        // forward: first stage on load, middle stages, last stage in
        // registers; the Green's function on the registers; backward: first
        // stage in registers (when it is radix 8 it reads exactly these
        // positions), middle stages, last stage on store.
        col_pipeline<T, N, CP>(pa, pb, tid, [&](const ColTile<T, N, CP>& t, cx<T>* v) {
        const GreenGeom gg = t.part->gg;
        const T kf = (T)gg.kf;
        col_first_store<T, N, -1, CP>(sh, v, tid);
        lds_sync();
        col_middle<T, N, -1, CP>(sh, smt, tid);
        col_last_regs<T, N, -1, CP>(sh, stw, tid, v);
        const int pc = tid % C::C, j = tid / C::C;
        const long col = col_of<T>(t.c0, pc, t.part->g);
        // green_value's sum kx^2 [+ ky^2] + kn^2 for the pencil's fixed modes
        // and the mode n along it: the fixed part once per tile
        int ix, iy = 0;
        if (gg.dim == 3) {
            ix = gg.nyq ? H : (int)(col & (H - 1));
            iy = gg.nyq ? (int)col : (int)(col / H);
        } else {
            ix = gg.nyq ? H : (int)col;
        }
        const T kx = (T)signed_mode(ix, N) * kf;
        T k2 = (T)0;
        k2 += kx * kx;
        if (gg.dim == 3) {
            const T ky = (T)signed_mode(iy, N) * kf;
            k2 += ky * ky;
        }
        const bool fixed_zero = ix == 0 && iy == 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int n = j + q * (N / 8);
            const T kn = (T)signed_mode(n, N) * kf;
            const T kk = k2 + kn * kn;
            T gv;
            if constexpr (sizeof(T) == 4) {
                gv = -__builtin_amdgcn_rcpf(kk);
            } else {
                gv = (T)-1. / kk;
            }
            v[q] = (fixed_zero && n == 0) ? mk<T>(0, 0) : scale(v[q], gv);
        }
        lds_sync();  // every thread has read its last-stage inputs
        if constexpr (Fused<N>::R0 == 8) {
            dft<T, 8, +1>(v);  // backward first stage, butterfly j, no twiddles
#pragma unroll
            for (int q = 0; q < 8; ++q) sh[lidx<N>(pc, j * 8 + q)] = v[q];
            lds_sync();
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) sh[lidx<N>(pc, j + q * (N / 8))] = v[q];
            lds_sync();
            lds_stage<T, N, C::E, C::NT, +1, Fused<N>::R0>(sh, 1, stw, 1, tid);
        }
        col_middle<T, N, +1, CP>(sh, smt, tid);
        col_last_regs<T, N, +1, CP>(sh, stw, tid, v);
        col_store_regs<T, N, CP>(v, t.part->data, t.part->g, t.bbase, t.c0, tid);
        });
        return;
    }
    const ColTile<T, N, CP> t = col_tile_at<T, N, CP>(pa, pb, blockIdx.x);
    const ColPart<T>& p = *t.part;
    const GreenGeom gg = p.gg;
    const T kf = (T)gg.kf;
    col_load<T, N>(sh, p.data, p.g, t.bbase, t.c0, tid);
    lds_sync();
    lds_fft<T, N, C::E, C::NT, -1>(sh, stw, 1, tid);
    for (int e = tid; e < C::C * N; e += C::NT) {
        const int c = e / N, n = e % N;
        const long col = t.c0 + c;
        int ix, iy, iz = 0;
        if (gg.dim == 3) {
            ix = gg.nyq ? H : (int)(col & (H - 1));
            iy = gg.nyq ? (int)col : (int)(col / H);
            iz = n;
        } else {
            ix = gg.nyq ? H : (int)col;
            iy = n;
        }
        const T gv = green_value<T>(ix, iy, iz, gg.dim, N, kf);
        const int li = lidx<N>(c, n);
        sh[li] = (ix == 0 && iy == 0 && iz == 0) ? mk<T>(0, 0) : scale(sh[li], gv);
    }
    lds_sync();
    lds_fft<T, N, C::E, C::NT, +1>(sh, stw, 1, tid);
    col_store<T, N>(sh, p.data, p.g, t.bbase, t.c0, tid);
}



// ------------------------------------------------------------ dispatch
#define CPM_FOR_N(MACRO) \
    MACRO(4) MACRO(8) MACRO(16) MACRO(32) MACRO(64) MACRO(128) MACRO(256) MACRO(512) MACRO(1024)

template <typename T>
void launch_rows_r2c(int N, const T* in, cx<T>* spec, cx<T>* nyq, long nrows, T shift,
                     const double* mean_ptr, const cx<T>* tw, hipStream_t st) {
    switch (N) {
#define CASE(n)                                                                        \
    case n: {                                                                          \
        using C = RowCfg<T, n>;                                                        \
        dim3 grid((unsigned)((nrows + C::P - 1) / C::P));                              \
        hipLaunchKernelGGL((k_rows_r2c<T, n>), grid, dim3(C::NT), 0, st, in, spec, nyq, \
                           nrows, shift, mean_ptr, tw);                                \
        break;                                                                         \
    }
        CPM_FOR_N(CASE)
#undef CASE
        default:
            throw Error(-2, "unsupported FFT length " + std::to_string(N));
    }
    CPM_LAUNCH_CHECK();
}

template <typename T>
void launch_rows_c2r(int N, const cx<T>* spec, const cx<T>* nyq, T* out, long nrows, double s,
                     const cx<T>* tw, hipStream_t st) {
    switch (N) {
#define CASE(n)                                                                               \
    case n: {                                                                                 \
        using C = RowCfg<T, n>;                                                               \
        dim3 grid((unsigned)((nrows + C::P - 1) / C::P));                                     \
        hipLaunchKernelGGL((k_rows_c2r<T, n>), grid, dim3(C::NT), 0, st, spec, nyq, out, nrows, \
                           s, tw);                                                            \
        break;                                                                                \
    }
        CPM_FOR_N(CASE)
#undef CASE
        default:
            throw Error(-2, "unsupported FFT length " + std::to_string(N));
    }
    CPM_LAUNCH_CHECK();
}

// Workgroups of a Green's-function launch: one per tile, or for the fused
// (persistent) kernel as many as the GPU holds at once.
template <typename T, int N, int CP>
dim3 col_grid(long ntiles) {
    if (!Fused<N>::ok) return dim3((unsigned)ntiles);
    static int per_cu = 0;  // workgroups of this kernel one CU holds (registers, LDS, threads)
    if (!per_cu) {
        CPM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)&k_cols_green<T, N>,
                                                             ColCfg<T, N, CP>::NT, 0));
        if (per_cu < 1) per_cu = 1;
    }
    const long g = device_cus() * per_cu;
    return dim3((unsigned)(ntiles < g ? ntiles : g));
}

template <typename T>
void launch_cols(int N, int dir, const ColPart<T>& a, const ColPart<T>& b, const cx<T>* tw, hipStream_t st) {
    const long ntiles = a.nblk + b.nblk;
    switch (N) {
#define CASE(n)                                                                             \
    case n: {                                                                               \
        using C = ColCfg<T, n>;                                                             \
        const dim3 grid((unsigned)ntiles);                                                  \
        if (dir < 0)                                                                        \
            hipLaunchKernelGGL((k_cols<T, n, -1>), grid, dim3(C::NT), 0, st, a, b, tw);     \
        else                                                                                \
            hipLaunchKernelGGL((k_cols<T, n, +1>), grid, dim3(C::NT), 0, st, a, b, tw);     \
        break;                                                                              \
    }
        CPM_FOR_N(CASE)
#undef CASE
        default:
            throw Error(-2, "unsupported FFT length " + std::to_string(N));
    }
    CPM_LAUNCH_CHECK();
}

template <typename T>
void launch_cols_green(int N, const ColPart<T>& a, const ColPart<T>& b, const cx<T>* tw, hipStream_t st) {
    const long ntiles = a.nblk + b.nblk;
    switch (N) {
#define CASE(n)                                                                                \
    case n: {                                                                                  \
        using C = ColCfg<T, n, CPM_GREEN_C>;                                                   \
        const dim3 grid = col_grid<T, n, CPM_GREEN_C>(ntiles);                                 \
        hipLaunchKernelGGL((k_cols_green<T, n>), grid, dim3(C::NT), 0, st, a, b, tw);          \
        break;                                                                                 \
    }
        CPM_FOR_N(CASE)
#undef CASE
        default:
            throw Error(-2, "unsupported FFT length " + std::to_string(N));
    }
    CPM_LAUNCH_CHECK();
}

// One part of a column launch: gx workgroups per batch, gx * nb in all.
template <typename T, int N, int CP>
ColPart<T> col_part_n(cx<T>* d, const ColGeom& g, const GreenGeom& gg) {
    const long gx = (g.nc + ColCfg<T, N, CP>::C - 1) / ColCfg<T, N, CP>::C;
    return ColPart<T>{d, g, gg, gx, gx * g.nb};
}
// green: a part of the Green's-function pass (its own tile width)
template <typename T>
ColPart<T> col_part(int N, cx<T>* d, const ColGeom& g, const GreenGeom& gg = GreenGeom{}, bool green = false) {
    switch (N) {
#define CASE(n) \
    case n:     \
        return green ? col_part_n<T, n, CPM_GREEN_C>(d, g, gg) : col_part_n<T, n, 0>(d, g, gg);
        CPM_FOR_N(CASE)
#undef CASE
        default:
            throw Error(-2, "unsupported FFT length " + std::to_string(N));
    }
}

// ------------------------------------------------------------ the plan
template <typename T>
FFTPlan<T>::FFTPlan(int dim_, int N_) : dim(dim_), N(N_), tw(nullptr) {
    if (dim != 2 && dim != 3) throw Error(-2, "FFT supports 2-D and 3-D meshes only");
    if (N < 4 || N > 1024 || (N & (N - 1))) throw Error(-2, "mesh size must be a power of two in [4, 1024]");
    H = N / 2;
    nrows = 1;
    for (int k = 1; k < dim; ++k) nrows *= N;
    M = nrows * N;
    spec_elems = (size_t)H * nrows;
    nyq_elems = (size_t)nrows;
    std::vector<cx<T>> h(N);
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int m = 0; m < N; ++m) {
        long double a = twopi * (long double)m / (long double)N;
        h[m] = mk<T>((T)cosl(a), (T)(-sinl(a)));
    }
    void* q = nullptr;
    device_alloc(&q, sizeof(cx<T>) * N);
    tw = static_cast<cx<T>*>(q);
    const hipError_t e = hipMemcpy(tw, h.data(), sizeof(cx<T>) * N, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        device_free(tw, sizeof(cx<T>) * N);
        tw = nullptr;
        throw Error(-3, std::string("hipMemcpy of the FFT twiddles: ") + hipGetErrorString(e));
    }
}

template <typename T>
FFTPlan<T>::~FFTPlan() {
    device_free(tw, sizeof(cx<T>) * N);
}

template <typename T>
void FFTPlan<T>::geoms(ColGeom* ymain, ColGeom* ynyq, ColGeom* zmain, ColGeom* znyq) const {
    if (dim == 3) {
        // main: index kx + H*(ky + N*kz); nyq: ky + N*kz
        *ymain = ColGeom{(long)H, 1, (long)H, (long)H * N, N};
        *ynyq = ColGeom{1, (long)N, (long)N, 0, 1};
        *zmain = ColGeom{(long)H * N, 1, (long)H * N, 0, 1};
        *znyq = ColGeom{(long)N, 1, (long)N, 0, 1};
    } else {
        // main: kx + H*ky; nyq: ky.  The y pass is the last axis.
        *ymain = ColGeom{(long)H, 1, (long)H, 0, 1};
        *ynyq = ColGeom{1, (long)N, 1, 0, 1};
        *zmain = *ymain;
        *znyq = *ynyq;
    }
}

template <typename T>
void FFTPlan<T>::forward(const T* in, cx<T>* spec, cx<T>* nyq, T shift, const double* mean_ptr,
                         hipStream_t st) const {
    launch_rows_r2c<T>(N, in, spec, nyq, nrows, shift, mean_ptr, tw, st);
    ColGeom ym, yn, zm, zn;
    geoms(&ym, &yn, &zm, &zn);
    launch_cols<T>(N, -1, col_part(N, spec, ym), col_part(N, nyq, yn), tw, st);
    if (dim == 3) launch_cols<T>(N, -1, col_part(N, spec, zm), col_part(N, nyq, zn), tw, st);
}

template <typename T>
void FFTPlan<T>::backward(cx<T>* spec, cx<T>* nyq, T* out, double s, hipStream_t st) const {
    ColGeom ym, yn, zm, zn;
    geoms(&ym, &yn, &zm, &zn);
    if (dim == 3) launch_cols<T>(N, +1, col_part(N, spec, zm), col_part(N, nyq, zn), tw, st);
    launch_cols<T>(N, +1, col_part(N, spec, ym), col_part(N, nyq, yn), tw, st);
    launch_rows_c2r<T>(N, spec, nyq, out, nrows, s, tw, st);
}

template <typename T>
void FFTPlan<T>::poisson_stage(int stage, int part, const T* rho, T* phi, cx<T>* spec, cx<T>* nyq, T shift,
                               const double* mean_ptr, double kf, double s, hipStream_t st) const {
    ColGeom ym, yn, zm, zn;
    geoms(&ym, &yn, &zm, &zn);
    const GreenGeom gmain{dim, 0, kf}, gnyq{dim, 1, kf};
    switch (stage) {
        case 0:
            if (part == 0) launch_rows_r2c<T>(N, rho, spec, nyq, nrows, shift, mean_ptr, tw, st);
            break;
        case 1:
            if (dim == 3 && part == 0) launch_cols<T>(N, -1, col_part(N, spec, ym), col_part(N, nyq, yn), tw, st);
            break;
        case 2:
            if (part == 0)
                launch_cols_green<T>(N, col_part(N, spec, dim == 3 ? zm : ym, gmain, true),
                                     col_part(N, nyq, dim == 3 ? zn : yn, gnyq, true), tw, st);
            break;
        case 3:
            if (dim == 3 && part == 0) launch_cols<T>(N, +1, col_part(N, spec, ym), col_part(N, nyq, yn), tw, st);
            break;
        case 4:
            if (part == 0) launch_rows_c2r<T>(N, spec, nyq, phi, nrows, s, tw, st);
            break;
        default:
            throw Error(-1, "bad Poisson stage");
    }
}

template <typename T>
void FFTPlan<T>::poisson(const T* rho, T* phi, cx<T>* spec, cx<T>* nyq, T shift,
                         const double* mean_ptr, double kf, double s, hipStream_t st) const {
    for (int stage = 0; stage < 5; ++stage)
        for (int part = 0; part < 2; ++part)
            poisson_stage(stage, part, rho, phi, spec, nyq, shift, mean_ptr, kf, s, st);
}

template struct FFTPlan<float>;
template struct FFTPlan<double>;

}  // namespace cpm
