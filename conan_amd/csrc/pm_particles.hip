// pm_particles.hip -- per-particle kernels of the PM step on gfx950:
// cell keys, the deterministic CIC deposit, the fused FD-gradient gather +
// kick (+ drift), the standalone drift, the Zel'dovich start, and the FP64
// reductions.
//
// Two forms share the kernels through a position policy:
//   F64: positions are IEEE doubles, unwrapped, exactly as the reference keeps
//        them (src/run/nbody.hh:65); every per-particle operation repeats the
//        reference's arithmetic in its order, so keys, weights, the gather,
//        the kick and the drift are bit-identical to the reference for equal
//        inputs.
//   FX:  positions are signed 32-bit fixed-point coordinates (x = q * L / 2^30,
//        unwrapped within [-2L, 2L)), so the periodic wrap is exact integer
//        arithmetic; mesh and gather in FP32, momenta and reductions in FP64.
// Particle arrays are SoA: component k of particle i at [k * np + i].
#include "pm_brick.hpp"

#include <algorithm>
#include <cstddef>
#include <type_traits>

namespace cpm {

// ------------------------------------------------------------ keys
template <typename PP, int R>
__global__ __launch_bounds__(256) void k_keys(const typename PP::Pos* __restrict__ X, long np, int lgN,
                                               PosArgs pa, uint32_t* __restrict__ keys,
                                               uint32_t* __restrict__ pids) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int mask = (1 << lgN) - 1;
    uint32_t key = 0;
    int oc[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int o;
        typename PP::Mesh f;
        PP::dep(X[k * np + i], pa, o, f);
        oc[k] = o & mask;
        key |= (uint32_t)oc[k] << (k * lgN);  // Box::idx of the origin (box.hh:75-81)
    }
    keys[i] = pa.brick ? brick_key(oc[0], oc[1], oc[2], R, pa) : key;
    pids[i] = (uint32_t)i;
}

// FX brick keys, for a full sort (prevkeys == nullptr: pids = iota) or an
// incremental re-sort (movers counted into their new group and out of their old one).
template <int R>
__global__ __launch_bounds__(256) void k_keys_fx(const uint32_t* __restrict__ X, long np, int lgN, PosArgs pa,
                                                  const uint32_t* __restrict__ prevkeys, uint32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ pids, uint32_t* __restrict__ in_count,
                                                  uint32_t* __restrict__ out_count) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    const int mask = (1 << lgN) - 1;
    int oc[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < R; ++k) oc[k] = (int)(X[k * np + s] >> pa.fshift) & mask;
    const uint32_t key = brick_key(oc[0], oc[1], oc[2], R, pa);
    keys[s] = key;
    fx_note_key(key, s, prevkeys, pids, in_count, out_count);
}

// row_start[r] = first sorted position whose key lies in mesh row r or later.
__global__ __launch_bounds__(256) void k_row_bounds(const uint32_t* __restrict__ keys, long np, int lgN,
                                                    long nrows, uint32_t* __restrict__ row_start) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > np) return;
    const long rp = (s == 0) ? -1 : (long)(keys[s - 1] >> lgN);
    const long rc = (s == np) ? nrows : (long)(keys[s] >> lgN);
    for (long r = rp + 1; r <= rc; ++r) row_start[r] = (uint32_t)s;
}

// Sum of v over the NT threads of a workgroup, returned to thread 0, in a
// fixed order (wave trees, then the waves in order); red: NT / 64 LDS words.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red, int tid) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double t = 0;
    if (tid == 0)
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) t += red[w];
    return t;
}

// ------------------------------------------------------------ deposit
// One workgroup per mesh row.  Cell c receives, from each origin cell
// c - block[j], the weight m * prod_k A[bit_k(j)][k] of every particle there
// (src/run/numeric.hh:124-137).  The 2^R per-origin lists are each sorted by
// particle id, and they are merged by id, so each cell sums its contributions
// in ascending particle order: the order of the reference's serial loop.
static constexpr int DEP_NT = 256;

// The fractional cell coordinates of every particle, in sorted order
// (frac[k * np + s] for particle pids[s]): computed once per particle by the
// same expression the merge would evaluate (PP::dep), so that the merge reads
// them in sorted order instead of loading a particle's position 2^R times at
// its storage index.
// Also, when Xs and inv are given, the positions in sorted order and the
// inverse order (inv[pids[s]] = s), for the FP64 kick's two passes.
template <typename PP, int R>
__global__ __launch_bounds__(256) void k_dep_frac(const uint32_t* __restrict__ pids,
                                                   const typename PP::Pos* __restrict__ X, long np, PosArgs pa,
                                                   typename PP::Mesh* __restrict__ frac,
                                                   typename PP::Pos* __restrict__ Xs, uint32_t* __restrict__ inv) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    const long m = pids[s];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int o;
        typename PP::Mesh f;
        const typename PP::Pos x = X[k * np + m];
        PP::dep(x, pa, o, f);
        if (frac) frac[k * np + s] = f;
        if (Xs) Xs[k * np + s] = x;
    }
    if (inv) inv[m] = (uint32_t)s;
}

template <typename PP, int R>
__global__ __launch_bounds__(DEP_NT) void k_deposit(const uint32_t* __restrict__ row_start,
                                                     const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ pids,
                                                     const typename PP::Mesh* __restrict__ frac, long np,
                                                     int lgN, typename PP::Mesh mass,
                                                     typename PP::Mesh* __restrict__ mesh,
                                                     double* __restrict__ row_partial) {
    using M = typename PP::Mesh;
    constexpr int NL = 1 << (R - 1);
    constexpr int NC = 1 << R;
    __shared__ uint32_t cs[NL][1025];
    __shared__ double red[DEP_NT / 64];
    const int N = 1 << lgN, msk = N - 1;
    const long row = blockIdx.x;
    const int cy = (int)(row & msk);
    const int cz = (R == 3) ? (int)(row >> lgN) : 0;
    const int tid = threadIdx.x;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        const int by = l & 1, bz = (l >> 1) & 1;
        long orow = (long)((cy - by) & msk);
        if (R == 3) orow += (long)((cz - bz) & msk) << lgN;
        const uint32_t s0 = row_start[orow], s1 = row_start[orow + 1];
        for (uint32_t s = s0 + tid; s <= s1; s += DEP_NT) {
            const int cp = (s == s0) ? -1 : (int)(keys[s - 1] & msk);
            const int cc = (s == s1) ? N : (int)(keys[s] & msk);
            for (int c = cp + 1; c <= cc; ++c) cs[l][c] = s;
        }
    }
    __syncthreads();
    double rsum = 0;
    for (int cx = tid; cx < N; cx += DEP_NT) {
        uint32_t cur[NC], end[NC], hp[NC];
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            const int ox = (cx - (j & 1)) & msk;
            cur[j] = cs[j >> 1][ox];
            end[j] = cs[j >> 1][ox + 1];
            hp[j] = cur[j] < end[j] ? pids[cur[j]] : 0xFFFFFFFFu;
        }
        M acc = 0;
        while (true) {
            uint32_t m = 0xFFFFFFFFu;
            int jm = -1;
#pragma unroll
            for (int j = 0; j < NC; ++j)
                if (hp[j] < m) {
                    m = hp[j];
                    jm = j;
                }
            if (jm < 0) break;
            uint32_t sp = 0;  // sorted position of the particle taken (m = pids[sp])
#pragma unroll
            for (int j = 0; j < NC; ++j)
                if (j == jm) sp = cur[j];
            M z = mass;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const M f = frac[k * np + sp];
                z *= ((jm >> k) & 1) ? f : (M)1 - f;
            }
            acc += z;
#pragma unroll
            for (int j = 0; j < NC; ++j)
                if (j == jm) {
                    ++cur[j];
                    hp[j] = cur[j] < end[j] ? pids[cur[j]] : 0xFFFFFFFFu;
                }
        }
        mesh[row * N + cx] = acc;
        rsum += (double)acc;
    }
    const double tot = block_sum<DEP_NT>(rsum, red, tid);
    if (tid == 0) row_partial[row] = tot;
}

// 128-bit fixed point of the FP64 deposit (k_deposit_f64_brick): a weight z
// in [0, 2^31) rounded down to a multiple of 2^-96, held as two 64-bit words
// (the carry of the low word added to the high one).  Integer addition is
// associative, so a cell's sum is the exact sum of its rounded weights in
// any order: deterministic, and within an ulp or two of the reference's
// serial double sum (which rounds at every addition).
static constexpr double FX128_ONE = 79228162514264337593543950336.0;  // 2^96
__device__ __forceinline__ void fx128_split(double z, unsigned long long& hi, unsigned long long& lo) {
    // z in [0, 2^31): hi = floor(z 2^32), lo = floor(frac(z 2^32) 2^64)
    const double t = z * 4294967296.0;
    const double h = floor(t);
    hi = (unsigned long long)h;
    lo = (unsigned long long)((t - h) * 18446744073709551616.0);
}
__device__ __forceinline__ double fx128_value(unsigned long long hi, unsigned long long lo) {
    return (double)hi * (1.0 / 4294967296.0) + (double)lo * (1.0 / FX128_ONE);
}

// ------------------------------------------------------------ FX: persistent brick order
// Particles move less than a cell per step, so `perm` (pm_brick.hpp) is close
// to the identity and reads through it are nearly contiguous.

// Deposit, one workgroup per brick.  The brick's cells receive the CIC
// weights of the particles of the brick and of its lower neighbours (a
// particle reaches its origin cell and the next one along each axis); only
// weights that land inside the brick are added, into an LDS copy of the
// brick.  The accumulators are 64-bit fixed point (2^-40 mass units):
// integer addition is associative, so the sum does not depend on the order of
// the LDS atomics and the deposit is deterministic; each cell is rounded once
// to the output type (FP32 for the step) on the way out and written by
// exactly one workgroup.
template <int R, typename MT>
__global__ __launch_bounds__(FXB_NT) void k_deposit_fx_brick(const uint32_t* __restrict__ bstart,
                                                             const uint32_t* __restrict__ perm,
                                                             const uint32_t* __restrict__ X, long np, int lgN,
                                                             PosArgs pa, float mass, MT* __restrict__ mesh,
                                                             double* __restrict__ partial) {
    __shared__ unsigned long long acc[FXB_CELLS];
    __shared__ double red[FXB_NT / 64];
    const int N = 1 << lgN, msk = N - 1;
    const uint32_t bid = blockIdx.x;
    const BrickCoord bc = brick_coord(bid, pa);
    const int ncell = bc.B[0] * bc.B[1] * bc.B[2];
    const int tid = threadIdx.x;
    for (int e = tid; e < ncell; e += FXB_NT) acc[e] = 0ull;
    __shared__ BrickRanges rg;
    brick_ranges<R>(rg, bstart, bc, pa, tid);
    {
        const uint32_t total = rg.roff[rg.n];
        // DEP_BATCH particles per thread at a time: their perm and position
        // loads are all issued before the first is used.
        constexpr int DEP_BATCH = 4;
        for (uint32_t t0 = tid; t0 < total; t0 += DEP_BATCH * FXB_NT) {
            uint32_t pi[DEP_BATCH];
#pragma unroll
            for (int q = 0; q < DEP_BATCH; ++q) {
                const uint32_t t = t0 + q * FXB_NT;
                pi[q] = t < total ? perm[brick_range_pos(rg, t)] : 0u;
            }
            uint32_t xq[DEP_BATCH][3];
#pragma unroll
            for (int q = 0; q < DEP_BATCH; ++q)
#pragma unroll
                for (int k = 0; k < 3; ++k) xq[q][k] = (k < R && t0 + q * FXB_NT < total) ? X[k * np + pi[q]] : 0u;
#pragma unroll
            for (int q = 0; q < DEP_BATCH; ++q) {
                if (t0 + q * FXB_NT >= total) break;
                int o[3] = {0, 0, 0};
                float f[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int k = 0; k < R; ++k) PosFX::dep(xq[q][k], pa, o[k], f[k]);
#pragma unroll
                for (int j = 0; j < (1 << R); ++j) {
                    int l[3];
                    bool in = true;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        l[k] = (k < R) ? (((o[k] + ((j >> k) & 1)) - bc.x0[k]) & msk) : 0;
                        in &= l[k] < bc.B[k];
                    }
                    if (!in) continue;
                    float w = mass;
#pragma unroll
                    for (int k = 0; k < R; ++k) w *= ((j >> k) & 1) ? f[k] : 1.f - f[k];
                    atomicAdd(&acc[l[0] + bc.B[0] * (l[1] + bc.B[1] * l[2])], fx_quantum(w));
                }
            }
        }
    }
    __syncthreads();
    double rsum = 0;
    if (R == 3 && bc.B[0] * bc.B[1] == FXB_NT) {
        // a whole brick layer per round: thread tid owns the column (lx, ly)
        // and walks it along the last axis (cells tid + FXB_NT lz, as below)
        const int lx = tid & (bc.B[0] - 1), ly = tid >> pa.lb[0];
        MT* __restrict__ out = mesh + (long)(bc.x0[0] + lx) + (long)N * ((bc.x0[1] + ly) + (long)N * bc.x0[2]);
        const long plane = (long)N * N;
        for (int lz = 0; lz < bc.B[2]; ++lz) {
            const MT v = (MT)((double)(long long)acc[tid + lz * FXB_NT] * (1.0 / FX_ONE));
            out[lz * plane] = v;
            rsum += (double)v;
        }
    } else {
        for (int e = tid; e < ncell; e += FXB_NT) {
            const int lx = e & (bc.B[0] - 1), ly = (e >> pa.lb[0]) & (bc.B[1] - 1), lz = e >> (pa.lb[0] + pa.lb[1]);
            const MT v = (MT)((double)(long long)acc[e] * (1.0 / FX_ONE));
            const long idx = (long)(bc.x0[0] + lx) + (long)N * ((bc.x0[1] + ly) + (long)N * (bc.x0[2] + lz));
            mesh[idx] = v;
            rsum += (double)v;
        }
    }
    if (partial) {  // (uniform)
        const double tot = block_sum<FXB_NT>(rsum, red, tid);
        if (tid == 0) partial[bid] = tot;
    }
}

// FX force at one particle: the trilinear interpolation (weights from the
// fractions a1) of the 4-point FD gradient at the 2^R nodes around its origin
// o.  ld(tx, ty, tz) returns the potential at node o - 2 + (tx, ty, tz)
// (tz ignored in 2-D): the brick kernels read it from their LDS tile, the
// others from the mesh, with the same operations in the same order, so every
// FX path computes the same force bit for bit.
template <int R, typename Ld>
__device__ __forceinline__ void force_fx(const float* a1, Ld ld, float inv12res, float* F) {
    constexpr int NC = 1 << R;
    float z[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        float w = 1.f;
#pragma unroll
        for (int k = 0; k < R; ++k) w *= ((c >> k) & 1) ? a1[k] : 1.f - a1[k];
        z[c] = w;
    }
    const int zc = (R == 3) ? 2 : 0;  // tz of the origin's own plane
    F[0] = F[1] = F[2] = 0.f;
    // axis 0: lines along x through the (y, z) corners
#pragma unroll
    for (int l = 0; l < NC / 2; ++l) {
        const int by = l & 1, bz = (l >> 1) & 1;
        float f[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) f[t] = ld(t, 2 + by, zc + bz);
#pragma unroll
        for (int bx = 0; bx < 2; ++bx)
            F[0] = fmaf(fd4(f[bx], f[bx + 1], f[bx + 3], f[bx + 4], inv12res), z[bx | (by << 1) | (bz << 2)], F[0]);
    }
#pragma unroll
    for (int l = 0; l < NC / 2; ++l) {
        const int bx = l & 1, bz = (l >> 1) & 1;
        float f[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) f[t] = ld(2 + bx, t, zc + bz);
#pragma unroll
        for (int by = 0; by < 2; ++by)
            F[1] = fmaf(fd4(f[by], f[by + 1], f[by + 3], f[by + 4], inv12res), z[bx | (by << 1) | (bz << 2)], F[1]);
    }
    if constexpr (R == 3) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const int bx = l & 1, by = (l >> 1) & 1;
            float f[6];
#pragma unroll
            for (int t = 0; t < 6; ++t) f[t] = ld(2 + bx, 2 + by, t);
#pragma unroll
            for (int bz = 0; bz < 2; ++bz)
                F[2] = fmaf(fd4(f[bz], f[bz + 1], f[bz + 3], f[bz + 4], inv12res), z[bx | (by << 1) | (bz << 2)], F[2]);
        }
    }
}

// FP64 force at one particle: the trilinear interpolation (weights from the
// fractions a1) of the 4-point FD gradient at the 2^R nodes around its origin
// o (src/run/numeric.hh:17-44,86-102), in the reference's operation order:
// corner weights z_j = prod_k A[bit_k(j)][k], F_0 summed over the corners in
// order as its lines are read, F_1 and F_2 from per-corner gradients summed
// in corner order.  ld(tx, ty, tz) returns the potential at node
// o - 2 + (tx, ty, tz) (tz ignored in 2-D): from the mesh (MeshNodes) or
// from a brick's LDS tile, with the same operations either way.
// The derivative's terms RN(f / 12) and RN(2 f / 3) are taken from one
// value per node, t = RN(f / 12) (div3 of f / 4): RN(2 f / 3) = 8 t exactly,
// since scaling by a power of two commutes with rounding while the results
// stay normal (|f| >= 2^-1016, the range div3 is exact in; and zeros).  A
// node's t serves both corners of its line and, for the 2^R nodes around
// the origin, the lines of every axis: 56 thirds per particle in 3-D
// instead of 96, the same value bit for bit.
template <int R, typename Ld>
__device__ __forceinline__ void force_f64(const double* a1, Ld ld, const PosArgs& pa, double* F) {
    constexpr int NC = 1 << R;
    double z[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        double w = 1.0;
#pragma unroll
        for (int k = 0; k < R; ++k) w *= ((j >> k) & 1) ? a1[k] : 1.0 - a1[k];
        z[j] = w;
    }
    auto third = [](double f) { return div3(f * 0.25); };  // RN(f / 12)
    // ((f[-2]/12 - 2 f[-1]/3) + 2 f[+1]/3) - f[+2]/12, then / res
    auto fd = [&](double tm2, double tm1, double tp1, double tp2) {
        const double v = tm2 - 8 * tm1 + 8 * tp1 - tp2;
        return pa.fastdiv ? div_fma(v, pa.res, pa.rinv) : v / pa.res;
    };
    // the nodes o + (bx, by, bz) (line positions 2 and 3 along every axis)
    double c[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) c[j] = third(ld(2 + (j & 1), 2 + ((j >> 1) & 1), 2 + ((j >> 2) & 1)));
    // the line along axis k through core corner j: its six thirds
    auto line = [&](int k, int j, double* t) {
        int p[3] = {2 + (j & 1), 2 + ((j >> 1) & 1), 2 + ((j >> 2) & 1)};
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            if (u == 2 || u == 3) {
                t[u] = c[(j & ~(1 << k)) | ((u - 2) << k)];
            } else {
                p[k] = u;
                t[u] = third(ld(p[0], p[1], p[2]));
            }
        }
    };
#pragma unroll
    for (int k = 0; k < R; ++k) F[k] = 0;
#pragma unroll
    for (int l = 0; l < NC / 2; ++l) {
        const int by = l & 1, bz = (l >> 1) & 1;
        double t[6];
        line(0, (by << 1) | (bz << 2), t);
#pragma unroll
        for (int bx = 0; bx < 2; ++bx) F[0] += fd(t[bx], t[bx + 1], t[bx + 3], t[bx + 4]) * z[bx | (by << 1) | (bz << 2)];
    }
    {
        double g1[NC];
#pragma unroll
        for (int l = 0; l < NC / 2; ++l) {
            const int bx = l & 1, bz = (l >> 1) & 1;
            double t[6];
            line(1, bx | (bz << 2), t);
#pragma unroll
            for (int by = 0; by < 2; ++by) g1[bx | (by << 1) | (bz << 2)] = fd(t[by], t[by + 1], t[by + 3], t[by + 4]);
        }
#pragma unroll
        for (int j = 0; j < NC; ++j) F[1] += g1[j] * z[j];
    }
    if constexpr (R == 3) {
        double g2[NC];
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const int bx = l & 1, by = (l >> 1) & 1;
            double t[6];
            line(2, bx | (by << 1), t);
#pragma unroll
            for (int bz = 0; bz < 2; ++bz) g2[bx | (by << 1) | (bz << 2)] = fd(t[bz], t[bz + 1], t[bz + 3], t[bz + 4]);
        }
#pragma unroll
        for (int j = 0; j < NC; ++j) F[2] += g2[j] * z[j];
    }
}

// The potential at node o - 2 + (tx, ty, tz) of an N^R mesh, wrapped.
template <int R, typename T = float>
struct MeshNodes {
    const T* __restrict__ phi;
    int o0, o1, o2, lgN, msk;
    __device__ __forceinline__ T operator()(int tx, int ty, int tz) const {
        uint32_t idx = (uint32_t)((o0 - 2 + tx) & msk) + ((uint32_t)((o1 - 2 + ty) & msk) << lgN);
        if (R == 3) idx += (uint32_t)((o2 - 2 + tz) & msk) << (2 * lgN);
        return phi[idx];
    }
};

// The per-particle steps of the brick gathers, for both position forms.
// FX: the kick adds F ck rounded to the momentum grid (ckq = ck / pq), the
// drift is integer; FP64: the reference's P - F ck and X + P cd (ckq = ck,
// pq unused).  The sort cell of a position is the deposit's (PP::dep).
template <typename PP>
__device__ __forceinline__ double brick_kick(double p, double F, double ckq, double pq) {
    if constexpr (std::is_same<PP, PosFX>::value) {
        return p - fx_kick_increment(F, ckq, pq);
    } else {
        return p - F * ckq;
    }
}
template <typename PP>
__device__ __forceinline__ typename PP::Pos brick_drift(typename PP::Pos x, double p, double cd) {
    if constexpr (std::is_same<PP, PosFX>::value) {
        return x + fx_round_u32(p * cd);
    } else {
        return x + p * cd;
    }
}
template <typename PP>
__device__ __forceinline__ int brick_cell(typename PP::Pos x, const PosArgs& pa, int msk) {
    if constexpr (std::is_same<PP, PosFX>::value) {
        return (int)(x >> pa.fshift) & msk;
    } else {
        int o;
        double f;
        PosF64::dep(x, pa, o, f);
        return o & msk;
    }
}
// The force of one particle from node values ld (force_fx or force_f64).
template <typename PP, int R, typename Ld>
__device__ __forceinline__ void brick_force(const typename PP::Mesh* a1, Ld ld, const PosArgs& pa,
                                            typename PP::Mesh* F) {
    if constexpr (std::is_same<PP, PosFX>::value) {
        force_fx<R>(a1, ld, (float)(1.0 / (12.0 * pa.res)), F);
    } else {
        force_f64<R>(a1, ld, pa, F);
    }
}

// Kick (and drift, with the new sort key and mover counts) of one particle
// whose force k_gather_kick_brick cannot take from its tile: input index
// i, sorted position s, the force read from the mesh by the same operations
// as in the tile (brick_force), the rest as in the tile kernel.
template <typename PP, int R>
__device__ __forceinline__ void kick_from_mesh(uint32_t s, uint32_t i, const typename PP::Pos* __restrict__ X,
                                               const double* __restrict__ P, typename PP::Pos* __restrict__ Xo,
                                               double* __restrict__ Po, long np,
                                               const typename PP::Mesh* __restrict__ phi, int lgN, const PosArgs& pa,
                                               double ckq, double pq, double cd, int cd_on,
                                               uint32_t* __restrict__ keys, uint32_t* __restrict__ in_count,
                                               uint32_t* __restrict__ out_count) {
    using Pos = typename PP::Pos;
    using T = typename PP::Mesh;
    const int msk = (1 << lgN) - 1;
    Pos xs[3] = {0, 0, 0};
    double pv[3] = {0, 0, 0};
    int o[3] = {0, 0, 0};
    T a1[3] = {0, 0, 0}, F[3];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        xs[k] = X[k * np + i];
        pv[k] = P[k * np + i];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) PP::gat(xs[k], pa, o[k], a1[k]);
    brick_force<PP, R>(a1, MeshNodes<R, T>{phi, o[0], o[1], o[2], lgN, msk}, pa, F);
    Pos xn[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double p = brick_kick<PP>(pv[k], (double)F[k], ckq, pq);
        Po[k * np + s] = p;
        xn[k] = cd_on ? brick_drift<PP>(xs[k], p, cd) : xs[k];
        Xo[k * np + s] = xn[k];
    }
    if (keys) {
        int co[3] = {0, 0, 0}, cn[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < R; ++k) {
            co[k] = brick_cell<PP>(xs[k], pa, msk);
            cn[k] = brick_cell<PP>(xn[k], pa, msk);
        }
        const uint32_t ko = brick_key(co[0], co[1], co[2], R, pa), kn = brick_key(cn[0], cn[1], cn[2], R, pa);
        keys[s] = kn;
        if (kn != ko) {
            atomicAdd(&in_count[kn], 1u);
            atomicAdd(&out_count[ko], 1u);
        }
    }
}

// Gather + kick (+ drift).  A workgroup takes a segment of 2^lseg bricks
// that follow each other along the last axis (same x0 and, in 3-D, y0).  A
// brick reads the potential of its cells and a halo of 2 cells below / 3
// above along each axis (the 4-point stencil at the 2^R corners): a tile of
// TP = B + 5 planes across the last axis (B: the brick's extent along it).
// Consecutive bricks share 5 of those planes, so the planes live in an LDS
// ring of TR = TP + B slots: plane p of the current brick's tile (global
// plane x0 - 2 + p) sits in slot (p + rot) mod TR, and rot advances by B per
// brick.  The B slots the current tile leaves free receive the next brick's
// new planes by direct global->LDS loads issued while the current brick's
// particles are processed, so the loads overlap that work and occupy no
// registers.  A slot holds PS = (cells of a plane) floats, x fastest; planes
// in consecutive slots are contiguous, which the direct load (64 lanes to
// 64 consecutive dwords) needs; a wave whose 64 cells straddle the ring's
// wrap issues its two parts separately.  Ring and guard stay below 64 KiB.
// Particle s (storage perm[s]) is written to position s of the output arrays.
// With keys != nullptr (drift fused), keys[s] is the sort key of the
// particle's new position, and a particle whose key changed is counted into
// its new group (in_count) and out of its old one (out_count): what keys_fx
// computes for an incremental re-sort, whose previous keys are those of the
// positions read here.  FULL: every brick has the full size of the form (the
// mesh is at least a brick along each axis), so the tile strides are
// compile-time constants and the stencil reads take their offsets as
// immediates.
template <typename PP, int R, int NT, bool FULL>
__global__ __launch_bounds__(NT) void k_gather_kick_brick(
    const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ perm, const typename PP::Pos* __restrict__ X,
    const double* __restrict__ P, const uint32_t* __restrict__ idi, typename PP::Pos* __restrict__ Xo,
    double* __restrict__ Po, uint32_t* __restrict__ ido, long np, const typename PP::Mesh* __restrict__ phi, int lgN,
    PosArgs pa, double ckq, double pq, double cd, int cd_on, int lseg, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ in_count, uint32_t* __restrict__ out_count, uint32_t* __restrict__ defer) {
    using Pos = typename PP::Pos;
    using T = typename PP::Mesh;                  // potential (float or double)
    using BD = BrickDims<PP>;
    constexpr int DW = (int)(sizeof(T) / 4);      // dwords per cell
    constexpr int LA = R - 1;                     // the axis the tile slides along
    constexpr int BL = (R == 3) ? BD::Z : BD::X2;   // brick extent along it, at most
    constexpr int PSMAX = (R == 3) ? (BD::X + 5) * (BD::Y + 5) : BD::X2 + 5;
    // The guard keeps the ring at least 256 bytes above LDS address 0, so the
    // base of the second part of a straddling wave (up to 63 cells before the
    // ring) is a valid address.
    constexpr int DQ = 512;  // LDS queue of particles whose force is read from the mesh
    struct Lds {
        float guard[64];
        T ring[(2 * BL + 5) * PSMAX];
        uint32_t dq[DQ];
        uint32_t dqn;
    };
    __shared__ Lds sm;
    static_assert(offsetof(Lds, dq) <= 65536, "the ring must stay addressable below 64 KiB");
    T* const ring = sm.ring;
    const int N = 1 << lgN, msk = N - 1;
    const int tid = threadIdx.x;
    const int lcol = pa.lnb[0] + ((R == 3) ? pa.lnb[1] : 0);
    const uint32_t cstride = 1u << lcol;  // brick-index step along the last axis
    // Workgroup w runs on XCD w mod 8; XCD k takes logical workgroups
    // k G/8 .. (k + 1) G/8 - 1 (columns fastest), so the columns running on
    // one XCD at a time are neighbours and share their halo lines in its L2.
    const uint32_t G = gridDim.x;
    const uint32_t lw = (G % 8 == 0) ? (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3) : blockIdx.x;
    uint32_t bid = (lw & (cstride - 1)) | ((uint32_t)(lw >> lcol) << (lcol + lseg));
    BrickCoord bc = brick_coord(bid, pa);
    constexpr int FB0 = (R == 3) ? BD::X : BD::X2, FB1 = (R == 3) ? BD::Y : BD::X2, FB2 = (R == 3) ? BD::Z : 1;
    const int B0 = FULL ? FB0 : bc.B[0], B1 = FULL ? FB1 : bc.B[1], B2 = FULL ? FB2 : bc.B[2];
    const int B = (R == 3) ? B2 : B1;  // extent along the last axis
    const int TX = B0 + 5, TP = B + 5, TR = TP + B;
    const int PS = (R == 3) ? TX * (B1 + 5) : TX;  // cells of a plane
    // e / PS and r / TX by multiply-high: exact for the e, r < 2^14 used here
    const uint32_t psinv = (uint32_t)((0x100000000ull + PS - 1) / PS);
    const uint32_t txinv = (uint32_t)((0x100000000ull + TX - 1) / TX);
    const uint32_t ring_lds = (uint32_t)(uintptr_t)ring;  // LDS byte address of slot 0
    int rot = 0;
    auto slot = [&](int p) {
        const int q = p + rot;
        return q >= TR ? q - TR : q;
    };
    // (index products in 24-bit multiplies: every factor is below 2^12)
    auto Tn = [&](int tx, int ty, int tz) -> T {
        return (R == 3) ? ring[__umul24(PS, slot(tz)) + tx + __umul24(TX, ty)] : ring[__umul24(PS, slot(ty)) + tx];
    };
    // Planes p0 .. p0 + npl - 1 of the tile of the brick at last-axis origin
    // zb, into slots (p + rp) mod TR.  Cell e of the run (plane e / PS) goes
    // to LDS slot-start * PS + e before the wrap and e - E after it.
    // (In dwords: a cell is DW of them, dword d of the run belongs to cell
    // d / DW, and half d mod DW = lane mod DW of it.)
    auto load_planes = [&](int p0, int npl, int zb, int rp) {
        constexpr int NLMAX = ((BL + 5) * PSMAX * DW + NT - 1) / NT;
        const int nd = npl * PS * DW;
        int start = p0 + rp;
        start = start >= TR ? start - TR : start;
        const int E = (TR - start) * PS * DW;  // first dword whose plane wraps to slot 0
        const int wbase = tid & ~63, lane = tid & 63;
#pragma unroll 1
        for (int l = 0; l < NLMAX; ++l) {
            const int d0 = wbase + l * NT;  // the wave's first dword
            if (d0 >= nd) break;
            const int d = d0 + lane;
            if (d < nd) {
                const int e = d / DW;
                const int q = (int)__umulhi((uint32_t)e, psinv), r = e - q * PS;
                const int ty = (R == 3) ? (int)__umulhi((uint32_t)r, txinv) : 0;
                const int tx = r - ty * TX;
                const uint32_t gl = (uint32_t)((zb - 2 + p0 + q) & msk);
                uint32_t off = (uint32_t)((bc.x0[0] - 2 + tx) & msk);
                off += (R == 3) ? (uint32_t)N * ((uint32_t)((bc.x0[1] - 2 + ty) & msk) + (uint32_t)N * gl)
                                : (uint32_t)N * gl;
                const float* src = reinterpret_cast<const float*>(phi + off) + (d & (DW - 1));
                if (d < E)
                    lds_dma_dword(src, ring_lds + 4u * (uint32_t)(start * PS * DW + d0));
                else
                    lds_dma_dword(src, ring_lds + 4u * (uint32_t)(d0 - E));
            }
        }
    };
    // The next brick's B new planes: cell e = wbase + l NT + lane of the
    // run (plane e / PS) has the same in-plane mesh offset for every brick of
    // the segment, so it is computed once, packed as (plane << 27) | offset
    // (lanes with tid + l NT >= NLIM have no cell; for full bricks NLIM is a
    // constant and the test vanishes from the loads that are always whole).
    constexpr int NLN = (BL * PSMAX * DW + NT - 1) / NT;
    const int wbase = __builtin_amdgcn_readfirstlane(tid & ~63), lane = tid & 63;
    const int lgP = (R == 3) ? 2 * lgN : lgN;  // log2 of the mesh stride between planes
    uint32_t noff[NLN];
#pragma unroll
    for (int l = 0; l < NLN; ++l) {
        const int d = wbase + l * NT + lane, e = d / DW;
        const int q = (int)__umulhi((uint32_t)e, psinv), r = e - q * PS;
        const int ty = (R == 3) ? (int)__umulhi((uint32_t)r, txinv) : 0;
        const int tx = r - ty * TX;
        uint32_t xy = (uint32_t)((bc.x0[0] - 2 + tx) & msk);
        if (R == 3) xy += (uint32_t)N * (uint32_t)((bc.x0[1] - 2 + ty) & msk);
        noff[l] = ((uint32_t)q << 27) | xy;
    }
    const int NLIM = B * PS * DW;  // dwords of the B new planes
    // the next brick's new planes (last-axis origin zb) into slots
    // (p + rp) mod TR, p = 5 .. TP - 1
    auto load_new_planes = [&](int zb, int rp) {
        int start = 5 + rp;
        start = start >= TR ? start - TR : start;
        const int E = (TR - start) * PS * DW;  // first dword whose plane wraps to slot 0
        // the wave's LDS bases before and after the wrap (load l adds l NT dwords)
        const uint32_t ba = ring_lds + 4u * (uint32_t)(start * PS * DW + wbase);
        const uint32_t bb = ring_lds + 4u * (uint32_t)(wbase - E);
#pragma unroll
        for (int l = 0; l < NLN; ++l) {
            if (tid + l * NT < NLIM) {
                const uint32_t v = noff[l];
                const uint32_t gl = (uint32_t)((zb + 3 + (int)(v >> 27)) & msk);
                const uint32_t off = (v & 0x7FFFFFFu) + (gl << lgP);
                const float* src = reinterpret_cast<const float*>(phi + off) + (lane & (DW - 1));
                if (tid + l * NT < E)
                    lds_dma_dword(src, ba + 4u * (uint32_t)(l * NT));
                else
                    lds_dma_dword(src, bb + 4u * (uint32_t)(l * NT));
            }
        }
    };
    // Particles are processed in rounds of NT (round r of a brick: sorted
    // position s0 + r NT + tid; an empty brick still takes one round, which
    // advances the ring).  The loads are pipelined two rounds deep: a round
    // first issues the loads of the thread's particle of the next round (its
    // storage index came with the round before) and the storage index of the
    // particle of the round after that, then processes its own particle, so
    // the loads overlap the round's work.  The two particle register sets are
    // used in turn (a copy of a loaded set would wait for its loads at once),
    // and every load is unconditional (invalid slots read the last particle
    // again), so no value is merged from two paths.  The next brick's planes
    // are issued at the start of the brick's last round, after a wait for
    // everything loaded so far: no load the compiler waits for is then older
    // than them, so they stay in flight through the last round instead of
    // being drained by the first wait.
    const int nseg = 1 << lseg;
    const uint32_t plast = (uint32_t)(np - 1);
    // sorted ranges of the current brick and the next two along the segment
    // (past the segment: empty)
    const uint32_t bid0 = bid;  // the segment's first brick
    uint32_t c0 = bstart[bid << R], c1 = bstart[(bid + 1) << R];
    uint32_t n0 = c1, n1 = c1, m0 = c1, m1 = c1;
    if (nseg > 1) {
        n0 = bstart[(bid + cstride) << R];
        n1 = bstart[(bid + cstride + 1) << R];
    }
    if (nseg > 2) {
        m0 = bstart[(bid + 2 * cstride) << R];
        m1 = bstart[(bid + 2 * cstride + 1) << R];
    }
    auto rounds = [](uint32_t a, uint32_t b) -> int {
        const int n = (int)((b - a + NT - 1) / NT);
        return n > 0 ? n : 1;
    };
    struct PSet {
        Pos x[3];
        double p[3];
        uint32_t id;
    };
    auto fetch = [&](PSet& d, uint32_t i) {
        d.id = idi[i];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            d.x[k] = X[k * np + i];
            d.p[k] = P[k * np + i];
        }
    };
    int j = 0, r = 0, nrr = rounds(c0, c1);
    // slot of the thread in round r + dr (dr = 1, 2) of the pipeline, counted
    // over the current brick and the next two; ~0u: no such particle
    auto slot_ahead = [&](int dr) -> uint32_t {
        int rr = r + dr;
        if (rr < nrr) {
            const uint32_t s = c0 + (uint32_t)rr * NT + tid;
            return s < c1 ? s : ~0u;
        }
        rr -= nrr;
        if (j + 1 >= nseg) return ~0u;
        const int nr1 = rounds(n0, n1);
        if (rr < nr1) {
            const uint32_t s = n0 + (uint32_t)rr * NT + tid;
            return s < n1 ? s : ~0u;
        }
        rr -= nr1;
        if (j + 2 >= nseg) return ~0u;
        const uint32_t s = m0 + (uint32_t)rr * NT + tid;  // rr = 0 here (dr <= 2)
        return s < m1 ? s : ~0u;
    };
    PSet pa0, pa1;
    uint32_t q0, q1;  // storage index of the particle of the next round
    {
        const uint32_t s = c0 + tid;
        fetch(pa0, perm[s < c1 ? s : plast]);
        const uint32_t s1 = slot_ahead(1);
        q0 = perm[s1 != ~0u ? s1 : plast];
    }
    if (tid == 0) sm.dqn = 0u;
    load_planes(0, TP, bc.x0[LA], 0);
    __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
    __syncthreads();
    auto process = [&](uint32_t s, const PSet& d) {
        ido[s] = d.id;
        int o[3] = {0, 0, 0};
        T a1[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < R; ++k) PP::gat(d.x[k], pa, o[k], a1[k]);
        // tile coordinates of the stencil origin (cell o - 2 along each axis)
        const int lx = (o[0] - bc.x0[0]) & msk, ly = (o[1] - bc.x0[1]) & msk;
        const int lz = (R == 3) ? ((o[2] - bc.x0[2]) & msk) : 0;
        // A particle whose origin is the cell above its own (a negative
        // unwrapped coordinate, see PosFX::gat; FP64: also a rounding of the
        // unwrapped coordinate that differs from the folded one's) on the
        // brick's upper face needs one plane beyond the tile: it is queued,
        // and kicked from the mesh when the segment is done (kick_from_mesh).
        if (lx >= B0 || ly >= B1 || (R == 3 && lz >= B2)) {
            const uint32_t q = atomicAdd(&sm.dqn, 1u);
            if (q < (uint32_t)DQ)
                sm.dq[q] = s;
            else
                defer[s] = 1u;  // beyond the LDS queue: flagged at its sorted position
            return;
        }
        T F[3];
        brick_force<PP, R>(a1, [&](int tx, int ty, int tz) { return Tn(lx + tx, ly + ty, lz + tz); }, pa, F);
        Pos xn[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const double p = brick_kick<PP>(d.p[k], (double)F[k], ckq, pq);
            Po[k * np + s] = p;
            xn[k] = cd_on ? brick_drift<PP>(d.x[k], p, cd) : d.x[k];
            Xo[k * np + s] = xn[k];
        }
        if (keys) {
            int co[3] = {0, 0, 0}, cn[3] = {0, 0, 0};
#pragma unroll
            for (int k = 0; k < R; ++k) {
                co[k] = brick_cell<PP>(d.x[k], pa, msk);
                cn[k] = brick_cell<PP>(xn[k], pa, msk);
            }
            const uint32_t ko = brick_key(co[0], co[1], co[2], R, pa), kn = brick_key(cn[0], cn[1], cn[2], R, pa);
            keys[s] = kn;
            if (kn != ko) {
                atomicAdd(&in_count[kn], 1u);
                atomicAdd(&out_count[ko], 1u);
            }
        }
    };
    // One round with the particle in cur (its storage index came with the
    // round before); qc: storage index of the next round's particle.  Returns
    // false after the segment's last round.
    auto round = [&](const PSet& cur, PSet& nxt, uint32_t qc, uint32_t& qn) -> bool {
        const bool last = r + 1 == nrr;
        const bool more = j + 1 < nseg;
        if (last) {
            __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
            if (more) {
                int rn = rot + B;
                rn = rn >= TR ? rn - TR : rn;
                load_new_planes(bc.x0[LA] + B, rn);
            }
        }
        fetch(nxt, qc);
        const uint32_t s2 = slot_ahead(2);
        qn = perm[s2 != ~0u ? s2 : plast];
        const uint32_t s = c0 + (uint32_t)r * NT + tid;
        if (s < c1) process(s, cur);
        if (!last) {
            ++r;
            return true;
        }
        if (!more) return false;
        // the next brick: its planes have landed once every wave has waited
        // for its own direct loads and passed the barrier
        c0 = n0;
        c1 = n1;
        n0 = m0;
        n1 = m1;
        if (j + 3 < nseg) {
            m0 = bstart[(bid + 3 * cstride) << R];
            m1 = bstart[(bid + 3 * cstride + 1) << R];
        } else {
            m0 = m1 = n1;
        }
        ++j;
        r = 0;
        nrr = rounds(c0, c1);
        __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);  // this wave's direct loads have landed
        __syncthreads();
        rot += B;
        rot = rot >= TR ? rot - TR : rot;
        bid += cstride;
        bc.x0[LA] += B;
        return true;
    };
    for (;;) {
        if (!round(pa0, pa1, q0, q1)) break;
        if (!round(pa1, pa0, q1, q0)) break;
    }
    // the queued particles of the segment, from the mesh
    __syncthreads();
    const uint32_t nq = sm.dqn < (uint32_t)DQ ? sm.dqn : (uint32_t)DQ;
    for (uint32_t q = tid; q < nq; q += NT) {
        const uint32_t s = sm.dq[q];
        kick_from_mesh<PP, R>(s, perm[s], X, P, Xo, Po, np, phi, lgN, pa, ckq, pq, cd, cd_on, keys, in_count,
                              out_count);
    }
    // those beyond the LDS queue (rare): the segment's bricks are walked for
    // their flags, each cleared after its particle is kicked, so the flags
    // are all zero again when the kernel ends
    if (sm.dqn > (uint32_t)DQ) {  // (uniform)
        for (int jj = 0; jj < nseg; ++jj) {
            const uint32_t b = bid0 + (uint32_t)jj * cstride;
            const uint32_t r0 = bstart[b << R], r1 = bstart[(b + 1) << R];
            for (uint32_t s = r0 + tid; s < r1; s += NT) {
                if (defer[s]) {
                    defer[s] = 0u;
                    kick_from_mesh<PP, R>(s, perm[s], X, P, Xo, Po, np, phi, lgN, pa, ckq, pq, cd, cd_on, keys,
                                          in_count, out_count);
                }
            }
        }
    }
}

// ------------------------------------------------------------ FP64 brick order
static constexpr int F64B_NT = 256;  // FP64 brick deposit workgroup size
static constexpr int F64G_NT = 256;  // FP64 brick gather workgroup size
static constexpr int F64B_CELLS = F64B_X * F64B_Y * F64B_Z;
static_assert(F64B_X2 * F64B_X2 == F64B_CELLS, "2-D and 3-D FP64 bricks hold the same number of cells");

// FP64 brick keys (keys_fx for FP64 positions: the brick of the deposit's cell).
template <int R>
__global__ __launch_bounds__(256) void k_keys_f64_brick(const double* __restrict__ X, long np, int lgN, PosArgs pa,
                                                        const uint32_t* __restrict__ prevkeys,
                                                        uint32_t* __restrict__ keys, uint32_t* __restrict__ pids,
                                                        uint32_t* __restrict__ in_count,
                                                        uint32_t* __restrict__ out_count) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    const int mask = (1 << lgN) - 1;
    int oc[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < R; ++k) oc[k] = brick_cell<PosF64>(X[k * np + s], pa, mask);
    const uint32_t key = brick_key(oc[0], oc[1], oc[2], R, pa);
    keys[s] = key;
    fx_note_key(key, s, prevkeys, pids, in_count, out_count);
}

// FP64 deposit in exact fixed point, one workgroup per brick: the particles
// of the brick and the face classes of its lower neighbours (brick_ranges),
// each adding the weights of its 2^R corners that land in the brick: the
// reference's z = ((m A_x) A_y) A_z (numeric.hh:124-131), split by
// fx128_split into the cell's two accumulator words.
template <int R>
__global__ __launch_bounds__(F64B_NT) void k_deposit_f64_brick(const uint32_t* __restrict__ bstart,
                                                               const uint32_t* __restrict__ perm,
                                                               const double* __restrict__ X, long np, int lgN,
                                                               PosArgs pa, double mass, double* __restrict__ mesh,
                                                               double* __restrict__ partial) {
    __shared__ unsigned long long alo[F64B_CELLS], ahi[F64B_CELLS];
    __shared__ double red[F64B_NT / 64];
    __shared__ BrickRanges rg;
    const int N = 1 << lgN, msk = N - 1;
    const uint32_t bid = blockIdx.x;
    const BrickCoord bc = brick_coord(bid, pa);
    const int ncell = bc.B[0] * bc.B[1] * bc.B[2];
    const int tid = threadIdx.x;
    for (int e = tid; e < ncell; e += F64B_NT) alo[e] = ahi[e] = 0ull;
    brick_ranges<R>(rg, bstart, bc, pa, tid);
    const uint32_t total = rg.roff[rg.n];
    // two particles per thread at a time: their perm and position loads are
    // issued before the first is used
    constexpr int BATCH = 2;
    for (uint32_t t0 = tid; t0 < total; t0 += BATCH * F64B_NT) {
        uint32_t pi[BATCH];
#pragma unroll
        for (int q = 0; q < BATCH; ++q) {
            const uint32_t t = t0 + q * F64B_NT;
            pi[q] = t < total ? perm[brick_range_pos(rg, t)] : 0u;
        }
        double xq[BATCH][3];
#pragma unroll
        for (int q = 0; q < BATCH; ++q)
#pragma unroll
            for (int k = 0; k < 3; ++k) xq[q][k] = (k < R && t0 + q * F64B_NT < total) ? X[k * np + pi[q]] : 0.0;
#pragma unroll
        for (int q = 0; q < BATCH; ++q) {
            if (t0 + q * F64B_NT >= total) break;
            int o[3] = {0, 0, 0};
            double f[3] = {0, 0, 0};
#pragma unroll
            for (int k = 0; k < R; ++k) {
                PosF64::dep(xq[q][k], pa, o[k], f[k]);
                o[k] &= msk;
            }
#pragma unroll
            for (int j = 0; j < (1 << R); ++j) {
                int l[3];
                bool in = true;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    l[k] = (k < R) ? (((o[k] + ((j >> k) & 1)) - bc.x0[k]) & msk) : 0;
                    in &= l[k] < bc.B[k];
                }
                if (!in) continue;
                double z = mass;
                z *= (j & 1) ? f[0] : 1 - f[0];
                z *= (j & 2) ? f[1] : 1 - f[1];
                if (R == 3) z *= (j & 4) ? f[2] : 1 - f[2];
                unsigned long long hi, lo;
                fx128_split(z, hi, lo);
                const int c = l[0] + bc.B[0] * (l[1] + bc.B[1] * l[2]);
                const unsigned long long old = atomicAdd(&alo[c], lo);
                atomicAdd(&ahi[c], hi + (old + lo < old ? 1ull : 0ull));
            }
        }
    }
    __syncthreads();
    double rsum = 0;
    for (int e = tid; e < ncell; e += F64B_NT) {
        const int lx = e & (bc.B[0] - 1), ly = (e >> pa.lb[0]) & (bc.B[1] - 1), lz = e >> (pa.lb[0] + pa.lb[1]);
        const long idx = (long)(bc.x0[0] + lx) + (long)N * ((bc.x0[1] + ly) + (long)N * (bc.x0[2] + lz));
        const double v = fx128_value(ahi[e], alo[e]);
        mesh[idx] = v;
        rsum += v;
    }
    if (partial) {  // (uniform)
        const double tot = block_sum<F64B_NT>(rsum, red, tid);
        if (tid == 0) partial[bid] = tot;
    }
}

// ------------------------------------------------------------ gather + kick (+ drift)
// P -= grad(phi)(X) * ck with the trilinear interpolation of the 4-point FD
// gradient at the 2^R surrounding nodes (src/run/nbody.hh:112-117,
// src/run/numeric.hh:17-44,86-102); then, when cd_on, the next drift
// X += P * cd (src/run/nbody.hh:126-130) on the updated momentum.
template <typename T>
__device__ __forceinline__ T load_mesh(const T* __restrict__ phi, long idx) {
    return phi[idx];
}

// The force at one particle: the trilinear interpolation of the 4-point FD
// gradient at the 2^R surrounding nodes (src/run/numeric.hh:17-44,86-102),
// in the reference's operation order.
template <typename PP, int R>
__device__ __forceinline__ void gather_force(const typename PP::Pos* xs, const typename PP::Mesh* __restrict__ phi,
                                             int lgN, const PosArgs& pa, typename PP::Mesh* F) {
    const int N = 1 << lgN, msk = N - 1;
    if constexpr (std::is_same<PP, PosFX>::value) {
        // the brick kernel's operations (force_fx), from the mesh
        int of[3] = {0, 0, 0};
        float af[3] = {0.f, 0.f, 0.f}, Ff[3];
#pragma unroll
        for (int k = 0; k < R; ++k) PosFX::gat(xs[k], pa, of[k], af[k]);
        force_fx<R>(af, MeshNodes<R>{phi, of[0], of[1], of[2], lgN, msk}, (float)(1.0 / (12.0 * pa.res)), Ff);
#pragma unroll
        for (int k = 0; k < R; ++k) F[k] = Ff[k];
    } else {
        int o[3] = {0, 0, 0};
        double a1[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < R; ++k) PosF64::gat(xs[k], pa, o[k], a1[k]);
        force_f64<R>(a1, MeshNodes<R, double>{phi, o[0], o[1], o[2], lgN, msk}, pa, F);
    }
}


// With ro.perm set, thread s reads particle perm[s] and writes the updated
// particle (and its id) to position s of the output arrays: the kick also
// moves the particles into the cell order of the deposit.  With ro.in_place
// as well, thread s updates particle perm[s] where it is: the threads of a
// wave then handle particles of neighbouring cells, whose potential loads
// coincide, while the particles stay in their order.
template <typename PP, int R>
__global__ __launch_bounds__(256) void k_gather_kick(typename PP::Pos* __restrict__ X, double* __restrict__ P,
                                                      long np, const typename PP::Mesh* __restrict__ phi,
                                                      int lgN, PosArgs pa, double ck, double cd, int cd_on,
                                                      double* __restrict__ Fout, Reorder<PP> ro) {
    using M = typename PP::Mesh;
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    const bool move = ro.perm && !ro.in_place;
    const long i = ro.perm ? (long)ro.perm[s] : s;  // input index
    const long w = move ? s : i;                      // output index
    typename PP::Pos* __restrict__ Xo = move ? ro.Xo : X;
    double* __restrict__ Po = move ? ro.Po : P;
    if (move) ro.ido[s] = ro.idi[i];
    typename PP::Pos xs[R];
#pragma unroll
    for (int k = 0; k < R; ++k) xs[k] = X[k * np + i];
    M F[R];
    gather_force<PP, R>(xs, phi, lgN, pa, F);
    if (Fout) {
#pragma unroll
        for (int k = 0; k < R; ++k) Fout[k * np + i] = (double)F[k];
        return;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double p = P[k * np + i] - (double)F[k] * ck;
        Po[k * np + w] = p;
        typename PP::Pos xn = xs[k];
        if (cd_on) {
            if constexpr (sizeof(M) == 8) {
                xn = xs[k] + p * cd;
            } else {
                xn = xs[k] + fx_round_u32(p * cd);
            }
        }
        if (cd_on || move) Xo[k * np + w] = xn;
    }
}


// The FP64 kick in two passes.  k_gather_force_sorted: the force at every
// particle in the deposit's cell order (positions copied into that order by
// k_dep_frac), so the threads of a wave read the same potential lines and
// their positions contiguously; k_kick_drift_inv: the kick (+ drift) in the
// particles' own order, reading each force through the inverse order.  The
// arithmetic is k_gather_kick's, operation for operation.
template <typename PP, int R>
__global__ __launch_bounds__(256) void k_gather_force_sorted(const typename PP::Pos* __restrict__ Xs, long np,
                                                              const typename PP::Mesh* __restrict__ phi, int lgN,
                                                              PosArgs pa, double* __restrict__ Fs) {
    using M = typename PP::Mesh;
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    typename PP::Pos xs[R];
#pragma unroll
    for (int k = 0; k < R; ++k) xs[k] = Xs[k * np + s];
    M F[R];
    gather_force<PP, R>(xs, phi, lgN, pa, F);
#pragma unroll
    for (int k = 0; k < R; ++k) Fs[s * R + k] = (double)F[k];
}

template <typename PP, int R>
__global__ __launch_bounds__(256) void k_kick_drift_inv(typename PP::Pos* __restrict__ X, double* __restrict__ P,
                                                         long np, const double* __restrict__ Fs,
                                                         const uint32_t* __restrict__ inv, double ck, double cd,
                                                         int cd_on) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const long s = inv[i];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double p = P[k * np + i] - Fs[s * R + k] * ck;
        P[k * np + i] = p;
        if (cd_on) X[k * np + i] = X[k * np + i] + p * cd;
    }
}

// X += P * cd (src/run/nbody.hh:122-133); FX: cd is in units of L / 2^30.
template <typename PP>
__global__ __launch_bounds__(256) void k_drift(typename PP::Pos* __restrict__ X, const double* __restrict__ P,
                                                long n, double cd) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (sizeof(typename PP::Mesh) == 8) {
        X[i] = X[i] + P[i] * cd;
    } else {
        X[i] = X[i] + fx_round_u32(P[i] * cd);
    }
}

// ------------------------------------------------------------ Zel'dovich start
// u = FD gradient of phi on the mass grid (src/run/numeric.hh:48-70);
// X = q * (L/N) - u * a_pos, P = u * -a_mom (src/run/nbody.hh:40-50).
// FX stores x / L as a 32-bit fraction of the box.
template <typename PP, int R>
__global__ __launch_bounds__(256) void k_zeldovich(const double* __restrict__ phi, int lgN, double L,
                                                    double a_pos, double a_mom,
                                                    typename PP::Pos* __restrict__ X, double* __restrict__ P) {
    const int N = 1 << lgN, msk = N - 1;
    long np = 1;
#pragma unroll
    for (int k = 0; k < R; ++k) np <<= lgN;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double res = L / N;
    int q[R];
#pragma unroll
    for (int k = 0; k < R; ++k) q[k] = (int)((i >> (k * lgN)) & msk);
    long stride[3] = {1, (long)N, (long)N * N};
#pragma unroll
    for (int k = 0; k < R; ++k) {
        long base = i - (long)q[k] * stride[k];
        double f[4];
        const int off[4] = {-2, -1, 1, 2};
#pragma unroll
        for (int s = 0; s < 4; ++s) f[s] = phi[base + (long)((q[k] + off[s]) & msk) * stride[k]];
        const double u = fd4(f[0], f[1], f[2], f[3], res);
        const double G = q[k] * (L / N);
        const double x = G - u * a_pos;
        P[k * np + i] = u * -a_mom;
        if constexpr (sizeof(typename PP::Mesh) == 8) {
            X[k * np + i] = x;
        } else {
            X[k * np + i] = fx_word(x / L);
        }
    }
}

// ------------------------------------------------------------ FP64 reductions
static constexpr int RED_NT = 256;

template <typename T>
__global__ __launch_bounds__(RED_NT) void k_partial_sum(const T* __restrict__ a, long n, long chunk,
                                                         int square, double* __restrict__ partials) {
    __shared__ double red[RED_NT];
    const long b0 = (long)blockIdx.x * chunk;
    const long b1 = b0 + chunk < n ? b0 + chunk : n;
    double s = 0;
    for (long i = b0 + threadIdx.x; i < b1; i += RED_NT) {
        const double v = (double)a[i];
        s += square ? v * v : v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = RED_NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(RED_NT) void k_final_sum(const double* __restrict__ partials, long n,
                                                      double scale, double* __restrict__ out) {
    __shared__ double red[RED_NT];
    double s = 0;
    for (long i = threadIdx.x; i < n; i += RED_NT) s += partials[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = RED_NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] * scale;
}

__global__ __launch_bounds__(RED_NT) void k_partial_max_abs(const double* __restrict__ a, long n, long chunk,
                                                            double* __restrict__ partials) {
    __shared__ double red[RED_NT];
    const long b0 = (long)blockIdx.x * chunk;
    const long b1 = b0 + chunk < n ? b0 + chunk : n;
    double m = 0;
    for (long i = b0 + threadIdx.x; i < b1; i += RED_NT) m = fmax(m, fabs(a[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = RED_NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(RED_NT) void k_final_max(const double* __restrict__ partials, long n,
                                                      double* __restrict__ out) {
    __shared__ double red[RED_NT];
    double m = 0;
    for (long i = threadIdx.x; i < n; i += RED_NT) m = fmax(m, partials[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = RED_NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

__global__ __launch_bounds__(256) void k_quantize(double* __restrict__ a, long n, double q) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = rint(a[i] * (1.0 / q)) * q;
}

// FX positions -> unwrapped doubles in [-2L, 2L)
__global__ __launch_bounds__(256) void k_fx_to_f64(const uint32_t* __restrict__ q, long n, double L,
                                                   double* __restrict__ x) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = fx_coord(q[i], L);
}
__global__ __launch_bounds__(256) void k_f64_to_fx(const double* __restrict__ x, long n, double L,
                                                   uint32_t* __restrict__ q) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = fx_word(x[i] / L);
}

// ------------------------------------------------------------ conan cic -s
// Sub-sampled Lagrangian grid of cic_run<R> (src/cic/cic.hh:64-77): the
// displacement D = X - q L/Nm on the mass grid is interpolated at the points
// y of an (s Nm)^R grid (LinearInterpolation, numeric.hh:17-44; y >= 0 so no
// truncation issue), giving positions y + D(y).
template <int R>
__global__ __launch_bounds__(256) void k_subsample(const double* __restrict__ X, long npm, int Nm, int Ns,
                                                    double L, double* __restrict__ Y, long nps) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nps) return;
    const double res_m = L / Nm;
    int qs[R];
    long t = i;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        qs[k] = (int)(t % Ns);
        t /= Ns;
    }
    double y[R], a1[R];
    int o[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        y[k] = qs[k] * (L / Ns);
        const double x = y[k] / res_m;
        o[k] = (int)(x);
        a1[k] = x - o[k];
    }
    double v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = 0;
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        double z = 1;
        long idx = 0, stride = 1;
        int q[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int b = (j >> k) & 1;
            z *= b ? a1[k] : 1 - a1[k];
            q[k] = (o[k] + b) % Nm;
            idx += (long)q[k] * stride;
            stride *= Nm;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const double d = X[k * npm + idx] - q[k] * (L / Nm);
            v[k] += d * z;
        }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) Y[k * nps + i] = y[k] + v[k];
}

void subsample(int R, const double* X, int Nm, int s, double L, double* Y, hipStream_t st) {
    long npm = 1, nps = 1;
    for (int k = 0; k < R; ++k) {
        npm *= Nm;
        nps *= (long)Nm * s;
    }
    const unsigned nb = (unsigned)((nps + 255) / 256);
    if (R == 3)
        hipLaunchKernelGGL(k_subsample<3>, dim3(nb), dim3(256), 0, st, X, npm, Nm, Nm * s, L, Y, nps);
    else if (R == 2)
        hipLaunchKernelGGL(k_subsample<2>, dim3(nb), dim3(256), 0, st, X, npm, Nm, Nm * s, L, Y, nps);
    else
        throw Error(-2, "only 2-D and 3-D meshes are supported");
    CPM_LAUNCH_CHECK();
}

// ------------------------------------------------------------ launchers

template <typename PP>
void ParticleOps<PP>::keys(int R, const Pos* X, long np, int lgN, const PosArgs& pa, uint32_t* keys,
                           uint32_t* pids, hipStream_t st) {
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_keys<PP, RR>), dim3(nblk(np, 256)), dim3(256), 0, st, X, np,
                                         lgN, pa, keys, pids));
    CPM_LAUNCH_CHECK();
}

void keys_fx(int R, const uint32_t* X, long np, int lgN, const PosArgs& pa, const uint32_t* prevkeys, uint32_t* keys,
             uint32_t* pids, uint32_t* in_count, uint32_t* out_count, hipStream_t st) {
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_keys_fx<RR>), dim3(nblk(np, 256)), dim3(256), 0, st, X, np, lgN, pa,
                                         prevkeys, keys, pids, in_count, out_count));
    CPM_LAUNCH_CHECK();
}

void row_bounds(const uint32_t* keys, long np, int lgN, long nrows, uint32_t* row_start, hipStream_t st) {
    hipLaunchKernelGGL(k_row_bounds, dim3(nblk(np + 1, 256)), dim3(256), 0, st, keys, np, lgN, nrows,
                       row_start);
    CPM_LAUNCH_CHECK();
}

template <typename PP>
void ParticleOps<PP>::deposit(int R, const uint32_t* row_start, const uint32_t* keys, const uint32_t* pids,
                              const Pos* X, long np, int lgN, const PosArgs& pa, Mesh mass, Mesh* mesh,
                              double* row_partial, long nrows, Mesh* frac, Pos* Xs, uint32_t* inv,
                              hipStream_t st) {
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_dep_frac<PP, RR>), dim3(nblk(np, 256)), dim3(256), 0, st, pids, X, np,
                                         pa, frac, Xs, inv));
    CPM_LAUNCH_CHECK();
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_deposit<PP, RR>), dim3((unsigned)nrows), dim3(DEP_NT), 0, st,
                                         row_start, keys, pids, (const Mesh*)frac, np, lgN, mass, mesh,
                                         row_partial));
    CPM_LAUNCH_CHECK();
}

template <typename PP>
void ParticleOps<PP>::gather_kick_sorted(int R, Pos* X, double* P, long np, const Mesh* phi, int lgN,
                                         const PosArgs& pa, double ck, double cd, int cd_on, const Pos* Xs,
                                         double* Fs, const uint32_t* inv, hipStream_t st) {
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_gather_force_sorted<PP, RR>), dim3(nblk(np, 256)), dim3(256), 0, st, Xs,
                                         np, phi, lgN, pa, Fs));
    CPM_LAUNCH_CHECK();
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_kick_drift_inv<PP, RR>), dim3(nblk(np, 256)), dim3(256), 0, st, X, P, np,
                                         (const double*)Fs, inv, ck, cd, cd_on));
    CPM_LAUNCH_CHECK();
}

template <typename PP>
void ParticleOps<PP>::gather_kick(int R, Pos* X, double* P, long np, const Mesh* phi, int lgN,
                                  const PosArgs& pa, double ck, double cd, int cd_on, double* Fout,
                                  const Reorder<PP>& ro, hipStream_t st) {
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_gather_kick<PP, RR>), dim3(nblk(np, 256)), dim3(256), 0, st, X,
                                         P, np, phi, lgN, pa, ck, cd, cd_on, Fout, ro));
    CPM_LAUNCH_CHECK();
}

void brick_layout(int R, int lgN, PosArgs& pa, bool f64) {
    const int bx = f64 ? (R == 3 ? F64B_X : F64B_X2) : FXB_X;
    const int by = f64 ? (R == 3 ? F64B_Y : F64B_X2) : (R == 3 ? FXB_Y : FXB_X);
    const int bz = f64 ? F64B_Z : FXB_Z;
    const int want[3] = {ilog2c(bx), ilog2c(by), R == 3 ? ilog2c(bz) : 0};
    for (int k = 0; k < 3; ++k) {
        const int lg = (k < R) ? lgN : 0;
        pa.lb[k] = want[k] < lg ? want[k] : lg;
        pa.lnb[k] = lg - pa.lb[k];
    }
    pa.brick = 1;
}

long brick_count(const PosArgs& pa) { return 1L << (pa.lnb[0] + pa.lnb[1] + pa.lnb[2]); }

template <typename MT>
void deposit_fx_brick(int R, const uint32_t* bstart, const uint32_t* perm, const uint32_t* X, long np, int lgN,
                      const PosArgs& pa, float mass, MT* mesh, double* partial, hipStream_t st) {
    const unsigned nb = (unsigned)brick_count(pa);
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_deposit_fx_brick<RR, MT>), dim3(nb), dim3(FXB_NT), 0, st, bstart, perm,
                                         X, np, lgN, pa, mass, mesh, partial));
    CPM_LAUNCH_CHECK();
}
template void deposit_fx_brick<float>(int, const uint32_t*, const uint32_t*, const uint32_t*, long, int,
                                      const PosArgs&, float, float*, double*, hipStream_t);
template void deposit_fx_brick<double>(int, const uint32_t*, const uint32_t*, const uint32_t*, long, int,
                                       const PosArgs&, float, double*, double*, hipStream_t);

template <typename PP, int NT, int LSEG>
static void launch_gather_brick(int R, const uint32_t* bstart, const uint32_t* perm, const typename PP::Pos* X,
                                const double* P, const uint32_t* idi, typename PP::Pos* Xo, double* Po, uint32_t* ido,
                                long np, const typename PP::Mesh* phi, int lgN, const PosArgs& pa, double ckq,
                                double pq, double cd, int cd_on, uint32_t* keys, uint32_t* in_count,
                                uint32_t* out_count, uint32_t* defer, hipStream_t st) {
    // segments of up to 2^LSEG bricks along the last axis, as long as that
    // leaves at least 4 workgroups per compute unit
    const long nbr = brick_count(pa), minwg = 4 * device_cus();
    int lseg = 0;
    while (lseg < LSEG && lseg < pa.lnb[R - 1] && (nbr >> (lseg + 1)) >= minwg) ++lseg;
    const unsigned nb = (unsigned)(nbr >> lseg);
    // every brick at the full size of the form: compile-time tile strides
    using BD = BrickDims<PP>;
    const bool full = (R == 3) ? ((1 << pa.lb[0]) == BD::X && (1 << pa.lb[1]) == BD::Y && (1 << pa.lb[2]) == BD::Z)
                               : ((1 << pa.lb[0]) == BD::X2 && (1 << pa.lb[1]) == BD::X2);
#define CPM_GATHER_LAUNCH(F)                                                                                        \
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_gather_kick_brick<PP, RR, NT, F>), dim3(nb), dim3(NT), 0, st, bstart,    \
                                         perm, X, P, idi, Xo, Po, ido, np, phi, lgN, pa, ckq, pq, cd, cd_on, lseg,   \
                                         keys, in_count, out_count, defer))
    if (full)
        CPM_GATHER_LAUNCH(true);
    else
        CPM_GATHER_LAUNCH(false);
#undef CPM_GATHER_LAUNCH
    CPM_LAUNCH_CHECK();
}

void gather_kick_fx_brick(int R, const uint32_t* bstart, const uint32_t* perm, const uint32_t* X, const double* P,
                          const uint32_t* idi, uint32_t* Xo, double* Po, uint32_t* ido, long np, const float* phi,
                          int lgN, const PosArgs& pa, double ck, double pq, double cd, int cd_on, uint32_t* keys,
                          uint32_t* in_count, uint32_t* out_count, uint32_t* defer, hipStream_t st) {
    launch_gather_brick<PosFX, FXG_NT, CPM_GATHER_LSEG>(R, bstart, perm, X, P, idi, Xo, Po, ido, np, phi, lgN, pa, ck * (1.0 / pq), pq,
                                       cd, cd_on, keys, in_count, out_count, defer, st);
}

void gather_kick_f64_brick(int R, const uint32_t* bstart, const uint32_t* perm, const double* X, const double* P,
                           const uint32_t* idi, double* Xo, double* Po, uint32_t* ido, long np, const double* phi,
                           int lgN, const PosArgs& pa, double ck, double cd, int cd_on, uint32_t* keys,
                           uint32_t* in_count, uint32_t* out_count, uint32_t* defer, hipStream_t st) {
    launch_gather_brick<PosF64, F64G_NT, CPM_GATHER_LSEG64>(R, bstart, perm, X, P, idi, Xo, Po, ido, np, phi, lgN, pa, ck, 0.0, cd,
                                         cd_on, keys, in_count, out_count, defer, st);
}

void keys_f64_brick(int R, const double* X, long np, int lgN, const PosArgs& pa, const uint32_t* prevkeys,
                    uint32_t* keys, uint32_t* pids, uint32_t* in_count, uint32_t* out_count, hipStream_t st) {
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_keys_f64_brick<RR>), dim3(nblk(np, 256)), dim3(256), 0, st, X, np, lgN,
                                         pa, prevkeys, keys, pids, in_count, out_count));
    CPM_LAUNCH_CHECK();
}

void deposit_f64_brick(int R, const uint32_t* bstart, const uint32_t* perm, const double* X, long np, int lgN,
                       const PosArgs& pa, double mass, double* mesh, double* partial, hipStream_t st) {
    const unsigned nb = (unsigned)brick_count(pa);
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_deposit_f64_brick<RR>), dim3(nb), dim3(F64B_NT), 0, st, bstart, perm, X,
                                         np, lgN, pa, mass, mesh, partial));
    CPM_LAUNCH_CHECK();
}

template <typename PP>
void ParticleOps<PP>::drift(Pos* X, const double* P, long n, double cd, hipStream_t st) {
    hipLaunchKernelGGL((k_drift<PP>), dim3(nblk(n, 256)), dim3(256), 0, st, X, P, n, cd);
    CPM_LAUNCH_CHECK();
}

template <typename PP>
void ParticleOps<PP>::zeldovich(int R, const double* phi, int lgNm, double L, double a_pos, double a_mom,
                                Pos* X, double* P, hipStream_t st) {
    const long np = 1L << (R * lgNm);
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_zeldovich<PP, RR>), dim3(nblk(np, 256)), dim3(256), 0, st, phi,
                                         lgNm, L, a_pos, a_mom, X, P));
    CPM_LAUNCH_CHECK();
}

template <typename T>
void reduce_sum(const T* a, long n, int square, double scale, double* partials, long max_partials,
                double* out, hipStream_t st) {
    long nb = max_partials < 1 ? 1 : max_partials;
    long chunk = (n + nb - 1) / nb;
    if (chunk < RED_NT) chunk = RED_NT;
    nb = (n + chunk - 1) / chunk;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL((k_partial_sum<T>), dim3((unsigned)nb), dim3(RED_NT), 0, st, a, n, chunk, square,
                       partials);
    CPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(RED_NT), 0, st, (const double*)partials, nb, scale, out);
    CPM_LAUNCH_CHECK();
}

void sum_partials(const double* partials, long n, double scale, double* scratch, long scratch_n, double* out,
                  hipStream_t st) {
    reduce_sum<double>(partials, n, 0, scale, scratch, scratch_n, out, st);
}

void reduce_max_abs(const double* a, long n, double* partials, long max_partials, double* out, hipStream_t st) {
    long nb = max_partials < 1 ? 1 : max_partials;
    long chunk = (n + nb - 1) / nb;
    if (chunk < RED_NT) chunk = RED_NT;
    nb = (n + chunk - 1) / chunk;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_partial_max_abs, dim3((unsigned)nb), dim3(RED_NT), 0, st, a, n, chunk, partials);
    CPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_final_max, dim3(1), dim3(RED_NT), 0, st, (const double*)partials, nb, out);
    CPM_LAUNCH_CHECK();
}

void quantize(double* a, long n, double q, hipStream_t st) {
    hipLaunchKernelGGL(k_quantize, dim3(nblk(n, 256)), dim3(256), 0, st, a, n, q);
    CPM_LAUNCH_CHECK();
}

void fx_to_f64(const uint32_t* q, long n, double L, double* x, hipStream_t st) {
    hipLaunchKernelGGL(k_fx_to_f64, dim3(nblk(n, 256)), dim3(256), 0, st, q, n, L, x);
    CPM_LAUNCH_CHECK();
}
void f64_to_fx(const double* x, long n, double L, uint32_t* q, hipStream_t st) {
    hipLaunchKernelGGL(k_f64_to_fx, dim3(nblk(n, 256)), dim3(256), 0, st, x, n, L, q);
    CPM_LAUNCH_CHECK();
}

template struct ParticleOps<PosF64>;
template struct ParticleOps<PosFX>;
template void reduce_sum<float>(const float*, long, int, double, double*, long, double*, hipStream_t);
template void reduce_sum<double>(const double*, long, int, double, double*, long, double*, hipStream_t);

}  // namespace cpm
