// pm_adjoint.hip -- adjoint kernels of the FX (fixed-point position) PM step.
//
// The forward step, per particle p with position Y_p (after the drift) and
// momentum P_p:
//   rho(c)  = m sum_p W(Y_p, c)                    CIC deposit
//   phi     = G rho                                Poisson (symmetric, DC = 0)
//   F_p     = sum_c W(Y_p, c) (D phi)(c)           FD gradient, CIC gather
//   P_p'    = P_p - ck F_p,   Y_p' = Y_p + cd' P_p'
// Its adjoint, given lY' = dL/dY', lP' = dL/dP':
//   lP~     = lP' + cd' lY'                        drift
//   s_k(c)  = sum_p lP~_pk W(Y_p, c)               deposit of lP~ (R meshes)
//   chi     = G (-sum_k D_k s_k)                   D^T = -D, G^T = G
//   lY_pk   = lY'_pk - ck (sum_i lP~_pi dF_pi/dY_pk + m sum_c dW(Y_p, c)/dY_pk chi(c))
//   lP_p    = lP~_p
// The deposit's adjoint is a gather (of chi, with the derivative of the CIC
// weights) and the gather's adjoint is a deposit (of lP~): each kernel here
// is the transpose of a forward kernel in pm_particles.hip, over the same
// brick order.  The reverse pass also reconstructs the state of the previous
// step: the drift is undone exactly in integer arithmetic, and the kick is
// undone with the force recomputed from the same positions (the FX deposit
// is order independent, so it reproduces the forward's mesh bit for bit).
#include "pm_adjoint.hpp"
#include "pm_brick.hpp"
#include "pm_sort.hpp"

namespace cpm {

// The adjoint state (lY, lP) is stored in FP32; every update is computed in
// FP64 from the stored values and rounded once.  The drift's adjoint
// lP + cd lY, in one place so that the fused path (k_adj_gather_fx_brick) and
// the separate one (k_rev_drift_keys) round it identically.
__device__ __forceinline__ float adj_drift(float lp, double cd, float ly) {
    return (float)((double)lp + cd * (double)ly);
}

// splitmix64's output function (Steele, Lea, Flood 2014; constants of
// Stafford's "Mix13"): a bijective 64-bit mixer.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// The checksum term of one particle: id, then position and momentum bits
// axis by axis (k_state_hash and the fused adjoint gather).
__device__ __forceinline__ unsigned long long hash_start(uint32_t id) {
    return mix64((unsigned long long)id + 0x9E3779B97F4A7C15ull);
}
__device__ __forceinline__ unsigned long long hash_axis(unsigned long long h, uint32_t x, double p) {
    h = mix64(h ^ x);
    return mix64(h ^ (unsigned long long)__double_as_longlong(p));
}

// ------------------------------------------------------------ reverse drift + keys
template <int R>
__global__ __launch_bounds__(256) void k_rev_drift_keys(uint32_t* __restrict__ Y, const double* __restrict__ P,
                                                         const float* __restrict__ lY, float* __restrict__ lP,
                                                         long np, double cd_fx, double cd, int lgN, PosArgs pa,
                                                         const uint32_t* __restrict__ prevkeys,
                                                         uint32_t* __restrict__ keys, uint32_t* __restrict__ pids,
                                                         uint32_t* __restrict__ in_count, uint32_t* __restrict__ out_count,
                                                         uint32_t* __restrict__ bmax) {
    __shared__ float wmax[4];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float mx = 0.f;
    if (i < np) {
        const int mask = (1 << lgN) - 1;
        int oc[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < R; ++k) {
            uint32_t y = Y[k * np + i];
            if (cd_fx != 0.0) {
                y -= fx_round_u32(P[k * np + i] * cd_fx);
                Y[k * np + i] = y;
            }
            const float lp = adj_drift(lP[k * np + i], cd, lY[k * np + i]);
            lP[k * np + i] = lp;
            mx = fmaxf(mx, fabsf(lp));
            oc[k] = (int)(y >> pa.fshift) & mask;
        }
        const uint32_t key = brick_key(oc[0], oc[1], oc[2], R, pa);
        keys[i] = key;
        fx_note_key(key, i, prevkeys, pids, in_count, out_count);
    }
    // block max -> bmax[block] (no atomics: one contended address would
    // serialise millions of them)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0)
        bmax[blockIdx.x] = __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3])));
}

// *lmax = max of n non-negative float bit patterns (one workgroup)
__global__ __launch_bounds__(1024) void k_max_u32(const uint32_t* __restrict__ b, long n, uint32_t* __restrict__ lmax) {
    __shared__ uint32_t red[1024];
    uint32_t m = 0;
    for (long i = threadIdx.x; i < n; i += 1024) m = max(m, b[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *lmax = red[0];
}

// ------------------------------------------------------------ vector deposit
// The brick deposit of pm_particles.hip with R value channels instead of the
// mass: 64-bit fixed-point LDS accumulators, scaled so that the largest
// |value| maps below 2^40 (headroom for 2^23 contributions per cell).  The
// R accumulator tiles take ~96 KB of LDS, so one workgroup fills a CU: it
// runs 1024 threads to keep 4 waves per SIMD waiting on the LDS atomics.
static constexpr int VEC_NT = 1024;
template <int R>
__global__ __launch_bounds__(VEC_NT) void k_deposit_fx_vec(const uint32_t* __restrict__ bstart,
                                                           const uint32_t* __restrict__ perm,
                                                           const uint32_t* __restrict__ Y,
                                                           const float* __restrict__ lP, long np, int lgN,
                                                           PosArgs pa, const uint32_t* __restrict__ lmax,
                                                           float* __restrict__ s, long M) {
    __shared__ unsigned long long acc[R][FXB_CELLS];
    __shared__ BrickRanges rg;
    const int N = 1 << lgN, msk = N - 1;
    const BrickCoord bc = brick_coord(blockIdx.x, pa);
    const int ncell = bc.B[0] * bc.B[1] * bc.B[2];
    const int tid = threadIdx.x;
    for (int e = tid; e < ncell; e += VEC_NT)
#pragma unroll
        for (int c = 0; c < R; ++c) acc[c][e] = 0ull;
    const float mx = __uint_as_float(*lmax);
    double scale = 0.0;
    if (mx > 0.f) {
        int ex = 0;
        (void)frexpf(mx, &ex);  // mx < 2^ex
        scale = ldexp(1.0, 40 - ex);
    }
    brick_ranges<R>(rg, bstart, bc, pa, tid);
    const uint32_t total = rg.roff[rg.n];
    for (uint32_t t = tid; t < total; t += VEC_NT) {
        const uint32_t sp = brick_range_pos(rg, t);
        const uint32_t i = perm[sp];
        int o[3] = {0, 0, 0};
        float f[3] = {0.f, 0.f, 0.f};
        double lam[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < R; ++k) {
            PosFX::dep(Y[k * np + i], pa, o[k], f[k]);
            lam[k] = (double)lP[k * np + i] * scale;
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
            float w = 1.f;
#pragma unroll
            for (int k = 0; k < R; ++k) w *= ((j >> k) & 1) ? f[k] : 1.f - f[k];
            const int cell = l[0] + bc.B[0] * (l[1] + bc.B[1] * l[2]);
#pragma unroll
            for (int c = 0; c < R; ++c)
                atomicAdd(&acc[c][cell], (unsigned long long)__double2ll_rn((double)w * lam[c]));
        }
    }
    __syncthreads();
    const double inv = scale > 0 ? 1.0 / scale : 0.0;  // exact: scale is a power of two
    for (int e = tid; e < ncell; e += VEC_NT) {
        const int lx = e & (bc.B[0] - 1), ly = (e >> pa.lb[0]) & (bc.B[1] - 1), lz = e >> (pa.lb[0] + pa.lb[1]);
        const long idx = (long)(bc.x0[0] + lx) + (long)N * ((bc.x0[1] + ly) + (long)N * (bc.x0[2] + lz));
#pragma unroll
        for (int c = 0; c < R; ++c) s[c * M + idx] = (float)((double)(long long)acc[c][e] * inv);
    }
}

// The vector deposit of lP and the mass deposit of the same positions in one
// launch, for the backward pass that recomputes the potential: two
// workgroups per brick, each with two accumulator tiles (64 KB of LDS, so two
// 1024-thread workgroups share a CU where one k_deposit_fx_vec fills it).
// Workgroup 2b deposits lP components 0 and 1 of brick b; workgroup 2b + 1
// component 2 (3-D) and the mass.  The mass channel is k_deposit_fx_brick's:
// the same weights and quanta, the same FP32 cell values, and the brick's
// partial sum in the same order (threads 0..511 sum as that kernel's 512
// threads do, the waves in order), so the mesh and the partials are
// bit-identical to a separate mass deposit.
template <int R>
__global__ __launch_bounds__(VEC_NT) void k_deposit_fx_vm(const uint32_t* __restrict__ bstart,
                                                          const uint32_t* __restrict__ perm,
                                                          const uint32_t* __restrict__ Y,
                                                          const float* __restrict__ lP, long np, int lgN,
                                                          PosArgs pa, const uint32_t* __restrict__ lmax,
                                                          float* __restrict__ s, long M, float mass,
                                                          float* __restrict__ mesh, double* __restrict__ partial) {
    static_assert(VEC_NT >= FXB_NT, "the mass channel's output uses FXB_NT threads");
    __shared__ unsigned long long acc[2][FXB_CELLS];
    __shared__ BrickRanges rg;
    __shared__ double red[VEC_NT / 64];
    const int N = 1 << lgN, msk = N - 1;
    const uint32_t brick = blockIdx.x >> 1;
    const int half = (int)(blockIdx.x & 1);
    const BrickCoord bc = brick_coord(brick, pa);
    const int ncell = bc.B[0] * bc.B[1] * bc.B[2];
    const int tid = threadIdx.x;
    // this workgroup's lP components: c0, c0 + 1, .. (nl of them); the mass in acc[1] (half 1)
    const int c0 = 2 * half;
    const int nl = half ? (R == 3 ? 1 : 0) : 2;
    const bool hm = half == 1;
    for (int e = tid; e < ncell; e += VEC_NT) {
        acc[0][e] = 0ull;
        acc[1][e] = 0ull;
    }
    const float mx = __uint_as_float(*lmax);
    double scale = 0.0;
    if (mx > 0.f) {
        int ex = 0;
        (void)frexpf(mx, &ex);  // mx < 2^ex
        scale = ldexp(1.0, 40 - ex);
    }
    brick_ranges<R>(rg, bstart, bc, pa, tid);
    const uint32_t total = rg.roff[rg.n];
    // A wave's lanes are spread over the workgroup's chunk of VEC_NT
    // particles, 8 stretches of 8: neighbours in the sorted order, which
    // often share cells, then fall in different waves, so fewer of one
    // instruction's 64-bit LDS atomics go to the same address (4.44 -> 3.70
    // ms per launch at 512^3 particles on 512^3; 16 or more stretches lose
    // more to the scattered loads than they gain, 4 gain less).
    constexpr int SP = 8;
    const uint32_t tmap = (uint32_t)((tid & (SP - 1)) * (VEC_NT / SP) + (tid / SP));
    for (uint32_t t0 = 0; t0 < total; t0 += VEC_NT) {
        const uint32_t t = t0 + tmap;
        if (t >= total) continue;
        const uint32_t sp = brick_range_pos(rg, t);
        const uint32_t i = perm[sp];
        int o[3] = {0, 0, 0};
        float f[3] = {0.f, 0.f, 0.f};
        double lam[2] = {0, 0};
#pragma unroll
        for (int k = 0; k < R; ++k) PosFX::dep(Y[k * np + i], pa, o[k], f[k]);
#pragma unroll
        for (int q = 0; q < 2; ++q)
            if (q < nl) lam[q] = (double)lP[(c0 + q) * np + i] * scale;
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
            const int cell = l[0] + bc.B[0] * (l[1] + bc.B[1] * l[2]);
            float w = 1.f;  // as k_deposit_fx_vec
#pragma unroll
            for (int k = 0; k < R; ++k) w *= ((j >> k) & 1) ? f[k] : 1.f - f[k];
#pragma unroll
            for (int q = 0; q < 2; ++q)
                if (q < nl) atomicAdd(&acc[q][cell], (unsigned long long)__double2ll_rn((double)w * lam[q]));
            if (hm) {
                float wm = mass;  // as k_deposit_fx_brick
#pragma unroll
                for (int k = 0; k < R; ++k) wm *= ((j >> k) & 1) ? f[k] : 1.f - f[k];
                atomicAdd(&acc[1][cell], fx_quantum(wm));
            }
        }
    }
    __syncthreads();
    const double inv = scale > 0 ? 1.0 / scale : 0.0;  // exact: scale is a power of two
    auto cell_index = [&](int e) -> long {
        const int lx = e & (bc.B[0] - 1), ly = (e >> pa.lb[0]) & (bc.B[1] - 1), lz = e >> (pa.lb[0] + pa.lb[1]);
        return (long)(bc.x0[0] + lx) + (long)N * ((bc.x0[1] + ly) + (long)N * (bc.x0[2] + lz));
    };
    if (!hm) {
        for (int e = tid; e < ncell; e += VEC_NT) {
            const long idx = cell_index(e);
#pragma unroll
            for (int q = 0; q < 2; ++q) s[(c0 + q) * M + idx] = (float)((double)(long long)acc[q][e] * inv);
        }
        return;
    }
    // half 1: threads FXB_NT.. write lP component 2 (3-D); threads 0..FXB_NT-1
    // write the mass channel exactly as k_deposit_fx_brick does
    double rsum = 0;
    if (tid >= FXB_NT) {
        if (R == 3)
            for (int e = tid - FXB_NT; e < ncell; e += VEC_NT - FXB_NT)
                s[2 * M + cell_index(e)] = (float)((double)(long long)acc[0][e] * inv);
    } else if (R == 3 && bc.B[0] * bc.B[1] == FXB_NT) {
        const int lx = tid & (bc.B[0] - 1), ly = tid >> pa.lb[0];
        float* __restrict__ out = mesh + (long)(bc.x0[0] + lx) + (long)N * ((bc.x0[1] + ly) + (long)N * bc.x0[2]);
        const long plane = (long)N * N;
        for (int lz = 0; lz < bc.B[2]; ++lz) {
            const float v = (float)((double)(long long)acc[1][tid + lz * FXB_NT] * (1.0 / FX_ONE));
            out[lz * plane] = v;
            rsum += (double)v;
        }
    } else {
        for (int e = tid; e < ncell; e += FXB_NT) {
            const float v = (float)((double)(long long)acc[1][e] * (1.0 / FX_ONE));
            mesh[cell_index(e)] = v;
            rsum += (double)v;
        }
    }
    if (!partial) return;  // (uniform) no mesh sum wanted
    // block_sum<FXB_NT> over threads 0..FXB_NT-1 (the other waves add 0 and are not read)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) rsum += __shfl_down(rsum, off);
    if ((tid & 63) == 0) red[tid >> 6] = rsum;
    __syncthreads();
    if (tid == 0) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < FXB_NT / 64; ++w) t += red[w];
        partial[brick] = t;
    }
}

// ------------------------------------------------------------ -div (FP32 mesh)
template <int R>
__global__ __launch_bounds__(256) void k_neg_div_f32(const float* __restrict__ s, int lgN, long M,
                                                      float inv12res, float* __restrict__ psi) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M) return;
    const int msk = (1 << lgN) - 1;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const long stride = 1L << (k * lgN);
        const int q = (int)((c >> (k * lgN)) & msk);
        const long base = c - (long)q * stride;
        const float* sk = s + k * M;
        const float fm2 = sk[base + (long)((q - 2) & msk) * stride];
        const float fm1 = sk[base + (long)((q - 1) & msk) * stride];
        const float fp1 = sk[base + (long)((q + 1) & msk) * stride];
        const float fp2 = sk[base + (long)((q + 2) & msk) * stride];
        acc += fd4(fm2, fm1, fp1, fp2, inv12res);
    }
    psi[c] = -acc;
}

// ------------------------------------------------------------ fused adjoint gather
// One workgroup per brick.  phi (brick + halo -2..+3) and chi (brick + 1) are
// staged in LDS.  Per particle: the forward's FD-gradient gather F (same
// operations in the same order as k_gather_kick_fx_brick, so the force is
// bit-identical), its per-corner gradients g_i(c), and the derivative of the
// CIC weights dW(c)/dY_k = +-1/res * prod_{m != k} A_m(c).
static constexpr int FXB_CTILE = (FXB_X + 1) * (FXB_Y + 1) * (FXB_Z + 1);

// HASH: dr.hash is set (a separate instance, so that the checksum's
// registers do not weigh on the launches without it).  FULL: every brick has
// the full FX size (the mesh is at least a brick along each axis), so the
// tile strides are compile-time constants and the tile reads take their
// offsets as immediates.
template <int R, bool HASH, bool FULL>
__global__ __launch_bounds__(FXG_NT) __attribute__((amdgpu_waves_per_eu(4))) void k_adj_gather_fx_brick(
    const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ Y,
    const double* __restrict__ P, const float* __restrict__ lY, const float* __restrict__ lP,
    const uint32_t* __restrict__ idi, uint32_t* __restrict__ Yo, double* __restrict__ Po,
    float* __restrict__ lYo, float* __restrict__ lPo, uint32_t* __restrict__ ido, long np,
    const float* __restrict__ phi, const float* __restrict__ chi, int lgN, PosArgs pa, AdjCoef cf, AdjDrift dr) {
    __shared__ float tile[FXB_TILE];
    __shared__ float ctile[FXB_CTILE];
    __shared__ float wmax[FXG_NT / 64];
    __shared__ unsigned long long whash[FXG_NT / 64];
    const double cFq = cf.cF * cf.ipq;  // the forward's ck / pq for cF = ck (exact: ipq is a power of two)
    const int N = 1 << lgN, msk = N - 1;
    const uint32_t bid = blockIdx.x;
    const BrickCoord bc = brick_coord(bid, pa);
    constexpr int FB0 = FXB_X, FB1 = (R == 3) ? FXB_Y : FXB_X, FB2 = (R == 3) ? FXB_Z : 1;
    const int B0 = FULL ? FB0 : bc.B[0], B1 = FULL ? FB1 : bc.B[1], B2 = FULL ? FB2 : bc.B[2];
    const int TX = B0 + 5, TY = B1 + 5, TZ = (R == 3) ? B2 + 5 : 1;
    const int CX = B0 + 1, CY = B1 + 1, CZ = (R == 3) ? B2 + 1 : 1;
    const int tid = threadIdx.x;
    const bool hp = phi != nullptr, hc = chi != nullptr;
    // A thread's first particle is loaded before the tiles are staged, so the
    // latencies overlap; each later one at the end of the previous iteration.
    const uint32_t s0 = bstart[bid << R], s1 = bstart[(bid + 1) << R];
    uint32_t ys[3] = {0u, 0u, 0u}, idv = 0u, iv = 0u;
    // the particle's state, loaded one iteration ahead: lP (FP32 values), P, lY
    double lam[3] = {0, 0, 0};
    double pin[3] = {0, 0, 0};
    float lyin[3] = {0.f, 0.f, 0.f};
    auto fetch = [&](uint32_t sp) {
        iv = perm[sp];
        idv = idi[iv];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            ys[k] = Y[k * np + iv];
            lam[k] = (double)lP[k * np + iv];
            pin[k] = P[k * np + iv];
            lyin[k] = lY[k * np + iv];
        }
    };
    if (s0 + tid < s1) fetch(s0 + tid);
    if (hp) stage_box<FXG_NT, FXB_TILE, R>(tile, phi, N, bc.x0, -2, TX, TY, TZ, tid);
    if (hc) stage_box<FXG_NT, FXB_CTILE, R>(ctile, chi, N, bc.x0, 0, CX, CY, CZ, tid);
    __syncthreads();
    const float inv12res = (float)(1.0 / (12.0 * pa.res));
    const double ires = 1.0 / pa.res;
    constexpr int NC = 1 << R;
    float mx = 0.f;
    constexpr bool hh = HASH;
    unsigned long long hsum = 0;  // checksum of the state before the drift's undo (dr.hash)
    for (uint32_t s = s0 + tid; s < s1; s += FXG_NT) {
        ido[s] = idv;
        int o[3] = {0, 0, 0};
        float a1[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < R; ++k) PosFX::gat(ys[k], pa, o[k], a1[k]);
        const int lx = (o[0] - bc.x0[0]) & msk, ly = (o[1] - bc.x0[1]) & msk;
        const int lz = (R == 3) ? ((o[2] - bc.x0[2]) & msk) : 0;
        float z[NC];
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            float w = 1.f;
#pragma unroll
            for (int k = 0; k < R; ++k) w *= ((j >> k) & 1) ? a1[k] : 1.f - a1[k];
            z[j] = w;
        }
        // X[j] = cT h_j + cU chi(c_j): what multiplies dW(c_j)/dY_k in both
        // terms of lY's update, T = sum_j dW(c_j)/dY_k h_j (h_j = lP . g(c_j))
        // and U = sum_j dW(c_j)/dY_k chi(c_j)
        float F[3] = {0.f, 0.f, 0.f};
        double X[NC];
#pragma unroll
        for (int j = 0; j < NC; ++j) X[j] = 0;
        if (hp) {
            float g[3][NC];
            const int zc = (R == 3) ? 2 : 0;
            auto Tt = [&](int tx, int ty, int tz) -> float { return tile[tx + TX * (ty + TY * tz)]; };
#pragma unroll
            for (int l = 0; l < NC / 2; ++l) {
                const int by = l & 1, bz = (l >> 1) & 1;
                float f[6];
#pragma unroll
                for (int t = 0; t < 6; ++t) f[t] = Tt(lx + t, ly + 2 + by, lz + zc + bz);
#pragma unroll
                for (int bx = 0; bx < 2; ++bx) {
                    const int j = bx | (by << 1) | (bz << 2);
                    g[0][j] = fd4(f[bx], f[bx + 1], f[bx + 3], f[bx + 4], inv12res);
                    F[0] = fmaf(g[0][j], z[j], F[0]);  // as k_gather_kick_fx_brick
                }
            }
#pragma unroll
            for (int l = 0; l < NC / 2; ++l) {
                const int bx = l & 1, bz = (l >> 1) & 1;
                float f[6];
#pragma unroll
                for (int t = 0; t < 6; ++t) f[t] = Tt(lx + 2 + bx, ly + t, lz + zc + bz);
#pragma unroll
                for (int by = 0; by < 2; ++by) {
                    const int j = bx | (by << 1) | (bz << 2);
                    g[1][j] = fd4(f[by], f[by + 1], f[by + 3], f[by + 4], inv12res);
                    F[1] = fmaf(g[1][j], z[j], F[1]);  // as k_gather_kick_fx_brick
                }
            }
            if constexpr (R == 3) {
#pragma unroll
                for (int l = 0; l < 4; ++l) {
                    const int bx = l & 1, by = (l >> 1) & 1;
                    float f[6];
#pragma unroll
                    for (int t = 0; t < 6; ++t) f[t] = Tt(lx + 2 + bx, ly + 2 + by, lz + t);
#pragma unroll
                    for (int bz = 0; bz < 2; ++bz) {
                        const int j = bx | (by << 1) | (bz << 2);
                        g[2][j] = fd4(f[bz], f[bz + 1], f[bz + 3], f[bz + 4], inv12res);
                        F[2] = fmaf(g[2][j], z[j], F[2]);  // as k_gather_kick_fx_brick
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                double h = 0;
#pragma unroll
                for (int q = 0; q < R; ++q) h += lam[q] * (double)g[q][j];
                X[j] = cf.cT * h;
            }
        }
        if (hc) {
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                const int bx = j & 1, by = (j >> 1) & 1, bz = (j >> 2) & 1;
                const double cv = (double)ctile[(lx + bx) + CX * ((ly + by) + CY * (lz + (R == 3 ? bz : 0)))];
                X[j] += cf.cU * cv;
            }
        }
        // V[k] = sum_j dW(c_j)/dY_k X[j].  dW(c_j)/dY_k = +-(1/res) prod_{m != k}
        // A_m(c_j), + where c_j is the upper corner along k: the corners pair
        // up along k with one product, V[k] = (1/res) sum over the upper
        // corners j of prod_{m != k} A_m(c_j) (X[j] - X[j - 2^k]).
        double A[3][2];
#pragma unroll
        for (int m = 0; m < R; ++m) {
            A[m][1] = (double)a1[m];
            A[m][0] = 1.0 - (double)a1[m];
        }
        double V[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < R; ++k) {
            double v = 0;
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                if (!((j >> k) & 1)) continue;
                double w = 1.0;
#pragma unroll
                for (int m = 0; m < R; ++m)
                    if (m != k) w *= A[m][(j >> m) & 1];
                v += w * (X[j] - X[j ^ (1 << k)]);
            }
            V[k] = v * ires;
        }
        uint32_t yn[3] = {0u, 0u, 0u};
        unsigned long long h = hh ? hash_start(idv) : 0ull;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const double p = pin[k] + fx_kick_increment((double)F[k], cFq, cf.pq);
            const float lyk = (float)((double)lyin[k] + V[k]);
            float lpk = (float)lam[k];
            if (hh) h = hash_axis(h, ys[k], p);
            yn[k] = ys[k];
            if (dr.on) {  // the previous step's drift, undone, and its adjoint
                yn[k] -= fx_round_u32(p * dr.cd_fx);
                lpk = adj_drift(lpk, dr.cd, lyk);
            }
            Yo[k * np + s] = yn[k];
            Po[k * np + s] = p;
            lYo[k * np + s] = lyk;
            lPo[k * np + s] = lpk;
            mx = fmaxf(mx, fabsf(lpk));
        }
        hsum += h;
        if (dr.keys) {
            int co[3] = {0, 0, 0}, cn[3] = {0, 0, 0};
#pragma unroll
            for (int k = 0; k < R; ++k) {
                co[k] = (int)(ys[k] >> pa.fshift) & msk;
                cn[k] = (int)(yn[k] >> pa.fshift) & msk;
            }
            const uint32_t ko = brick_key(co[0], co[1], co[2], R, pa), kn = brick_key(cn[0], cn[1], cn[2], R, pa);
            dr.keys[s] = kn;
            if (kn != ko) {
                atomicAdd(&dr.in_count[kn], 1u);
                atomicAdd(&dr.out_count[ko], 1u);
            }
        }
        if (s + FXG_NT < s1) fetch(s + FXG_NT);
    }
    if constexpr (HASH) {  // integer sum: the same for any order of the particles
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) hsum += __shfl_xor(hsum, off);
        if ((tid & 63) == 0) whash[tid >> 6] = hsum;
        __syncthreads();
        if (tid == 0) {
            unsigned long long t = 0;
#pragma unroll
            for (int w = 0; w < FXG_NT / 64; ++w) t += whash[w];
            atomicAdd(dr.hash, t);
        }
    }
    if (dr.bmax) {  // block max of |lP| out -> bmax[block]
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
        if ((tid & 63) == 0) wmax[tid >> 6] = mx;
        __syncthreads();
        if (tid == 0) {
            float m = 0.f;
#pragma unroll
            for (int w = 0; w < FXG_NT / 64; ++w) m = fmaxf(m, wmax[w]);
            dr.bmax[bid] = __float_as_uint(m);
        }
    }
}

// ------------------------------------------------------------ state checksum
// *sum += sum over particles of a hash of (id, positions, momentum bits):
// integer addition, so the value does not depend on the storage order.  With
// Xc != nullptr the state is also copied to (Xc, Pc, idc) in the same pass.
template <int R>
__global__ __launch_bounds__(256) void k_state_hash(const uint32_t* __restrict__ X, const double* __restrict__ P,
                                                    const uint32_t* __restrict__ id, long np,
                                                    uint32_t* __restrict__ Xc, double* __restrict__ Pc,
                                                    uint32_t* __restrict__ idc, unsigned long long* __restrict__ sum) {
    __shared__ unsigned long long red[4];
    unsigned long long acc = 0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < np; i += (long)gridDim.x * 256) {
        const uint32_t d = id[i];
        unsigned long long h = hash_start(d);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const uint32_t x = X[k * np + i];
            const double pk = P[k * np + i];
            h = hash_axis(h, x, pk);
            if (Xc) {
                Xc[k * np + i] = x;
                Pc[k * np + i] = pk;
            }
        }
        if (idc) idc[i] = d;
        acc += h;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(sum, red[0] + red[1] + red[2] + red[3]);
}

// ------------------------------------------------------------ permutations by particle id
template <typename T>
__global__ __launch_bounds__(256) void k_scatter_by_id(const T* __restrict__ src, const uint32_t* __restrict__ id,
                                                        long np, int nc, T* __restrict__ dst) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    const long d = id[s];
    for (int k = 0; k < nc; ++k) dst[k * np + d] = src[k * np + s];
}
template <typename T>
__global__ __launch_bounds__(256) void k_gather_by_id(const T* __restrict__ src, const uint32_t* __restrict__ id,
                                                       long np, int nc, T* __restrict__ dst) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    const long d = id[s];
    for (int k = 0; k < nc; ++k) dst[k * np + s] = src[k * np + d];
}

// ------------------------------------------------------------ -div (FP64 mass grid)
template <int R>
__global__ __launch_bounds__(256) void k_neg_div_f64(const double* __restrict__ u, int lgN, long n, double res,
                                                      double* __restrict__ psi) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int msk = (1 << lgN) - 1;
    double acc = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const long stride = 1L << (k * lgN);
        const int q = (int)((c >> (k * lgN)) & msk);
        const long base = c - (long)q * stride;
        const double* uk = u + k * n;
        acc += fd4(uk[base + (long)((q - 2) & msk) * stride], uk[base + (long)((q - 1) & msk) * stride],
                   uk[base + (long)((q + 1) & msk) * stride], uk[base + (long)((q + 2) & msk) * stride], res);
    }
    psi[c] = -acc;
}

// ------------------------------------------------------------ small vector kernels
__global__ __launch_bounds__(256) void k_axpby_f32(long n, double a, const float* __restrict__ x, double b,
                                                   const float* __restrict__ y, double* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * (double)x[i] + b * (double)y[i];
}
__global__ __launch_bounds__(256) void k_axpby(long n, double a, const double* __restrict__ x, double b,
                                               const double* __restrict__ y, double* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i] + b * y[i];
}

static constexpr int DOT_NT = 256;
__global__ __launch_bounds__(DOT_NT) void k_dot_partial(const double* __restrict__ a, const double* __restrict__ b,
                                                        long n, long chunk, double* __restrict__ partials) {
    __shared__ double red[DOT_NT];
    const long b0 = (long)blockIdx.x * chunk;
    const long b1 = b0 + chunk < n ? b0 + chunk : n;
    double s = 0;
    for (long i = b0 + threadIdx.x; i < b1; i += DOT_NT) s += a[i] * b[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = DOT_NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(DOT_NT) void k_dot_final(const double* __restrict__ partials, long n,
                                                      double* __restrict__ out) {
    __shared__ double red[DOT_NT];
    double s = 0;
    for (long i = threadIdx.x; i < n; i += DOT_NT) s += partials[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = DOT_NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

__global__ __launch_bounds__(256) void k_f32_to_f64(const float* __restrict__ a, long n, double* __restrict__ b) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}
__global__ __launch_bounds__(256) void k_f64_to_f32(const double* __restrict__ a, long n, float* __restrict__ b) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ __launch_bounds__(256) void k_iota_u32(uint32_t* __restrict__ p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, long n, uint32_t v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_state_diff(const uint32_t* __restrict__ Ya, const uint32_t* __restrict__ Yb,
                                                    const double* __restrict__ Pa, const double* __restrict__ Pb,
                                                    long n, double* __restrict__ partials) {
    __shared__ double r0[256], r1[256], r2[256];
    double cnt = 0, mdp = 0, mp = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        cnt += (Ya[i] != Yb[i]) ? 1.0 : 0.0;
        mdp = fmax(mdp, fabs(Pa[i] - Pb[i]));
        mp = fmax(mp, fabs(Pb[i]));
    }
    r0[threadIdx.x] = cnt;
    r1[threadIdx.x] = mdp;
    r2[threadIdx.x] = mp;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            r0[threadIdx.x] += r0[threadIdx.x + off];
            r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + off]);
            r2[threadIdx.x] = fmax(r2[threadIdx.x], r2[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x] = r0[0];
        partials[3 * blockIdx.x + 1] = r1[0];
        partials[3 * blockIdx.x + 2] = r2[0];
    }
}

// ------------------------------------------------------------ launchers
void rev_drift_keys(int R, uint32_t* Y, const double* P, const float* lY, float* lP, long np, double cd_fx,
                    double cd, int lgN, const PosArgs& pa, const uint32_t* prevkeys, uint32_t* keys, uint32_t* pids,
                    uint32_t* in_count, uint32_t* out_count, uint32_t* bmax, uint32_t* lmax, hipStream_t st) {
    const unsigned nb = nblk(np, 256);
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_rev_drift_keys<RR>), dim3(nb), dim3(256), 0, st, Y, P, lY, lP, np, cd_fx,
                                         cd, lgN, pa, prevkeys, keys, pids, in_count, out_count, bmax));
    CPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_max_u32, dim3(1), dim3(1024), 0, st, (const uint32_t*)bmax, (long)nb, lmax);
    CPM_LAUNCH_CHECK();
}

void deposit_fx_vec(int R, const uint32_t* bstart, const uint32_t* perm, const uint32_t* Y, const float* lP,
                    long np, int lgN, const PosArgs& pa, const uint32_t* lmax, float* s, hipStream_t st) {
    const unsigned nb = (unsigned)brick_count(pa);
    const long M = 1L << (R * lgN);
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_deposit_fx_vec<RR>), dim3(nb), dim3(VEC_NT), 0, st, bstart, perm, Y,
                                         lP, np, lgN, pa, lmax, s, M));
    CPM_LAUNCH_CHECK();
}

void deposit_fx_vm(int R, const uint32_t* bstart, const uint32_t* perm, const uint32_t* Y, const float* lP, long np,
                   int lgN, const PosArgs& pa, const uint32_t* lmax, float* s, float mass, float* mesh,
                   double* partial, hipStream_t st) {
    const unsigned nb = (unsigned)brick_count(pa);
    const long M = 1L << (R * lgN);
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_deposit_fx_vm<RR>), dim3(2 * nb), dim3(VEC_NT), 0, st, bstart, perm, Y,
                                         lP, np, lgN, pa, lmax, s, M, mass, mesh, partial));
    CPM_LAUNCH_CHECK();
}

void neg_div_f32(int R, const float* s, int lgN, double res, float* psi, hipStream_t st) {
    const long M = 1L << (R * lgN);
    const float inv12res = (float)(1.0 / (12.0 * res));
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_neg_div_f32<RR>), dim3(nblk(M, 256)), dim3(256), 0, st, s, lgN, M,
                                         inv12res, psi));
    CPM_LAUNCH_CHECK();
}

void adj_gather_fx_brick(int R, const uint32_t* bstart, const uint32_t* perm, const uint32_t* Y, const double* P,
                         const float* lY, const float* lP, const uint32_t* idi, uint32_t* Yo, double* Po,
                         float* lYo, float* lPo, uint32_t* ido, long np, const float* phi, const float* chi,
                         int lgN, const PosArgs& pa, AdjCoef c, AdjDrift dr, uint32_t* lmax, hipStream_t st) {
    const unsigned nb = (unsigned)brick_count(pa);
    if (dr.hash) zero_u32(reinterpret_cast<uint32_t*>(dr.hash), 2, st);
    // every brick at the full FX size: compile-time tile strides
    const bool full = (1 << pa.lb[0]) == FXB_X && (1 << pa.lb[1]) == (R == 3 ? FXB_Y : FXB_X) &&
                      (R < 3 || (1 << pa.lb[2]) == FXB_Z);
#define CPM_ADJ_LAUNCH(H, F)                                                                                       \
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_adj_gather_fx_brick<RR, H, F>), dim3(nb), dim3(FXG_NT), 0, st, bstart, \
                                         perm, Y, P, lY, lP, idi, Yo, Po, lYo, lPo, ido, np, phi, chi, lgN, pa, c,  \
                                         dr))
    if (dr.hash) {
        if (full)
            CPM_ADJ_LAUNCH(true, true);
        else
            CPM_ADJ_LAUNCH(true, false);
    } else {
        if (full)
            CPM_ADJ_LAUNCH(false, true);
        else
            CPM_ADJ_LAUNCH(false, false);
    }
#undef CPM_ADJ_LAUNCH
    CPM_LAUNCH_CHECK();
    if (dr.bmax) {
        hipLaunchKernelGGL(k_max_u32, dim3(1), dim3(1024), 0, st, (const uint32_t*)dr.bmax, (long)nb, lmax);
        CPM_LAUNCH_CHECK();
    }
}

void state_hash(int R, const uint32_t* X, const double* P, const uint32_t* id, long np, uint32_t* Xc, double* Pc,
                uint32_t* idc, unsigned long long* sum, hipStream_t st) {
    zero_u32(reinterpret_cast<uint32_t*>(sum), 2, st);
    long nb = (np + 255) / 256;
    if (nb > 8192) nb = 8192;
    if (nb < 1) nb = 1;
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_state_hash<RR>), dim3((unsigned)nb), dim3(256), 0, st, X, P, id, np, Xc,
                                         Pc, idc, sum));
    CPM_LAUNCH_CHECK();
}

template <typename T>
void scatter_by_id(const T* src, const uint32_t* id, long np, int nc, T* dst, hipStream_t st) {
    hipLaunchKernelGGL((k_scatter_by_id<T>), dim3(nblk(np, 256)), dim3(256), 0, st, src, id, np, nc, dst);
    CPM_LAUNCH_CHECK();
}
template <typename T>
void gather_by_id(const T* src, const uint32_t* id, long np, int nc, T* dst, hipStream_t st) {
    hipLaunchKernelGGL((k_gather_by_id<T>), dim3(nblk(np, 256)), dim3(256), 0, st, src, id, np, nc, dst);
    CPM_LAUNCH_CHECK();
}
template void scatter_by_id<double>(const double*, const uint32_t*, long, int, double*, hipStream_t);
template void scatter_by_id<uint32_t>(const uint32_t*, const uint32_t*, long, int, uint32_t*, hipStream_t);
template void gather_by_id<double>(const double*, const uint32_t*, long, int, double*, hipStream_t);
template void gather_by_id<uint32_t>(const uint32_t*, const uint32_t*, long, int, uint32_t*, hipStream_t);
template void scatter_by_id<float>(const float*, const uint32_t*, long, int, float*, hipStream_t);
template void gather_by_id<float>(const float*, const uint32_t*, long, int, float*, hipStream_t);

void neg_div_f64(int R, const double* u, int lgN, double res, double* psi, hipStream_t st) {
    const long n = 1L << (R * lgN);
    CPM_R_DISPATCH(R, hipLaunchKernelGGL((k_neg_div_f64<RR>), dim3(nblk(n, 256)), dim3(256), 0, st, u, lgN, n, res,
                                         psi));
    CPM_LAUNCH_CHECK();
}

void axpby_f32(long n, double a, const float* x, double b, const float* y, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_axpby_f32, dim3(nblk(n, 256)), dim3(256), 0, st, n, a, x, b, y, out);
    CPM_LAUNCH_CHECK();
}
void axpby(long n, double a, const double* x, double b, const double* y, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_axpby, dim3(nblk(n, 256)), dim3(256), 0, st, n, a, x, b, y, out);
    CPM_LAUNCH_CHECK();
}

void dot(const double* a, const double* b, long n, double* partials, long max_partials, double* out,
         hipStream_t st) {
    long nb = max_partials < 1 ? 1 : max_partials;
    long chunk = (n + nb - 1) / nb;
    if (chunk < DOT_NT) chunk = DOT_NT;
    nb = (n + chunk - 1) / chunk;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_dot_partial, dim3((unsigned)nb), dim3(DOT_NT), 0, st, a, b, n, chunk, partials);
    CPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(DOT_NT), 0, st, (const double*)partials, nb, out);
    CPM_LAUNCH_CHECK();
}

void f32_to_f64(const float* a, long n, double* b, hipStream_t st) {
    hipLaunchKernelGGL(k_f32_to_f64, dim3(nblk(n, 256)), dim3(256), 0, st, a, n, b);
    CPM_LAUNCH_CHECK();
}
void f64_to_f32(const double* a, long n, float* b, hipStream_t st) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3(nblk(n, 256)), dim3(256), 0, st, a, n, b);
    CPM_LAUNCH_CHECK();
}
void iota_u32(uint32_t* p, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_iota_u32, dim3(nblk(n, 256)), dim3(256), 0, st, p, n);
    CPM_LAUNCH_CHECK();
}
void fill_u32(uint32_t* p, long n, uint32_t v, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(n, 256)), dim3(256), 0, st, p, n, v);
    CPM_LAUNCH_CHECK();
}
void state_diff(const uint32_t* Ya, const uint32_t* Yb, const double* Pa, const double* Pb, long n,
                double* partials, int nb, hipStream_t st) {
    hipLaunchKernelGGL(k_state_diff, dim3((unsigned)nb), dim3(256), 0, st, Ya, Yb, Pa, Pb, n, partials);
    CPM_LAUNCH_CHECK();
}

}  // namespace cpm
