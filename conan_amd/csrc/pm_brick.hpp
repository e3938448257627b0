// pm_brick.hpp -- device-side helpers of the FX brick order, shared by the
// forward kernels (pm_particles.hip) and their adjoints (pm_adjoint.hip).
//
// The FX form keeps its particle arrays grouped by brick (FXB_X x FXB_Y x
// FXB_Z cells, bricks in x-fastest order) in the order of the previous step:
// `perm` maps sorted position s -> storage index.  The sort key of a particle
// is (brick << R) | face class (brick_key), and bstart[g] is the first sorted
// position of key group g, so bstart[b << R] .. bstart[(b + 1) << R] is brick b.
#pragma once
#include "pm_particles.hpp"

namespace cpm {

static constexpr int FXB_NT = 512;  // deposit workgroup size
#ifndef CPM_FXG_NT
#define CPM_FXG_NT 512
#endif
static constexpr int FXG_NT = CPM_FXG_NT;  // gather workgroup size
#ifndef CPM_GATHER_LSEG
#define CPM_GATHER_LSEG 3  // log2 of the bricks one gather workgroup slides through (FP32)
#endif
#ifndef CPM_GATHER_LSEG64
#define CPM_GATHER_LSEG64 4  // the same for the FP64 gather (its bricks are 4 planes deep)
#endif
static constexpr int FXB_CELLS = FXB_X * FXB_Y * FXB_Z;
// potential tile of a brick: 2 cells of halo below and 3 above along each axis
static constexpr int FXB_TILE = (FXB_X + 5) * (FXB_Y + 5) * (FXB_Z + 5);
static constexpr double FX_ONE = 1099511627776.0;  // 2^40: fixed-point unit of the deposit accumulators

// round(w 2^40) (ties to even) as the 64-bit deposit quantum, for
// 0 <= w < 2^23, in FP32 operations: a = w 2^8 is exact, hi = floor(a) and
// a - hi are exact, (a - hi) 2^32 < 2^32 is exact, and rounding it to an
// integer rounds w 2^40, whose integer part hi 2^32 needs no rounding.
// Equal to __double2ll_rn((double)w * 2^40), without FP64 operations.
__device__ __forceinline__ unsigned long long fx_quantum(float w) {
    const float a = w * 256.f;
    const float hi = floorf(a);
    const float lo = rintf((a - hi) * 4294967296.f);
    return ((unsigned long long)(uint32_t)hi << 32) + (uint32_t)lo;
}

struct BrickCoord {
    int x0[3], B[3], nb[3], bi[3];
};

__device__ __forceinline__ BrickCoord brick_coord(uint32_t b, const PosArgs& pa) {
    BrickCoord c;
    uint32_t t = b;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c.nb[k] = 1 << pa.lnb[k];
        c.bi[k] = (int)(t & (uint32_t)(c.nb[k] - 1));
        t >>= pa.lnb[k];
        c.B[k] = 1 << pa.lb[k];
        c.x0[k] = c.bi[k] << pa.lb[k];
    }
    return c;
}

// Stages the box of cells x0[k] + lo + t, t < T[k] (wrapped), of an N^R mesh
// into tile[tx + T0 (ty + T1 tz)], with NT threads and at most MAXE cells.
// The tile is walked in steps of NT cells, carrying (tx, ty, tz): all
// offsets first, then all loads in flight together, then the LDS stores (a
// load-store pair per cell would wait out the memory latency once per cell).
// The caller synchronises before reading the tile.
template <int NT, int MAXE, int R>
__device__ __forceinline__ void stage_box(float* tile, const float* __restrict__ m, int N, const int* x0, int lo,
                                          int T0, int T1, int T2, int tid) {
    constexpr int NL = (MAXE + NT - 1) / NT;
    const int msk = N - 1, total = T0 * T1 * T2;
    uint32_t off[NL];
    int tx = tid % T0, r = tid / T0;
    int ty = r % T1, tz = r / T1;
    const int dtx = NT % T0, dr = NT / T0;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        const uint32_t gx = (uint32_t)((x0[0] + lo + tx) & msk), gy = (uint32_t)((x0[1] + lo + ty) & msk);
        const uint32_t gz = (R == 3) ? (uint32_t)((x0[2] + lo + tz) & msk) : 0u;
        off[l] = gx + (uint32_t)N * (gy + (uint32_t)N * gz);
        tx += dtx;
        int carry = dr;
        if (tx >= T0) {
            tx -= T0;
            ++carry;
        }
        ty += carry;
        while (ty >= T1) {
            ty -= T1;
            ++tz;
        }
    }
    float v[NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) v[l] = (tid + l * NT < total) ? m[off[l]] : 0.f;
#pragma unroll
    for (int l = 0; l < NL; ++l)
        if (tid + l * NT < total) tile[tid + l * NT] = v[l];
}

// One dword per active lane from src straight into LDS at byte address
// lds_base + 4 * lane (global_load_lds_dword; lds_base must be the same for
// the whole wave, lanes that are switched off write nothing).  The load writes no register.  It is issued from inline assembly so
// that the compiler does not make later LDS reads wait for it (it cannot tell
// which LDS it writes); the caller waits with __builtin_amdgcn_s_waitcnt
// (WAIT_VMCNT0) and then synchronises before the data is read.  The LDS
// address comes from M0, which needs one wait state after the SALU write
// before an LDS-DMA load reads it (the s_nop).
__device__ __forceinline__ void lds_dma_dword(const float* src, uint32_t lds_base) {
    const uint32_t a = __builtin_amdgcn_readfirstlane(lds_base);
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dword %0, off" : : "v"(src), "s"(a) : "memory", "m0");
}
// s_waitcnt vmcnt(0) with the other counters left alone (gfx9 encoding:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
static constexpr int WAIT_VMCNT0 = 0x0F70;

// The sorted-particle ranges whose CIC clouds reach a brick: the whole brick
// itself, and of each lower neighbour only the face classes towards it.
struct BrickRanges {
    uint32_t rs0[64], rlen[64], roff[65];
    int n;
};

// Lane `lane` of one wave resolves one (neighbour, class) pair of brick bc:
// whether it contributes (ok) and its sorted range [s0, s1).
template <int R>
__device__ __forceinline__ void brick_range_lane(const uint32_t* __restrict__ bstart, const BrickCoord& bc,
                                                 const PosArgs& pa, int lane, bool& ok, uint32_t& s0,
                                                 uint32_t& s1) {
    const int nbr = lane >> R, c = lane & ((1 << R) - 1);
    const int d[3] = {nbr & 1, (nbr >> 1) & 1, (nbr >> 2) & 1};
    ok = nbr < (1 << R);
#pragma unroll
    for (int k = 0; k < 3; ++k) ok &= !(d[k] && bc.nb[k] == 1);  // that brick would be this one
    ok &= (nbr == 0) ? (c == 0) : ((c & nbr) == nbr);
    s0 = s1 = 0;
    if (ok) {
        uint32_t nb = 0;
        int sh = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            nb |= (uint32_t)((bc.bi[k] - d[k]) & (bc.nb[k] - 1)) << sh;
            sh += pa.lnb[k];
        }
        const uint32_t g0 = (nb << R) + (nbr == 0 ? 0 : c);
        s0 = bstart[g0];
        s1 = bstart[nbr == 0 ? g0 + (1u << R) : g0 + 1];
    }
}
// One whole wave compacts its lanes' non-empty ranges into r (a ballot for
// the slots, a shuffle scan for the offsets).  No barrier: other waves read
// r after the caller's next one.
__device__ __forceinline__ void brick_ranges_store(BrickRanges& r, bool ok, uint32_t s0, uint32_t s1, int lane) {
    const bool have = ok && s1 > s0;
    const uint64_t m = __ballot(have);
    const int slot = __popcll(m & ((1ull << lane) - 1ull));
    const uint32_t len = have ? s1 - s0 : 0u;
    uint32_t incl = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
    }
    if (have) {
        r.rs0[slot] = s0;
        r.rlen[slot] = len;
        r.roff[slot] = incl - len;
    }
    if (lane == 63) {
        r.n = (int)__popcll(m);
        r.roff[__popcll(m)] = incl;
    }
}

// Fills `r` (in LDS) for brick bc; called by the whole workgroup (the first
// wave resolves every (neighbour, class) pair at once).  Ends with a barrier.
template <int R>
__device__ __forceinline__ void brick_ranges(BrickRanges& r, const uint32_t* __restrict__ bstart,
                                             const BrickCoord& bc, const PosArgs& pa, int tid) {
    if (tid < 64) {
        bool ok;
        uint32_t s0, s1;
        brick_range_lane<R>(bstart, bc, pa, tid, ok, s0, s1);
        brick_ranges_store(r, ok, s0, s1, tid);
    }
    __syncthreads();
}

// Bookkeeping of a new FX sort key at storage position s.  Full sort
// (prevkeys == nullptr): pids[s] = s for the radix sort.  Incremental re-sort:
// a particle whose key differs from the key sorted at s last time (a mover)
// is counted into its new group (in_count) and out of its old one (out_count).
__device__ __forceinline__ void fx_note_key(uint32_t key, long s, const uint32_t* __restrict__ prevkeys,
                                            uint32_t* __restrict__ pids, uint32_t* __restrict__ in_count,
                                            uint32_t* __restrict__ out_count) {
    if (prevkeys) {
        const uint32_t pk = prevkeys[s];
        if (key != pk) {
            atomicAdd(&in_count[key], 1u);
            atomicAdd(&out_count[pk], 1u);
        }
    } else {
        pids[s] = (uint32_t)s;
    }
}

// Sorted position of the t-th particle of the concatenated ranges.
__device__ __forceinline__ uint32_t brick_range_pos(const BrickRanges& r, uint32_t t) {
    int q = 0;
    while (r.roff[q + 1] <= t) ++q;
    return r.rs0[q] + (t - r.roff[q]);
}

// Host side: launch-grid size, and a dispatch on the number of dimensions
// that binds the compile-time RR for the kernel template.
static inline unsigned nblk(long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

#define CPM_R_DISPATCH(R, ...)                                             \
    do {                                                                   \
        if ((R) == 3) { constexpr int RR = 3; __VA_ARGS__; }               \
        else if ((R) == 2) { constexpr int RR = 2; __VA_ARGS__; }          \
        else throw Error(-2, "only 2-D and 3-D runs are supported");      \
    } while (0)

}  // namespace cpm
