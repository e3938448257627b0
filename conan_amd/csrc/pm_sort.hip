// pm_sort.hip -- stable LSD radix sort of (cell key, particle id) pairs and a
// device-wide exclusive scan, both deterministic (no floating point, no
// order-dependent results).
//
// The deposit needs particles grouped by cell with ascending particle id
// inside each cell; sorting (key, id) pairs that start in id order with a
// stable LSD sort gives exactly that.
#include "pm_sort.hpp"

namespace cpm {

static constexpr int SORT_NT = 256;
static constexpr int SORT_ITEMS = 16;
static constexpr int SORT_TILE = SORT_NT * SORT_ITEMS;
static constexpr int SORT_WAVES = SORT_NT / 64;
static constexpr int MAX_RADIX = 512;
#ifndef CPM_SORT_BITS
#define CPM_SORT_BITS 8  // most bits per pass (27-bit keys: 4 passes of 7)
#endif

__global__ __launch_bounds__(SORT_NT) void k_rs_hist(const uint32_t* __restrict__ keys, long n, int shift,
                                                     int dbits, uint32_t* __restrict__ counts, int nblocks) {
    __shared__ uint32_t hist[MAX_RADIX];
    const int radix = 1 << dbits;
    const uint32_t dmask = (uint32_t)radix - 1;
    for (int d = threadIdx.x; d < radix; d += SORT_NT) hist[d] = 0;
    __syncthreads();
    const long tile0 = (long)blockIdx.x * SORT_TILE;
#pragma unroll 4
    for (int r = 0; r < SORT_ITEMS; ++r) {
        long e = tile0 + (long)r * SORT_NT + threadIdx.x;
        if (e < n) atomicAdd(&hist[(keys[e] >> shift) & dmask], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < radix; d += SORT_NT) counts[(long)d * nblocks + blockIdx.x] = hist[d];
}

// Each tile is walked in index order (round r, then thread); the rank of an
// element among equal digits = count in earlier rounds + count in earlier
// waves of this round + count in earlier lanes of its wave.  Elements are
// first placed in LDS in (digit, rank) order and then written out so that
// consecutive threads write consecutive addresses of one digit's run.
__global__ __launch_bounds__(SORT_NT) void k_rs_scatter(const uint32_t* __restrict__ kin,
                                                        const uint32_t* __restrict__ vin,
                                                        uint32_t* __restrict__ kout,
                                                        uint32_t* __restrict__ vout, long n, int shift,
                                                        int dbits, const uint32_t* __restrict__ offsets,
                                                        int nblocks) {
    __shared__ uint32_t run[MAX_RADIX];   // next local position per digit
    __shared__ uint32_t loff[MAX_RADIX];  // local start of each digit's run
    __shared__ uint32_t goff[MAX_RADIX];  // global start of each digit's run
    __shared__ uint32_t wcnt[SORT_WAVES][MAX_RADIX];
    __shared__ uint32_t lk[SORT_TILE], lv[SORT_TILE];
    const int radix = 1 << dbits;
    const uint32_t dmask = (uint32_t)radix - 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long tile0 = (long)blockIdx.x * SORT_TILE;
    const int tn = (int)((n - tile0) < SORT_TILE ? (n - tile0) : SORT_TILE);
    // local histogram of the tile -> local offsets
    for (int d = tid; d < radix; d += SORT_NT) {
        run[d] = 0;
        goff[d] = offsets[(long)d * nblocks + blockIdx.x];
#pragma unroll
        for (int w = 0; w < SORT_WAVES; ++w) wcnt[w][d] = 0;
    }
    __syncthreads();
    uint32_t kr[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r) {
        const int le = r * SORT_NT + tid;
        kr[r] = le < tn ? kin[tile0 + le] : 0u;
        if (le < tn) atomicAdd(&run[(kr[r] >> shift) & dmask], 1u);
    }
    __syncthreads();
    {
        // exclusive scan of the radix counts: thread t owns digits 2t, 2t+1
        uint32_t* sc = wcnt[0];
        const int d0 = 2 * tid;
        const uint32_t c0 = d0 < radix ? run[d0] : 0u, c1 = d0 + 1 < radix ? run[d0 + 1] : 0u;
        sc[tid] = c0 + c1;
        __syncthreads();
        for (int off = 1; off < SORT_NT; off <<= 1) {
            const uint32_t t = tid >= off ? sc[tid - off] : 0u;
            __syncthreads();
            sc[tid] += t;
            __syncthreads();
        }
        const uint32_t excl = sc[tid] - (c0 + c1);
        __syncthreads();
        if (d0 < radix) loff[d0] = run[d0] = excl;
        if (d0 + 1 < radix) loff[d0 + 1] = run[d0 + 1] = excl + c0;
        sc[tid] = 0;
        if (SORT_NT + tid < MAX_RADIX) sc[SORT_NT + tid] = 0;
    }
    __syncthreads();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < SORT_ITEMS; ++r) {
        const int le = r * SORT_NT + tid;
        const bool valid = le < tn;
        const uint32_t key = kr[r];
        const uint32_t d = (key >> shift) & dmask;
        uint64_t mask = __ballot(valid);
        for (int b = 0; b < dbits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            mask &= bit ? bb : ~bb;
        }
        const int rank = __popcll(mask & below);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d];
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            pos += (uint32_t)rank;
            lk[pos] = key;
            lv[pos] = vin[tile0 + le];
        }
        __syncthreads();
        for (int dd = tid; dd < radix; dd += SORT_NT) {
            uint32_t s = 0;
#pragma unroll
            for (int w = 0; w < SORT_WAVES; ++w) {
                s += wcnt[w][dd];
                wcnt[w][dd] = 0;
            }
            run[dd] += s;
        }
        __syncthreads();
    }
    for (int p = tid; p < tn; p += SORT_NT) {
        const uint32_t key = lk[p];
        const uint32_t d = (key >> shift) & dmask;
        const uint32_t pos = goff[d] + ((uint32_t)p - loff[d]);
        kout[pos] = key;
        vout[pos] = lv[p];
    }
}

// ------------------------------------------------------------ scan
static constexpr int SCAN_NT = 256;
static constexpr int SCAN_ITEMS = 16;
static constexpr int SCAN_TILE = SCAN_NT * SCAN_ITEMS;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    // sh: SCAN_NT entries
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_NT; off <<= 1) {
        uint32_t t = tid >= off ? sh[tid - off] : 0u;
        __syncthreads();
        sh[tid] += t;
        __syncthreads();
    }
    uint32_t incl = sh[tid];
    *total = sh[SCAN_NT - 1];
    __syncthreads();
    return incl - v;
}

// Exclusive scan of each tile; tile totals to sums (if non-null).
__global__ __launch_bounds__(SCAN_NT) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ out, long n,
                                                        uint32_t* __restrict__ sums) {
    __shared__ uint32_t sh[SCAN_NT];
    const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t loc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        loc += v[i];
    }
    uint32_t tot;
    uint32_t pre = block_exclusive_scan(loc, sh, &tot);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = pre;
        pre += v[i];
    }
    if (sums && threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_add(uint32_t* __restrict__ out, long n,
                                                      const uint32_t* __restrict__ sums) {
    const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    const uint32_t add = sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i < n) out[base + i] += add;
}

// Two-pass form for up to SCAN_FLAT tiles: the tile totals, then each tile
// scanned from the sum of the totals before it (read by the whole workgroup
// from the small sums array), so the scan of the totals and the pass that
// adds them need no launches of their own.
static constexpr long SCAN_FLAT = 4096;
__global__ __launch_bounds__(SCAN_NT) void k_scan_tile_sums(const uint32_t* __restrict__ in, long n,
                                                            uint32_t* __restrict__ sums) {
    __shared__ uint32_t sh[SCAN_NT];
    const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    uint32_t loc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) loc += (base + i < n) ? in[base + i] : 0u;
    uint32_t tot;
    (void)block_exclusive_scan(loc, sh, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
__global__ __launch_bounds__(SCAN_NT) void k_scan_tiles_from(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out, long n,
                                                             const uint32_t* __restrict__ sums) {
    __shared__ uint32_t sh[SCAN_NT];
    uint32_t x = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += SCAN_NT) x += sums[j];
    uint32_t off;
    (void)block_exclusive_scan(x, sh, &off);
    __syncthreads();  // sh is reused
    const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t loc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        loc += v[i];
    }
    uint32_t tot;
    uint32_t pre = off + block_exclusive_scan(loc, sh, &tot);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = pre;
        pre += v[i];
    }
}

size_t scan_scratch_elems(long n) {
    size_t tot = 0;
    long m = n;
    while (m > SCAN_TILE) {
        m = (m + SCAN_TILE - 1) / SCAN_TILE;
        tot += (size_t)m + (size_t)m;  // sums + scanned sums
    }
    return tot + 1;
}

void exclusive_scan(const uint32_t* in, uint32_t* out, long n, uint32_t* scratch, hipStream_t st) {
    const long nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nt <= 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_NT), 0, st, in, out, n, (uint32_t*)nullptr);
        CPM_LAUNCH_CHECK();
        return;
    }
    uint32_t* sums = scratch;
    if (nt <= SCAN_FLAT) {
        hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)nt), dim3(SCAN_NT), 0, st, in, n, sums);
        CPM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_scan_tiles_from, dim3((unsigned)nt), dim3(SCAN_NT), 0, st, in, out, n,
                           (const uint32_t*)sums);
        CPM_LAUNCH_CHECK();
        return;
    }
    uint32_t* ssc = scratch + nt;
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(SCAN_NT), 0, st, in, out, n, sums);
    CPM_LAUNCH_CHECK();
    exclusive_scan(sums, ssc, nt, scratch + 2 * nt, st);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nt), dim3(SCAN_NT), 0, st, out, n, (const uint32_t*)ssc);
    CPM_LAUNCH_CHECK();
}

// ------------------------------------------------------------ sort driver
RadixSorter::RadixSorter(long n_, int keybits_) : n(n_), keybits(keybits_) {
    passes = (keybits + CPM_SORT_BITS - 1) / CPM_SORT_BITS;
    if (passes < 1) passes = 1;
    dbits = (keybits + passes - 1) / passes;
    nblocks = (int)((n + SORT_TILE - 1) / SORT_TILE);
    if (nblocks < 1) nblocks = 1;
    ncounts = (long)(1 << dbits) * nblocks;
    void* c = nullptr;
    device_alloc(&c, sizeof(uint32_t) * (ncounts + 1));
    counts = static_cast<uint32_t*>(c);
    try {
        device_alloc(&c, sizeof(uint32_t) * scan_scratch_elems(ncounts));
    } catch (...) {
        device_free(counts, sizeof(uint32_t) * (ncounts + 1));
        throw;
    }
    scratch = static_cast<uint32_t*>(c);
}

RadixSorter::~RadixSorter() {
    device_free(counts, sizeof(uint32_t) * (ncounts + 1));
    device_free(scratch, sizeof(uint32_t) * scan_scratch_elems(ncounts));
}

int RadixSorter::sort(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, hipStream_t st) {
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * dbits;
        int db = dbits;
        if (shift + db > keybits) db = keybits - shift;
        if (db <= 0) break;
        hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(SORT_NT), 0, st, (const uint32_t*)ki, n, shift, db,
                           counts, nblocks);
        CPM_LAUNCH_CHECK();
        const long nc = (long)(1 << db) * nblocks;
        exclusive_scan(counts, counts, nc, scratch, st);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblocks), dim3(SORT_NT), 0, st, (const uint32_t*)ki,
                           (const uint32_t*)vi, ko, vo, n, shift, db, (const uint32_t*)counts, nblocks);
        CPM_LAUNCH_CHECK();
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    return ki == k0 ? 0 : 1;
}

}  // namespace cpm

// ------------------------------------------------------------ incremental re-sort
// Keys that were sorted one step ago change for few elements (particles that
// crossed into another brick or face class).  The storage order is the order
// of the last sort, whose group starts bprev give each position its previous
// key; the stayers of group g keep their order at the front of the group and
// the movers follow them.
namespace cpm {

__global__ __launch_bounds__(256) void k_resort_counts(const uint32_t* __restrict__ bprev,
                                                       uint32_t* __restrict__ in_count,
                                                       uint32_t* __restrict__ out_count,
                                                       uint32_t* __restrict__ fill, long ngroups,
                                                       uint32_t* __restrict__ stay, uint32_t* __restrict__ tot) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g > ngroups) return;
    if (g == ngroups) {
        tot[g] = 0;
        return;
    }
    const uint32_t in = in_count[g], out = out_count[g];
    const uint32_t s = (bprev[g + 1] - bprev[g]) - out;
    stay[g] = s;
    tot[g] = s + in;
    // the counts are consumed: cleared for the next gather, and the
    // placement's slot counters cleared for this sort
    in_count[g] = 0u;
    out_count[g] = 0u;
    fill[g] = 0u;
}

// One wave per brick (four bricks per workgroup): the brick's previous range
// (its G groups, one after the other) is walked in passes of E x 64
// elements, all loads of a pass in flight together; a scan of the stayer
// flags (wave ballots, no barriers, no LDS) gives each stayer its rank among
// the stayers of its group (they keep their order); a mover takes the next
// free slot after the stayers of its new group.  The previous key of
// position s is the group of bprev that holds s (the brick's G group starts,
// in scalar registers), so the previous keys are not read.  Lane c <= G
// reads the previous start of group g0 + c, and lanes c < G its new start and
// its stayers; nb[c] = new start of group c minus the stayers of the brick's
// groups before c, so a stayer's position is nb[c] plus its rank among the
// brick's stayers.  E: elements per lane in flight per pass.
static constexpr int RS_AGG = 8;  // movers of a 64-element slice above which they reserve slots by group
template <int E, bool AGG>
__global__ __launch_bounds__(256) void k_resort_place(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ bprev,
                                                           const uint32_t* __restrict__ bnew,
                                                           const uint32_t* __restrict__ stay,
                                                           uint32_t* __restrict__ fill, long nbricks, int G,
                                                           uint32_t* __restrict__ perm) {
    const int lane = threadIdx.x & 63;
    const long b = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= nbricks) return;  // the whole wave
    const uint32_t g0 = (uint32_t)b * (uint32_t)G;
    const uint32_t vo = lane <= G ? bprev[g0 + lane] : 0u;
    const uint32_t vn = lane < G ? bnew[g0 + lane] : 0u;
    const uint32_t vs = lane < G ? stay[g0 + lane] : 0u;
    uint32_t ost[8], nb[8];
    uint32_t before = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        ost[c] = __builtin_amdgcn_readlane(vo, c);
        nb[c] = __builtin_amdgcn_readlane(vn, c) - before;
        before += __builtin_amdgcn_readlane(vs, c);
    }
    const uint32_t r0 = ost[0], r1 = __builtin_amdgcn_readlane(vo, G);
    const uint64_t below = (1ull << lane) - 1ull;  // the lanes before this one
    uint32_t carry = 0;                            // stayers of the brick before the pass
    for (uint32_t base = r0; base < r1; base += E * 64) {
        uint32_t key[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t s = base + e * 64 + lane;
            key[e] = s < r1 ? keys[s] : 0u;
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t s = base + e * 64 + lane;
            uint32_t c = 0;  // the group that held s
#pragma unroll
            for (int k = 1; k < 8; ++k) c += (k < G && s >= ost[k]) ? 1u : 0u;
            const bool in = s < r1, st = in && key[e] == g0 + c;
            const uint64_t m = __ballot(st);
            const uint32_t rank = carry + (uint32_t)__popcll(m & below);
            carry += (uint32_t)__popcll(m);
            uint32_t n = nb[0];
#pragma unroll
            for (int k = 1; k < 8; ++k) n = (c == (uint32_t)k) ? nb[k] : n;
            const bool mv = in && !st;
            uint64_t pend = __ballot(mv);
            if (!AGG || __popcll(pend) <= RS_AGG) {
                // a few movers: each reserves its own slot
                if (in) perm[st ? n + rank : bnew[key[e]] + stay[key[e]] + atomicAdd(&fill[key[e]], 1u)] = s;
            } else {
                // many (dense bricks, large steps): by new group, the first pending lane's
                // group is taken, its movers ranked among themselves, and one
                // lane (the leader) reserves their slots with one atomic, so
                // the device-scope atomics number the slice's groups, not its
                // movers
                int ld = lane;
                uint32_t rk = 0u, cnt = 0u;
                while (pend) {
                    const int l = (int)__builtin_ctzll(pend);
                    const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)key[e], l);
                    const uint64_t same = __ballot(mv && key[e] == k);
                    if (mv && key[e] == k) {
                        ld = l;
                        rk = (uint32_t)__popcll(same & below);
                        cnt = lane == l ? (uint32_t)__popcll(same) : 0u;
                    }
                    pend &= ~same;
                }
                uint32_t lb = 0u;
                if (cnt) lb = bnew[key[e]] + stay[key[e]] + atomicAdd(&fill[key[e]], cnt);
                const uint32_t b = (uint32_t)__shfl((int)lb, ld);
                if (in) perm[st ? n + rank : b + rk] = s;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* __restrict__ p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

void resort_counts(const uint32_t* bprev, uint32_t* in_count, uint32_t* out_count, uint32_t* fill, long ngroups,
                   uint32_t* stay, uint32_t* tot, hipStream_t st) {
    hipLaunchKernelGGL(k_resort_counts, dim3((unsigned)((ngroups + 256) / 256)), dim3(256), 0, st, bprev, in_count,
                       out_count, fill, ngroups, stay, tot);
    CPM_LAUNCH_CHECK();
}

void resort_place(const uint32_t* keys, const uint32_t* bprev, const uint32_t* bnew, const uint32_t* stay,
                  uint32_t* fill, long nbricks, int G, long np, uint32_t* perm, hipStream_t st) {
    // passes of 128 or 512 elements: about one pass per brick on average;
    // bricks of thousands of particles (whose large steps move many of them)
    // reserve their movers' slots by group
    const unsigned nb = (unsigned)((nbricks + 3) / 4);
    if (np / nbricks >= 2048)
        hipLaunchKernelGGL((k_resort_place<8, true>), dim3(nb), dim3(256), 0, st, keys, bprev, bnew, stay, fill,
                           nbricks, G, perm);
    else if (np / nbricks >= 256)
        hipLaunchKernelGGL((k_resort_place<8, false>), dim3(nb), dim3(256), 0, st, keys, bprev, bnew, stay, fill,
                           nbricks, G, perm);
    else
        hipLaunchKernelGGL((k_resort_place<2, false>), dim3(nb), dim3(256), 0, st, keys, bprev, bnew, stay, fill,
                           nbricks, G, perm);
    CPM_LAUNCH_CHECK();
}

// skeys[s] = the group of bstart that holds position s, one workgroup per
// brick of G groups.
__global__ __launch_bounds__(256) void k_group_keys(const uint32_t* __restrict__ bstart, int G,
                                                    uint32_t* __restrict__ skeys) {
    __shared__ uint32_t gs[9];
    const uint32_t g0 = blockIdx.x * (uint32_t)G;
    if (threadIdx.x <= (unsigned)G) gs[threadIdx.x] = bstart[g0 + threadIdx.x];
    __syncthreads();
    for (uint32_t s = gs[0] + threadIdx.x; s < gs[G]; s += 256) {
        uint32_t c = 0;
        for (int k = 1; k < G; ++k) c += (s >= gs[k]) ? 1u : 0u;
        skeys[s] = g0 + c;
    }
}

void group_keys(const uint32_t* bstart, long nbricks, int G, uint32_t* skeys, hipStream_t st) {
    hipLaunchKernelGGL(k_group_keys, dim3((unsigned)nbricks), dim3(256), 0, st, bstart, G, skeys);
    CPM_LAUNCH_CHECK();
}

void zero_u32(uint32_t* p, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    CPM_LAUNCH_CHECK();
}

}  // namespace cpm
