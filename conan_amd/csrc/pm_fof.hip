// pm_fof.hip -- friends-of-friends groups on the GPU (src/fof/fof.cc:157-214)
// and their C ABI (pm_fof_* in include/conanpm.h).
//
// The reference tags particles by flood fills started in index order
// (fof.cc:175-214), so group t is the connected component, under the
// Proximity predicate (fof.cc:93-118), whose lowest particle index is the
// t-th smallest.  Here the components come from a lock-free union-find and
// are numbered the same way afterwards:
//
//   cells   cell of every particle on a periodic grid of cells no smaller
//           than the linking length; key = linear cell index mod 2^tbits
//           (a table about twice the particle count; cells that share a key
//           only add candidates, the predicate rejects them); counting sort
//           by key: per-key counts (atomics), scan -> bucket starts
//   place   particles into key order: positions (AoS) and original indices
//   link    each particle tests the particles of the half shell of its
//           neighbour cells (13 of 26 plus the later ones of its own cell, so
//           each pair is tested once; one thread per particle and row of
//           cells along x) with the reference's predicate in IEEE
//           double arithmetic (no contraction), and unites the linked pairs
//           by their original indices: roots are hooked under smaller roots
//           by compare-and-swap, so a component's root is its lowest index
//   roots   every particle finds its root; flag[i] = particle i is a root
//   number  the scan of the flags numbers the components in the order of
//           their lowest indices, the reference's order
//
// Positions as the snapshots hold them (unwrapped); the cell of a particle is
// taken from its position folded into [0, L).  The result does not depend on
// the order in which the atomics land: it is the same on every run.
#include "pm_brick.hpp"
#include "pm_impl.hpp"
#include "pm_sort.hpp"

#include <memory>

namespace cpm {

struct FofGrid {
    double L, invL, halfL, rsqr;
    double inv_cs[3];
    int nc[3];
    uint32_t kmask;  // table size - 1
    // rows of the half shell: row r is (dy, dz) = ((rows >> 3r) & 3) - 1,
    // (rows >> (3r + 2)) & 1; row 0 is the particle's own row
    uint32_t rows;
    int nrows;
};

static constexpr int FOF_NT = 256;

// Cell of a position on the grid (each coordinate folded into [0, L) first;
// clamped so that no input can index outside the grid).
template <int R>
__device__ __forceinline__ void fof_cell(const double* x, const FofGrid& g, int* c) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= R) {
            c[k] = 0;
            continue;
        }
        const double q = floor(x[k] * g.invL);
        double w = x[k] - q * g.L;
        w = fmin(fmax(w * g.inv_cs[k], 0.0), (double)g.nc[k]);
        int ck = (int)w;
        c[k] = ck >= g.nc[k] ? ck - g.nc[k] : ck;
    }
}

__device__ __forceinline__ uint32_t fof_key(int x, int y, int z, const FofGrid& g) {
    const uint64_t lin = (uint64_t)x + (uint64_t)g.nc[0] * ((uint64_t)y + (uint64_t)g.nc[1] * (uint64_t)z);
    return (uint32_t)lin & g.kmask;
}

__device__ __forceinline__ int fof_wrap(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

// Proximity (fof.cc:93-118): b = p - a per axis, wrapped once by L, and the
// squares summed from 0 in axis order (mvector.hh:128-133).  The library is
// built without FMA contraction, so this rounds like the reference.
template <int R>
__device__ __forceinline__ bool fof_near(const double* a, const double* p, const FofGrid& g) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double b = p[k] - a[k];
        const double d = b > g.halfL ? b - g.L : (b < -g.halfL ? b + g.L : b);
        v = v + d * d;
    }
    return v <= g.rsqr;
}

// Parent links are read and path-split with relaxed atomics of workgroup
// scope: plain cached loads and stores, so a thread may see an older link.
// That is safe: a link only ever changes from a root to a smaller position or
// from a non-root to one of its ancestors, so every value ever read is an
// ancestor of the same set, and no store of a path split can undo a hook
// (only roots are hooked, and a path split writes only to non-roots).  The
// hooks themselves are compare-and-swaps of device scope, whose result is
// current.
__device__ __forceinline__ uint32_t uf_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Root of x, splitting the path on the way (each node visited is linked to
// its grandparent).  The chain strictly decreases, so the walk ends.
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
    uint32_t p = uf_load(parent + x);
    while (p != x) {
        const uint32_t gp = uf_load(parent + p);
        if (gp != p) __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        x = p;
        p = gp;
    }
    return x;
}

// Joins the sets of a and b and returns a root (as of the join) of the
// joined set: the larger root is hooked under the smaller one, so a set's
// root is its smallest position.  A failed compare-and-swap returns the
// current link of the would-be root, and the walk continues from there: b
// strictly decreases on every retry, so the loop ends.
__device__ __forceinline__ uint32_t uf_unite(uint32_t* parent, uint32_t a, uint32_t b) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    while (a != b) {
        if (a > b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        uint32_t expect = b;
        if (__hip_atomic_compare_exchange_strong(parent + b, &expect, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return a;
        a = uf_find(parent, a);
        b = uf_find(parent, expect);
    }
    return a;
}

// Cell key of every particle and its slot among the particles of its key.
// Particles that follow each other in the input often share a cell (input
// order is mostly the Lagrangian grid's), so lanes with the key of the lane
// before them join its run and the run's first lane takes the slots of the
// whole run with one atomic.  Also starts the union-find: parent[i] = i.
template <int R>
__global__ __launch_bounds__(FOF_NT) void k_fof_cells(const double* __restrict__ pos, uint32_t n, FofGrid g,
                                                      uint32_t* __restrict__ key, uint32_t* __restrict__ slot,
                                                      uint32_t* __restrict__ cnt, uint32_t* __restrict__ parent) {
    const uint32_t i = blockIdx.x * FOF_NT + threadIdx.x;
    const int lane = threadIdx.x & 63;
    uint32_t kk = 0xFFFFFFFFu;
    if (i < n) {
        double x[3];
#pragma unroll
        for (int k = 0; k < R; ++k) x[k] = pos[(size_t)i * R + k];
        int c[3];
        fof_cell<R>(x, g, c);
        kk = fof_key(c[0], c[1], c[2], g);
    }
    const uint32_t prev = __shfl_up(kk, 1);
    const bool head = lane == 0 || prev != kk;
    const uint64_t heads = __ballot(head);
    // run of this lane: from its head (highest head at or below the lane) to the next head
    const int h = 63 - __clzll(heads & ((2ull << lane) - 1ull));
    const uint64_t above = heads & ~((2ull << lane) - 1ull);
    const int nexth = above ? __ffsll((unsigned long long)above) - 1 : 64;
    uint32_t base = 0;
    if (head && i < n) base = atomicAdd(&cnt[kk], (uint32_t)(nexth - lane));
    base = __shfl(base, h);
    if (i < n) {
        key[i] = kk;
        slot[i] = base + (uint32_t)(lane - h);
        parent[i] = i;
    }
}

template <int R>
__global__ __launch_bounds__(FOF_NT) void k_fof_place(const double* __restrict__ pos, uint32_t n,
                                                      const uint32_t* __restrict__ key,
                                                      const uint32_t* __restrict__ slot,
                                                      const uint32_t* __restrict__ bstart, double* __restrict__ spos,
                                                      uint32_t* __restrict__ sidx) {
    const uint32_t i = blockIdx.x * FOF_NT + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = bstart[key[i]] + slot[i];
#pragma unroll
    for (int k = 0; k < R; ++k) spos[(size_t)s * R + k] = pos[(size_t)i * R + k];
    sidx[s] = i;
}

// One thread per (particle, row of neighbour cells): blockIdx.y picks the row
// (dy, dz) of the half shell, the thread scans the row's cells along x.  Row
// 0 is the particle's own row: the later particles of its own cell and the
// +x cell; the other rows (dz > 0, or dz = 0 and dy > 0) take all three cells.
// A thread tests at most FOF_CAP candidates of one range itself; the rest of
// a longer range (the cells of a dense halo core hold hundreds of particles)
// goes to a queue that whole waves work through (k_fof_link_queue), so that no
// single thread runs a long serial loop.  Should the queue be full, the
// thread scans the whole range itself.
static constexpr uint32_t FOF_CAP = 32;
static constexpr uint32_t FOF_UNROLL = 2;  // candidates loaded together per round (4 measured slower)
// Counters and the queue are spread over many addresses: atomics on a single
// address from every wave of the grid would serialise.
static constexpr int FOF_NSLOT = 1024;  // counter slots (tests, links)
static constexpr int FOF_NQ = 256;      // sub-queues, each with its own fill count

// Wave sums of a thread's test and link counts, added to one counter slot.
__device__ __forceinline__ void fof_count(unsigned long long* stats, uint32_t ntest, uint32_t nlink, uint32_t slot) {
    unsigned long long t = ntest, l = nlink;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        t += __shfl_down(t, off);
        l += __shfl_down(l, off);
    }
    if ((threadIdx.x & 63) == 0 && (t | l)) {
        slot &= FOF_NSLOT - 1;
        atomicAdd(&stats[slot], t);
        atomicAdd(&stats[FOF_NSLOT + slot], l);
    }
}

template <int R>
__global__ __launch_bounds__(FOF_NT) void k_fof_link(const double* __restrict__ spos,
                                                     const uint32_t* __restrict__ sidx, uint32_t n,
                                                     const uint32_t* __restrict__ bstart, FofGrid g,
                                                     uint32_t* parent, unsigned long long* __restrict__ stats,
                                                     uint3* __restrict__ queue, uint32_t* __restrict__ qcount,
                                                     uint32_t qsub) {
    const uint32_t sub = (blockIdx.x + 97u * blockIdx.y) & (FOF_NQ - 1);
    const uint32_t s = blockIdx.x * FOF_NT + threadIdx.x;
    const uint32_t rc = g.rows >> (3 * blockIdx.y);
    const int dy = (int)(rc & 3u) - 1, dz = (int)((rc >> 2) & 1u);
    const bool own = blockIdx.y == 0;
    uint32_t ntest = 0, nlink = 0;
    if (s < n) {
        double a[3];
#pragma unroll
        for (int k = 0; k < R; ++k) a[k] = spos[(size_t)s * R + k];
        int c[3];
        fof_cell<R>(a, g, c);
        uint32_t rs = sidx[s];  // particle s, later an ancestor of it: where the next union starts
        auto scan = [&](uint32_t lo, uint32_t hi) {
            // queue slots for the lanes with long ranges: one atomic per wave
            const bool lng = hi - lo > FOF_CAP;
            const uint64_t m = __ballot(lng);
            if (lng) {
                const int lane = threadIdx.x & 63, lead = __ffsll((unsigned long long)m) - 1;
                uint32_t base = 0;
                if (lane == lead) base = atomicAdd(qcount + sub, (uint32_t)__popcll(m));
                base = __shfl(base, lead);
                const uint32_t q = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                if (q < qsub) {
                    queue[(size_t)sub * qsub + q] = make_uint3(s, lo + FOF_CAP, hi);
                    hi = lo + FOF_CAP;
                }
            }
            // FOF_UNROLL candidates per round: their positions in flight together
            uint32_t t = lo;
            for (; t + FOF_UNROLL <= hi; t += FOF_UNROLL) {
                double p[FOF_UNROLL][3];
#pragma unroll
                for (int u = 0; u < FOF_UNROLL; ++u)
#pragma unroll
                    for (int k = 0; k < R; ++k) p[u][k] = spos[(size_t)(t + u) * R + k];
                ntest += FOF_UNROLL;
                bool near[FOF_UNROLL];
#pragma unroll
                for (int u = 0; u < FOF_UNROLL; ++u) near[u] = fof_near<R>(a, p[u], g);
#pragma unroll
                for (int u = 0; u < FOF_UNROLL; ++u)
                    if (near[u]) {
                        ++nlink;
                        rs = uf_unite(parent, rs, sidx[t + u]);
                    }
            }
            for (; t < hi; ++t) {
                double p[3];
#pragma unroll
                for (int k = 0; k < R; ++k) p[k] = spos[(size_t)t * R + k];
                ++ntest;
                if (fof_near<R>(a, p, g)) {
                    ++nlink;
                    rs = uf_unite(parent, rs, sidx[t]);
                }
            }
        };
        const int cy = fof_wrap(c[1] + dy, g.nc[1]), cz = fof_wrap(c[2] + dz, g.nc[2]);
        const uint32_t k0 = fof_key(c[0], cy, cz, g);
        if (g.nc[0] < 3) {  // one cell along x
            scan(own ? s + 1 : bstart[k0], bstart[k0 + 1]);
        } else {
            const uint32_t kp = fof_key(fof_wrap(c[0] + 1, g.nc[0]), cy, cz, g);
            if (own) {
                if (kp == k0 + 1) {
                    scan(s + 1, bstart[kp + 1]);
                } else {
                    scan(s + 1, bstart[k0 + 1]);
                    scan(bstart[kp], bstart[kp + 1]);
                }
            } else {
                const uint32_t km = fof_key(fof_wrap(c[0] - 1, g.nc[0]), cy, cz, g);
                if (k0 == km + 1 && kp == k0 + 1) {  // the row's three cells are adjacent in the table
                    scan(bstart[km], bstart[kp + 1]);
                } else {
                    scan(bstart[km], bstart[km + 1]);
                    scan(bstart[k0], bstart[k0 + 1]);
                    scan(bstart[kp], bstart[kp + 1]);
                }
            }
        }
    }
    fof_count(stats, ntest, nlink, (blockIdx.x + gridDim.x * blockIdx.y) * (FOF_NT / 64) + (threadIdx.x >> 6));
}

// The queued range tails: one wave per item at a time, lane j testing
// candidates lo + j, lo + j + 64, ...  The items of the sub-queues are
// numbered through a prefix sum of their fill counts (in LDS).
template <int R>
__global__ __launch_bounds__(FOF_NT) void k_fof_link_queue(const double* __restrict__ spos,
                                                           const uint32_t* __restrict__ sidx,
                                                           const uint3* __restrict__ queue,
                                                           const uint32_t* __restrict__ qcount, uint32_t qsub,
                                                           FofGrid g, uint32_t* parent,
                                                           unsigned long long* __restrict__ stats) {
    static_assert(FOF_NQ == FOF_NT, "one sub-queue count per thread");
    __shared__ uint32_t qend[FOF_NQ];
    const int tid = threadIdx.x, lane = tid & 63;
    uint32_t v = min(qcount[tid], qsub), incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    qend[tid] = incl;  // inclusive within the wave
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < (tid >> 6); ++w) add += qend[w * 64 + 63];
    __syncthreads();
    qend[tid] = incl + add;  // inclusive over the sub-queues
    __syncthreads();
    const uint32_t nitems = qend[FOF_NQ - 1];
    const uint32_t wave = (blockIdx.x * FOF_NT + tid) >> 6, nwaves = (gridDim.x * FOF_NT) >> 6;
    uint32_t ntest = 0, nlink = 0;
    for (uint32_t it = wave; it < nitems; it += nwaves) {
        int lo = 0, hi = FOF_NQ - 1;  // first sub-queue whose end exceeds it
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (qend[mid] > it) hi = mid; else lo = mid + 1;
        }
        const uint32_t start = lo ? qend[lo - 1] : 0u;
        const uint3 q = queue[(size_t)lo * qsub + (it - start)];
        double a[3];
#pragma unroll
        for (int k = 0; k < R; ++k) a[k] = spos[(size_t)q.x * R + k];
        uint32_t rs = sidx[q.x];
        for (uint32_t t = q.y + lane; t < q.z; t += 64) {
            double p[3];
#pragma unroll
            for (int k = 0; k < R; ++k) p[k] = spos[(size_t)t * R + k];
            ++ntest;
            if (fof_near<R>(a, p, g)) {
                ++nlink;
                rs = uf_unite(parent, rs, sidx[t]);
            }
        }
    }
    fof_count(stats, ntest, nlink, blockIdx.x * (FOF_NT / 64) + (tid >> 6));
}

// After the links: root[i] of every particle (its group's lowest index), and
// flag[i] = 1 for the roots; flag[n] = 0 closes the array for the scan.
__global__ __launch_bounds__(FOF_NT) void k_fof_roots(uint32_t n, uint32_t* parent, uint32_t* __restrict__ root,
                                                      uint32_t* __restrict__ flag) {
    const uint32_t i = blockIdx.x * FOF_NT + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        flag[n] = 0;
        return;
    }
    const uint32_t r = uf_find(parent, i);
    root[i] = r;
    flag[i] = r == i ? 1u : 0u;
}

// group[i] = number of roots below i's root = rank of i's group in the
// order of the groups' lowest indices (the reference's tag order).
__global__ __launch_bounds__(FOF_NT) void k_fof_number(uint32_t n, const uint32_t* __restrict__ root,
                                                       const uint32_t* __restrict__ rank, long long* __restrict__ group) {
    const uint32_t i = blockIdx.x * FOF_NT + threadIdx.x;
    if (i < n) group[i] = (long long)rank[root[i]];
}

__global__ __launch_bounds__(FOF_NT) void k_fof_fill(uint32_t* p, long n, uint32_t v) {
    const long i = (long)blockIdx.x * FOF_NT + threadIdx.x;
    if (i < n) p[i] = v;
}

static void fof_fill(uint32_t* p, long n, uint32_t v, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fof_fill, dim3(nblk(n, FOF_NT)), dim3(FOF_NT), 0, st, p, n, v);
    CPM_LAUNCH_CHECK();
}

// Device buffers of one friends-of-friends run, grown on demand.
struct FofPlan {
    int R;
    double L, ll;
    FofGrid g{};
    int tbits = 0;
    long cap = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf<double> pos, spos;
    DevBuf<uint32_t> key, slot, sidx, parent, flag, cnt, scratch;
    DevBuf<long long> group;
    DevBuf<unsigned long long> stats;  // FOF_NSLOT test counts, then FOF_NSLOT link counts
    DevBuf<uint3> queue;                // FOF_NQ sub-queues of qsub items
    DevBuf<uint32_t> qcount;            // fill count of each sub-queue
    uint32_t qsub = 0;
    pm_fof_stats last{};

    FofPlan(int R_, double L_, double ll_) : R(R_), L(L_), ll(ll_) {
        CPM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        CPM_HIP(hipEventCreate(&ev0));
        CPM_HIP(hipEventCreate(&ev1));
        stats.alloc(2 * FOF_NSLOT);
        qcount.alloc(FOF_NQ);
        // cells a little wider than the linking length, so that rounding in
        // a cell index can never put a linked pair two cells apart; an axis
        // with fewer than 3 cells is one cell (every cell would be adjacent)
        for (int k = 0; k < 3; ++k) {
            double m = k < R ? std::floor(L / (ll * (1.0 + 1e-9))) : 1.0;
            if (!(m >= 3.0)) m = 1.0;
            if (m > 1048576.0) m = 1048576.0;
            g.nc[k] = (int)m;
            g.inv_cs[k] = (double)g.nc[k] / L;
        }
        const bool yfull = g.nc[1] >= 3, zfull = R == 3 && g.nc[2] >= 3;
        int rr[5][2] = {{0, 0}}, nr = 1;
        if (yfull) rr[nr][0] = 1, rr[nr++][1] = 0;
        if (zfull)
            for (int dy = yfull ? -1 : 0; dy <= (yfull ? 1 : 0); ++dy) rr[nr][0] = dy, rr[nr++][1] = 1;
        g.rows = 0;
        for (int r = 0; r < nr; ++r) g.rows |= (uint32_t)((rr[r][0] + 1) | (rr[r][1] << 2)) << (3 * r);
        g.nrows = nr;
        g.L = L;
        g.invL = 1.0 / L;
        g.halfL = L / 2;
        g.rsqr = ll * ll;
    }
    ~FofPlan() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (st) (void)hipStreamDestroy(st);
    }

    // Device buffers for n particles; the staging copies of host positions
    // and groups only when the caller passes host memory.
    void reserve(long n, bool device) {
        // table: about two keys per particle, no more than there are cells
        const double cells = (double)g.nc[0] * g.nc[1] * g.nc[2];
        int tb = ilog2(2 * std::max(n, 1L));
        while (tb > 1 && std::ldexp(1.0, tb - 1) >= cells) --tb;
        tbits = std::max(tb, 1);
        g.kmask = (uint32_t)((1ull << tbits) - 1);
        const long T = 1L << tbits;
        if (!device && pos.n < (size_t)R * n) {
            pos.alloc((size_t)R * n);
            group.alloc(n);
        }
        if (n > cap) {
            cap = n;
            spos.alloc((size_t)R * n);
            key.alloc(n);
            slot.alloc(n);
            sidx.alloc(n);
            parent.alloc(n);
            flag.alloc(n + 1);
            qsub = (uint32_t)((n + FOF_NQ - 1) / FOF_NQ);
            queue.alloc((size_t)FOF_NQ * qsub);
        }
        if (cnt.n < (size_t)T + 1) cnt.alloc(T + 1);
        const size_t sc = scan_scratch_elems(std::max(T + 1, n + 1));
        if (scratch.n < sc) scratch.alloc(sc);
    }

    // pos, group: host or device (current device) memory as `device` says.
    long long run(long n, const double* hpos, long long* hgroup, bool device) {
        if (n == 0) {
            last = pm_fof_stats{};
            fill_stats(0, 0.0);
            return 0;
        }
        reserve(n, device);
        const uint32_t un = (uint32_t)n;
        const double* dpos = hpos;
        if (!device) {
            CPM_HIP(hipMemcpyAsync(pos.p, hpos, sizeof(double) * R * n, hipMemcpyHostToDevice, st));
            dpos = pos.p;
        }
        long long* dgroup = device ? hgroup : group.p;
        const long T = 1L << tbits;
        CPM_HIP(hipEventRecord(ev0, st));
        fof_fill(cnt.p, T + 1, 0u, st);
        fof_fill((uint32_t*)stats.p, 4 * FOF_NSLOT, 0u, st);
        fof_fill(qcount.p, FOF_NQ, 0u, st);
        CPM_R_DISPATCH(R, hipLaunchKernelGGL(k_fof_cells<RR>, dim3(nblk(n, FOF_NT)), dim3(FOF_NT), 0, st, dpos,
                                             un, g, key.p, slot.p, cnt.p, parent.p));
        CPM_LAUNCH_CHECK();
        exclusive_scan(cnt.p, cnt.p, T + 1, scratch.p, st);
        CPM_R_DISPATCH(R, hipLaunchKernelGGL(k_fof_place<RR>, dim3(nblk(n, FOF_NT)), dim3(FOF_NT), 0, st, dpos,
                                             un, key.p, slot.p, cnt.p, spos.p, sidx.p));
        CPM_LAUNCH_CHECK();
        CPM_R_DISPATCH(R, hipLaunchKernelGGL(k_fof_link<RR>, dim3(nblk(n, FOF_NT), g.nrows), dim3(FOF_NT), 0, st, spos.p,
                                             sidx.p, un, cnt.p, g, parent.p, stats.p, queue.p, qcount.p, qsub));
        CPM_LAUNCH_CHECK();
        CPM_R_DISPATCH(R, hipLaunchKernelGGL(k_fof_link_queue<RR>, dim3((unsigned)(8 * device_cus())), dim3(FOF_NT), 0,
                                             st, spos.p, sidx.p, queue.p, qcount.p, qsub, g, parent.p, stats.p));
        CPM_LAUNCH_CHECK();
        // key -> root of every particle
        hipLaunchKernelGGL(k_fof_roots, dim3(nblk(n + 1, FOF_NT)), dim3(FOF_NT), 0, st, un, parent.p, key.p, flag.p);
        CPM_LAUNCH_CHECK();
        exclusive_scan(flag.p, flag.p, n + 1, scratch.p, st);
        hipLaunchKernelGGL(k_fof_number, dim3(nblk(n, FOF_NT)), dim3(FOF_NT), 0, st, un, key.p, flag.p, dgroup);
        CPM_LAUNCH_CHECK();
        CPM_HIP(hipEventRecord(ev1, st));
        uint32_t ng = 0;
        std::vector<unsigned long long> hs(2 * FOF_NSLOT);
        CPM_HIP(hipMemcpyAsync(&ng, flag.p + n, sizeof ng, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipMemcpyAsync(hs.data(), stats.p, sizeof(unsigned long long) * hs.size(), hipMemcpyDeviceToHost, st));
        if (!device) CPM_HIP(hipMemcpyAsync(hgroup, group.p, sizeof(long long) * n, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
        float ms = 0.f;
        CPM_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        last.candidates = last.links = 0;
        for (int i = 0; i < FOF_NSLOT; ++i) {
            last.candidates += (long long)hs[i];
            last.links += (long long)hs[FOF_NSLOT + i];
        }
        last.ngroups = ng;
        fill_stats(n, ms);
        return ng;
    }

    void fill_stats(long n, double ms) {
        last.device_ms = ms;
        last.particles = n;
        for (int k = 0; k < 3; ++k) last.cells[k] = g.nc[k];
        last.table_bits = tbits;
    }
};

}  // namespace cpm

struct pm_fof_plan {
    std::unique_ptr<cpm::FofPlan> plan;
    std::string err;
};

static thread_local std::string g_fof_err;

#define PMF_TRY try {
#define PMF_CATCH(F)                                                    \
    }                                                                   \
    catch (const cpm::Error& e) {                                       \
        if (F) (F)->err = e.what(); else g_fof_err = e.what();          \
        return e.code;                                                  \
    }                                                                   \
    catch (const std::exception& e) {                                   \
        if (F) (F)->err = e.what(); else g_fof_err = e.what();          \
        return PM_EDEVICE;                                              \
    }                                                                   \
    return PM_OK;

static const long FOF_MAX_N = 2147483647L;  // 32-bit sorted positions and keys

extern "C" {

int pm_fof_create(int dim, double L, double ll, pm_fof_plan** out) {
    pm_fof_plan* f = nullptr;
    PMF_TRY
    if (!out) throw cpm::Error(PM_EINVAL, "null argument");
    *out = nullptr;
    if (dim != 2 && dim != 3) throw cpm::Error(PM_EUNSUP, "dim must be 2 or 3");
    if (!(L > 0) || !std::isfinite(L)) throw cpm::Error(PM_EINVAL, "L must be positive and finite");
    if (!(ll > 0) || !std::isfinite(ll)) throw cpm::Error(PM_EINVAL, "the linking length must be positive and finite");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw cpm::Error(PM_EDEVICE, "no HIP device available");
    std::unique_ptr<pm_fof_plan> h(new pm_fof_plan);
    h->plan.reset(new cpm::FofPlan(dim, L, ll));
    *out = h.release();
    PMF_CATCH(f)
}

void pm_fof_destroy(pm_fof_plan* f) { delete f; }

const char* pm_fof_last_error(const pm_fof_plan* f) { return f ? f->err.c_str() : g_fof_err.c_str(); }

int pm_fof_run(pm_fof_plan* f, long long n, const double* pos, long long* group, long long* ngroups, int device) {
    if (!f) {
        g_fof_err = "null handle";
        return PM_EINVAL;
    }
    PMF_TRY
    if (n < 0 || n > FOF_MAX_N) throw cpm::Error(PM_EINVAL, "particle count must lie in [0, 2^31)");
    if (n > 0 && (!pos || !group)) throw cpm::Error(PM_EINVAL, "null argument");
    if (!device && n > 0) cpm::check_finite(pos, (size_t)f->plan->R * n, "positions");
    const long long ng = f->plan->run((long)n, pos, group, device != 0);
    if (ngroups) *ngroups = ng;
    PMF_CATCH(f)
}

int pm_fof_get_stats(const pm_fof_plan* f, pm_fof_stats* s) {
    if (!f || !s) {
        g_fof_err = "null argument";
        return PM_EINVAL;
    }
    *s = f->plan->last;
    return PM_OK;
}

int pm_fof(int dim, long long n, double L, double ll, const double* pos, long long* group, long long* ngroups) {
    pm_fof_plan* f = nullptr;
    int rc = pm_fof_create(dim, L, ll, &f);
    if (rc != PM_OK) return rc;
    rc = pm_fof_run(f, n, pos, group, ngroups, 0);
    if (rc != PM_OK) g_fof_err = f->err;
    pm_fof_destroy(f);
    return rc;
}

}  // extern "C"
