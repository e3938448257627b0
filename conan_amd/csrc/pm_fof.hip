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
//           each pair is tested once) with the reference's predicate in IEEE
//           double arithmetic (no contraction), and unites the linked pairs:
//           roots are hooked under smaller roots by compare-and-swap
//   roots   every particle finds its root (the component's first sorted
//           position) and lowers the component's minimum original index
//   number  flag[i] = particle i is its component's lowest index; the scan of
//           the flags numbers the components in the reference's order
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

__device__ __forceinline__ uint32_t uf_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x.  Parent links only ever point to smaller positions of the same
// set, so a stale read is still an ancestor; path halving stores such an
// ancestor into a non-root, which every thread may do.
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
    uint32_t p = uf_load(parent + x);
    while (p != x) {
        const uint32_t gp = uf_load(parent + p);
        if (gp != p) __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
        p = gp;
    }
    return x;
}

// Joins the sets of a and b: the larger root is hooked under the smaller
// one, so a set's root is its smallest position.  A failed compare-and-swap
// means that root was hooked meanwhile; the union is retried from there.
__device__ __forceinline__ void uf_unite(uint32_t* parent, uint32_t a, uint32_t b) {
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
            return;
        a = uf_find(parent, a);
        b = uf_find(parent, expect);
    }
}

template <int R>
__global__ __launch_bounds__(FOF_NT) void k_fof_cells(const double* __restrict__ pos, uint32_t n, FofGrid g,
                                                      uint32_t* __restrict__ key, uint32_t* __restrict__ slot,
                                                      uint32_t* __restrict__ cnt) {
    const uint32_t i = blockIdx.x * FOF_NT + threadIdx.x;
    if (i >= n) return;
    double x[3];
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = pos[(size_t)i * R + k];
    int c[3];
    fof_cell<R>(x, g, c);
    const uint32_t kk = fof_key(c[0], c[1], c[2], g);
    key[i] = kk;
    slot[i] = atomicAdd(&cnt[kk], 1u);
}

template <int R>
__global__ __launch_bounds__(FOF_NT) void k_fof_place(const double* __restrict__ pos, uint32_t n,
                                                      const uint32_t* __restrict__ key,
                                                      const uint32_t* __restrict__ slot,
                                                      const uint32_t* __restrict__ bstart, double* __restrict__ spos,
                                                      uint32_t* __restrict__ sidx, uint32_t* __restrict__ parent) {
    const uint32_t i = blockIdx.x * FOF_NT + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = bstart[key[i]] + slot[i];
#pragma unroll
    for (int k = 0; k < R; ++k) spos[(size_t)s * R + k] = pos[(size_t)i * R + k];
    sidx[s] = i;
    parent[s] = s;
}

template <int R>
__global__ __launch_bounds__(FOF_NT) void k_fof_link(const double* __restrict__ spos, uint32_t n,
                                                     const uint32_t* __restrict__ bstart, FofGrid g,
                                                     uint32_t* parent, unsigned long long* __restrict__ stats) {
    const uint32_t s = blockIdx.x * FOF_NT + threadIdx.x;
    uint32_t ntest = 0, nlink = 0;
    if (s < n) {
        double a[3];
#pragma unroll
        for (int k = 0; k < R; ++k) a[k] = spos[(size_t)s * R + k];
        int c[3];
        fof_cell<R>(a, g, c);
        auto scan = [&](uint32_t lo, uint32_t hi) {
            for (uint32_t t = lo; t < hi; ++t) {
                double p[3];
#pragma unroll
                for (int k = 0; k < R; ++k) p[k] = spos[(size_t)t * R + k];
                ++ntest;
                if (fof_near<R>(a, p, g)) {
                    ++nlink;
                    uf_unite(parent, s, t);
                }
            }
        };
        const bool xfull = g.nc[0] >= 3;
        const int ylo = g.nc[1] >= 3 ? -1 : 0, yhi = -ylo;
        const int zhi = (R == 3 && g.nc[2] >= 3) ? 1 : 0;  // half shell: dz >= 0
        const int xp = fof_wrap(c[0] + 1, g.nc[0]), xm = fof_wrap(c[0] - 1, g.nc[0]);
        for (int dz = 0; dz <= zhi; ++dz) {
            const int cz = fof_wrap(c[2] + dz, g.nc[2]);
            for (int dy = (dz == 0 ? 0 : ylo); dy <= yhi; ++dy) {
                const int cy = fof_wrap(c[1] + dy, g.nc[1]);
                const uint32_t k0 = fof_key(c[0], cy, cz, g);
                if (dz == 0 && dy == 0) {  // own row: the later particles of this cell and the +x cell
                    if (!xfull) {
                        scan(s + 1, bstart[k0 + 1]);
                        continue;
                    }
                    const uint32_t kp = fof_key(xp, cy, cz, g);
                    if (kp == k0 + 1) {
                        scan(s + 1, bstart[kp + 1]);
                    } else {
                        scan(s + 1, bstart[k0 + 1]);
                        scan(bstart[kp], bstart[kp + 1]);
                    }
                    continue;
                }
                if (!xfull) {
                    scan(bstart[k0], bstart[k0 + 1]);
                    continue;
                }
                const uint32_t km = fof_key(xm, cy, cz, g), kp = fof_key(xp, cy, cz, g);
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
    // wave sums of the counters, one atomic per wave
    unsigned long long t = ntest, l = nlink;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        t += __shfl_down(t, off);
        l += __shfl_down(l, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], t);
        atomicAdd(&stats[1], l);
    }
}

__global__ __launch_bounds__(FOF_NT) void k_fof_roots(uint32_t n, uint32_t* parent,
                                                      const uint32_t* __restrict__ sidx, uint32_t* __restrict__ root,
                                                      uint32_t* __restrict__ cmin) {
    const uint32_t s = blockIdx.x * FOF_NT + threadIdx.x;
    if (s >= n) return;
    const uint32_t r = uf_find(parent, s);
    root[s] = r;
    atomicMin(&cmin[r], sidx[s]);
}

__global__ __launch_bounds__(FOF_NT) void k_fof_flags(uint32_t n, const uint32_t* __restrict__ sidx,
                                                      const uint32_t* __restrict__ root,
                                                      const uint32_t* __restrict__ cmin, uint32_t* __restrict__ flag) {
    const uint32_t s = blockIdx.x * FOF_NT + threadIdx.x;
    if (s > n) return;
    if (s == n) {
        flag[n] = 0;
        return;
    }
    const uint32_t i = sidx[s];
    flag[i] = cmin[root[s]] == i ? 1u : 0u;
}

__global__ __launch_bounds__(FOF_NT) void k_fof_number(uint32_t n, const uint32_t* __restrict__ sidx,
                                                       const uint32_t* __restrict__ root,
                                                       const uint32_t* __restrict__ cmin,
                                                       const uint32_t* __restrict__ rank, long long* __restrict__ group) {
    const uint32_t s = blockIdx.x * FOF_NT + threadIdx.x;
    if (s >= n) return;
    group[sidx[s]] = (long long)rank[cmin[root[s]]];
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
    DevBuf<unsigned long long> stats;
    pm_fof_stats last{};

    FofPlan(int R_, double L_, double ll_) : R(R_), L(L_), ll(ll_) {
        CPM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        CPM_HIP(hipEventCreate(&ev0));
        CPM_HIP(hipEventCreate(&ev1));
        stats.alloc(2);
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

    void reserve(long n) {
        // table: about two keys per particle, no more than there are cells
        const double cells = (double)g.nc[0] * g.nc[1] * g.nc[2];
        int tb = ilog2(2 * std::max(n, 1L));
        while (tb > 1 && std::ldexp(1.0, tb - 1) >= cells) --tb;
        tbits = std::max(tb, 1);
        g.kmask = (uint32_t)((1ull << tbits) - 1);
        const long T = 1L << tbits;
        if (n > cap) {
            cap = n;
            pos.alloc((size_t)R * n);
            spos.alloc((size_t)R * n);
            key.alloc(n);
            slot.alloc(n);
            sidx.alloc(n);
            parent.alloc(n);
            flag.alloc(n + 1);
            group.alloc(n);
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
        reserve(n);
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
        fof_fill((uint32_t*)stats.p, 4, 0u, st);
        CPM_R_DISPATCH(R, hipLaunchKernelGGL(k_fof_cells<RR>, dim3(nblk(n, FOF_NT)), dim3(FOF_NT), 0, st, dpos,
                                             un, g, key.p, slot.p, cnt.p));
        CPM_LAUNCH_CHECK();
        exclusive_scan(cnt.p, cnt.p, T + 1, scratch.p, st);
        CPM_R_DISPATCH(R, hipLaunchKernelGGL(k_fof_place<RR>, dim3(nblk(n, FOF_NT)), dim3(FOF_NT), 0, st, dpos,
                                             un, key.p, slot.p, cnt.p, spos.p, sidx.p, parent.p));
        CPM_LAUNCH_CHECK();
        CPM_R_DISPATCH(R, hipLaunchKernelGGL(k_fof_link<RR>, dim3(nblk(n, FOF_NT)), dim3(FOF_NT), 0, st, spos.p,
                                             un, cnt.p, g, parent.p, stats.p));
        CPM_LAUNCH_CHECK();
        // key -> component minimum (cmin), slot -> root
        fof_fill(key.p, n, 0xFFFFFFFFu, st);
        hipLaunchKernelGGL(k_fof_roots, dim3(nblk(n, FOF_NT)), dim3(FOF_NT), 0, st, un, parent.p, sidx.p, slot.p,
                           key.p);
        CPM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_fof_flags, dim3(nblk(n + 1, FOF_NT)), dim3(FOF_NT), 0, st, un, sidx.p, slot.p, key.p,
                           flag.p);
        CPM_LAUNCH_CHECK();
        exclusive_scan(flag.p, flag.p, n + 1, scratch.p, st);
        hipLaunchKernelGGL(k_fof_number, dim3(nblk(n, FOF_NT)), dim3(FOF_NT), 0, st, un, sidx.p, slot.p, key.p,
                           flag.p, dgroup);
        CPM_LAUNCH_CHECK();
        CPM_HIP(hipEventRecord(ev1, st));
        uint32_t ng = 0;
        unsigned long long hs[2] = {0, 0};
        CPM_HIP(hipMemcpyAsync(&ng, flag.p + n, sizeof ng, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipMemcpyAsync(hs, stats.p, sizeof hs, hipMemcpyDeviceToHost, st));
        if (!device) CPM_HIP(hipMemcpyAsync(hgroup, group.p, sizeof(long long) * n, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
        float ms = 0.f;
        CPM_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        last.candidates = (long long)hs[0];
        last.links = (long long)hs[1];
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
