// pm_impl.hpp -- the device state of one PM run (the pm_handle of the C ABI)
// and the host-side step sequence over it; shared by the C ABI and the
// differentiable driver.
//
// The handle is the MI355X counterpart of Gravity<R> (src/run/nbody.hh:53-219):
// it keeps positions and momenta resident in HBM across steps, and one kick is
//   keys -> stable (key, id) sort -> row bounds -> deposit (rho)
//   -> rows r2c -> [y fwd] -> last axis fwd * Green * bwd -> [y bwd] -> rows c2r (phi)
//   -> gather + kick (+ the next step's drift).
// Scalars that the reference computes per call (adot, G, the drift and kick
// factors, the Poisson scale) are computed here on the host with the
// reference's expressions, so both sides see the same doubles.
#pragma once
#include "../../include/conanpm.h"
#include "pm_adjoint.hpp"
#include "pm_fft_api.hpp"
#include "pm_particles.hpp"
#include "pm_sort.hpp"
#include "pm_spectral.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace cpm {

// src/run/cosmology.hh:6-28
struct Cosmo {
    double H0, Om, OL, Ok, G;
    Cosmo(double H0_, double m_, double L_) : H0(H0_), Om(m_), OL(L_), Ok(1 - m_ - L_), G(3. / 2 * m_ * H0_ * H0_) {}
    double adot(double a) const { return H0 * a * sqrt(OL + Om * pow(a, -3) + Ok * pow(a, -2)); }
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n_) { alloc(n_); }
    void alloc(size_t n_) {
        release();
        void* q = nullptr;
        device_alloc(&q, sizeof(T) * n_);
        p = static_cast<T*>(q);
        n = n_;
    }
    void release() {
        device_free(p, sizeof(T) * n);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

inline void check_finite(const double* a, size_t n, const char* what) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) throw Error(PM_ENONFINITE, std::string("non-finite value in ") + what);
}

inline int ilog2_exact(long n, const char* what) {
    if (n < 1 || (n & (n - 1))) throw Error(PM_EUNSUP, std::string(what) + " must be a power of two");
    return ilog2(n);
}

// Per-phase device time from HIP event pairs recorded on the handle's
// stream.  Recording does not synchronise: pairs queue up in a pool and are
// resolved when the totals are read, so profiling can stay on in a timed run.
struct Prof {
    bool on = false;
    std::vector<hipEvent_t> evb, eve;
    std::vector<int> phase;
    size_t used = 0;
    double ms[PM_NPHASES] = {};
    long long cnt[PM_NPHASES] = {};
    void enable(bool v) {
        if (v) {
            for (int i = 0; i < PM_NPHASES; ++i) {
                ms[i] = 0;
                cnt[i] = 0;
            }
            used = 0;
        }
        on = v;
    }
    void begin(int ph, hipStream_t st) {
        if (!on) return;
        if (used == evb.size()) {
            hipEvent_t a, b;
            CPM_HIP(hipEventCreate(&a));
            CPM_HIP(hipEventCreate(&b));
            evb.push_back(a);
            eve.push_back(b);
            phase.push_back(ph);
        }
        phase[used] = ph;
        CPM_HIP(hipEventRecord(evb[used], st));
    }
    void end(int ph, hipStream_t st) {
        if (!on) return;
        CPM_HIP(hipEventRecord(eve[used], st));
        ++used;
        (void)ph;
    }
    // Folds the queued pairs into the totals (synchronises on the last event).
    void collect() {
        if (used == 0) return;
        CPM_HIP(hipEventSynchronize(eve[used - 1]));
        for (size_t i = 0; i < used; ++i) {
            float t = 0;
            CPM_HIP(hipEventElapsedTime(&t, evb[i], eve[i]));
            ms[phase[i]] += t;
            ++cnt[phase[i]];
        }
        used = 0;
    }
    ~Prof() {
        for (size_t i = 0; i < evb.size(); ++i) {
            (void)hipEventDestroy(evb[i]);
            (void)hipEventDestroy(eve[i]);
        }
    }
};

struct ImplBase {
    virtual ~ImplBase() = default;
    virtual void init_zeldovich(const double* phi, double a_pos, double a_mom) = 0;
    virtual void drift(double a, double da) = 0;
    virtual void kick(double a, double da) = 0;
    virtual double steps(double a, double da, int n) = 0;
    virtual void get_state(double* X, double* P) = 0;
    virtual void set_state(const double* X, const double* P) = 0;
    virtual void set_positions_device(const double* Xsoa) = 0;
    virtual void deposit(double* rho, int32_t* keys, double mass_override) = 0;
    virtual void potential(double a, double* phi) = 0;
    virtual void forces(double a, double* F) = 0;
    virtual void poisson_mesh(const double* rho, double a, double* phi) = 0;
    virtual void gather_mesh(const double* phi, double* F) = 0;
    virtual void power_spectrum(int nbins, const double* edges, double* k, double* pk, double* cnt) = 0;
    virtual void power_spectrum_noalloc(int nbins, const double* edges, double* k, double* pk, double* cnt) = 0;
    virtual void diagnostics(pm_diag* d) = 0;
    virtual double momentum_quantum() const = 0;
    virtual void set_momentum_quantum(double pq) = 0;
    double last_apos = 0, last_avel = 0;
    bool have_state = false;
    Prof prof;
    hipStream_t st = nullptr;
};

template <typename PP>
struct Impl : ImplBase {
    using Pos = typename PP::Pos;
    using Mesh = typename PP::Mesh;
    static constexpr bool FX = sizeof(Mesh) == 4;

    pm_config cfg;
    int R, lgNf;
    long Np, M, nrows;
    Cosmo cos;
    double mass;
    PosArgs pa;
    FFTPlan<Mesh> fft;
    RadixSorter sorter;
    DevBuf<Pos> X, X2;
    DevBuf<double> P, P2;
    // The brick forms (FX, and FP64 without PM_FLAG_SERIAL_DEPOSIT) keep the
    // particles in the brick order of the last deposit; id[s] is the
    // mass-grid index of the particle stored at s (X2/P2/id2: the next order).
    DevBuf<uint32_t> id, id2;
    DevBuf<uint32_t> k0, v0, k1, v1, row_start;
    // FX incremental re-sort (see fx_sort): new keys, sorted keys / perm
    // outputs, the other group-start buffer, per-group counters (movers in |
    // movers out | fill), stayers, group totals.
    DevBuf<uint32_t> knew, skb0, skb1, permb, bs2, cnt3, stay, tot, scan_scr;
    // brick gathers: one flag per sorted position for the particles that
    // overflow a workgroup's LDS queue (all zero between launches)
    DevBuf<uint32_t> defer;
    const uint32_t* bstart = nullptr;  // group starts of the last sort (row_start or bs2)
    bool resort_ok = false;            // storage order == output order of the last brick sort
    bool keys_ready = false;           // knew / cnt3 hold the keys of the current positions (fused gather)
    bool skeys_valid = false;          // skeys holds the keys of the last sort (the re-sort does not write them)
    // in_count / out_count / fill (cnt3) are all zero: set by a re-sort, which
    // clears them, and reset by every kernel that counts into them
    bool counts_clear = false;
    DevBuf<Mesh> mesh;
    DevBuf<Mesh> frac;  // FP64 deposit: fractional cell coordinates in sorted order
    DevBuf<Pos> xsort;     // FP64: positions in the deposit's sorted order
    DevBuf<double> fsort;  // FP64: forces in that order
    DevBuf<uint32_t> inv;  // FP64: sorted position of each particle
    DevBuf<cx<Mesh>> spec, nyq;
    DevBuf<double> row_partial, red, scal;
    const uint32_t* skeys = nullptr;
    const uint32_t* spids = nullptr;
    bool have_deposit = false;
    long npartials = 0;     // FP64 mesh-sum partials written by the last deposit
    long nsort_groups = 0;  // key groups of the sort: brick key groups, or mesh rows (serial-order deposit)
    double pq = 0;          // FX: momentum quantum (a power of two), see fx_quantize_momenta
    bool pq_pinned = false; // FX: pq was set by the caller (kept by set_state)
    Mesh* phi_out = nullptr;  // when set, the potential of a kick goes here instead of into mesh
    bool bricks;              // the brick order (brick_form)

    // the brick order (FX always; FP64 unless the serial-order deposit is asked for)
    static bool brick_form(const pm_config& c) { return FX || !(c.flags & PM_FLAG_SERIAL_DEPOSIT); }
    static int fx_keybits(int dim, int lgN) {
        PosArgs t{};
        brick_layout(dim, lgN, t, !FX);
        return t.lnb[0] + t.lnb[1] + t.lnb[2] + dim;  // brick, then face class
    }

    static long ipow(long n, int r) {
        long s = 1;
        for (int i = 0; i < r; ++i) s *= n;
        return s;
    }

    explicit Impl(const pm_config& c)
        : cfg(c), R(c.dim), lgNf(ilog2_exact(c.Nf, "Nf")), Np(ipow(c.Nm, c.dim)), M(ipow(c.Nf, c.dim)),
          nrows(M / c.Nf), cos(c.H0, c.Omega_m, c.Omega_L), fft(c.dim, c.Nf),
          // (key, index) pairs sorted by brick and face class, or (FP64 with
          // the serial-order deposit) by the full cell index
          sorter(Np, brick_form(c) ? fx_keybits(c.dim, lgNf) : c.dim * lgNf), bricks(brick_form(c)) {
        mass = pow(double(c.Nf) / c.Nm, R);  // src/run/nbody.hh:85
        pa.L = c.L;
        pa.res = c.L / c.Nf;
        pa.rinv = 1.0 / pa.res;
        pa.fastdiv = divides_by_fma(pa.res) ? 1 : 0;
        pa.Linv = 1.0 / pa.L;
        pa.fastdivL = divides_by_fma(pa.L) ? 1 : 0;
        pa.wrap = (c.flags & PM_FLAG_WRAP_GATHER) ? 1 : 0;
        pa.fshift = FX_BITS - lgNf;
        pa.fmask = (uint32_t)((1ull << pa.fshift) - 1);
        pa.fscale = (float)ldexp(1.0, -pa.fshift);
        pa.cmask = c.Nf - 1;
        pa.brick = 0;
        for (int k = 0; k < 3; ++k) pa.lb[k] = pa.lnb[k] = 0;
        if (bricks) brick_layout(R, lgNf, pa, !FX);
        nsort_groups = bricks ? (brick_count(pa) << R) : nrows;
        CPM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        try {
            allocate();
        } catch (...) {
            (void)hipStreamDestroy(st);  // the buffers allocated so far go with the members
            st = nullptr;
            throw;
        }
    }
    void allocate() {
        X.alloc((size_t)R * Np);
        P.alloc((size_t)R * Np);
        if (bricks) {
            X2.alloc((size_t)R * Np);
            P2.alloc((size_t)R * Np);
            id.alloc(Np);
            id2.alloc(Np);
        }
        k0.alloc(Np);
        v0.alloc(Np);
        k1.alloc(Np);
        v1.alloc(Np);
        row_start.alloc((nsort_groups > nrows ? nsort_groups : nrows) + 1);
        if (bricks) {
            knew.alloc(Np);
            skb0.alloc(Np);
            skb1.alloc(Np);
            permb.alloc(Np);
            bs2.alloc(nsort_groups + 1);
            cnt3.alloc(3 * nsort_groups);
            stay.alloc(nsort_groups);
            tot.alloc(nsort_groups + 1);
            scan_scr.alloc(scan_scratch_elems(nsort_groups + 1));
            defer.alloc(Np);
            zero_u32(defer.p, Np, st);
        }
        if (!bricks) {  // the serial-order deposit's and its kick's sorted-order copies
            frac.alloc((size_t)R * Np);
            xsort.alloc((size_t)R * Np);
            fsort.alloc((size_t)R * Np);
            inv.alloc(Np);
        }
        mesh.alloc(M);
        spec.alloc(fft.spec_elems);
        nyq.alloc(fft.nyq_elems);
        row_partial.alloc(bricks && brick_count(pa) > nrows ? brick_count(pa) : nrows);
        red.alloc(4096);
        scal.alloc(16);
    }
    ~Impl() override {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }

    void need_state() const {
        if (!have_state) throw Error(PM_ESTATE, "no particle state: call pm_init_zeldovich or pm_set_state first");
    }

    // Storage order = mass-grid order (set after ZA / set_state).
    void reset_ids() {
        invalidate_order();
        if (!bricks) return;
        std::vector<uint32_t> h(Np);
        for (long i = 0; i < Np; ++i) h[i] = (uint32_t)i;
        CPM_HIP(hipMemcpyAsync(id.p, h.data(), sizeof(uint32_t) * Np, hipMemcpyHostToDevice, st));
        CPM_HIP(hipStreamSynchronize(st));
    }
    // Mass-grid index of each stored particle.
    std::vector<uint32_t> host_ids() {
        std::vector<uint32_t> h(Np);
        if (bricks) {
            CPM_HIP(hipMemcpyAsync(h.data(), id.p, sizeof(uint32_t) * Np, hipMemcpyDeviceToHost, st));
            CPM_HIP(hipStreamSynchronize(st));
        } else {
            for (long i = 0; i < Np; ++i) h[i] = (uint32_t)i;
        }
        return h;
    }
    // SoA in storage order -> AoS in mass-grid order.
    void soa_to_aos(const std::vector<double>& soa, const std::vector<uint32_t>& ids, double* out) const {
        for (long s = 0; s < Np; ++s)
            for (int k = 0; k < R; ++k) out[(size_t)ids[s] * R + k] = soa[(size_t)k * Np + s];
    }
    // Scale factor at which the factors of a drift or kick over [a, a + da]
    // are evaluated: its start (the reference) or its middle (PM_FLAG_MIDPOINT).
    double a_eval(double a, double da) const { return (cfg.flags & PM_FLAG_MIDPOINT) ? a + da / 2 : a; }
    double drift_factor(double a_in, double da) const {
        const double a = a_eval(a_in, da);
        double adot = cos.adot(a);
        double c = da / (a * a * adot);  // src/run/nbody.hh:124,129
        return FX ? c * (FX_UNIT / cfg.L) : c;
    }

    // FX momenta live on a grid of pq = 2^(e - 36), max |P| < 2^e (2^-40 L
    // if all P are 0): kicks add multiples of pq (fx_kick_increment), which is
    // exact while |P| < 2^53 pq = 2^17 2^e, so every kick can be undone bit
    // for bit.  The grid is 2^-36 of the largest initial momentum: finer than
    // the FP32 force by far.
    void fx_quantize_momenta() {
        if constexpr (FX) {
            if (pq_pinned) {
                quantize(P.p, (long)R * Np, pq, st);
                CPM_HIP(hipStreamSynchronize(st));
                return;
            }
            reduce_max_abs(P.p, (long)R * Np, red.p, 4096, scal.p + 8, st);
            double mx = 0;
            CPM_HIP(hipMemcpyAsync(&mx, scal.p + 8, sizeof(double), hipMemcpyDeviceToHost, st));
            CPM_HIP(hipStreamSynchronize(st));
            int e = 0;
            if (mx > 0 && std::isfinite(mx)) {
                (void)frexp(mx, &e);
                pq = ldexp(1.0, e - 36);
            } else {
                pq = ldexp(1.0, ilogb(cfg.L) - 40);
            }
            quantize(P.p, (long)R * Np, pq, st);
            CPM_HIP(hipStreamSynchronize(st));
        }
    }

    double momentum_quantum() const override { return FX ? pq : 0.0; }
    // Pin the FX momentum grid (a power of two): set_state keeps it instead
    // of choosing one from the momenta, until init_zeldovich.  A run that
    // resumes from a snapshot sets it first and continues on its grid.
    void set_momentum_quantum(double q) override {
        if constexpr (FX) {
            int e = 0;
            if (!(q > 0) || !std::isfinite(q) || frexp(q, &e) != 0.5)
                throw Error(PM_EINVAL, "the momentum quantum must be a positive power of two");
            pq = q;
            pq_pinned = true;
            if (have_state) {
                quantize(P.p, (long)R * Np, pq, st);
                CPM_HIP(hipStreamSynchronize(st));
            }
        } else {
            if (q != 0.0) throw Error(PM_EUNSUP, "the FP64 form keeps plain double momenta (quantum 0)");
        }
    }

    void init_zeldovich(const double* phi, double a_pos, double a_mom) override {
        const long npm = Np;
        if (cfg.Nm < 4 || (cfg.Nm & (cfg.Nm - 1))) throw Error(PM_EUNSUP, "Zel'dovich start needs a power-of-two Nm >= 4");
        check_finite(phi, npm, "potential");
        DevBuf<double> d(npm);
        CPM_HIP(hipMemcpyAsync(d.p, phi, sizeof(double) * npm, hipMemcpyHostToDevice, st));
        ParticleOps<PP>::zeldovich(R, d.p, ilog2(cfg.Nm), cfg.L, a_pos, a_mom, X.p, P.p, st);
        CPM_HIP(hipStreamSynchronize(st));
        pq_pinned = false;  // a new run: its grid comes from its initial momenta
        fx_quantize_momenta();
        reset_ids();
        last_apos = a_pos;
        last_avel = a_mom;
        have_state = true;
        have_deposit = false;
    }

    // FX sort, in two steps around a keys kernel (keys_fx, or the adjoint's
    // reverse drift): fx_key_args() gives the kernel its outputs, fx_sort()
    // turns them into bstart and the permutation spids.  While the storage
    // order is the output order of the previous sort (resort_ok: a gather
    // moved the particles there), the sort is incremental: only particles
    // whose brick or face class changed are placed anew.  Otherwise a full
    // stable radix sort of (key, index) runs.
    struct FxKeyArgs {
        const uint32_t* prevkeys;
        uint32_t *keys, *pids, *in_count, *out_count;
    };
    // Positions or storage order changed outside a gather: the next FX sort
    // is a full one.
    void invalidate_order() {
        resort_ok = false;
        keys_ready = false;
    }
    FxKeyArgs fx_key_args() {
        keys_ready = false;
        if (resort_ok) {
            zero_u32(cnt3.p, 3 * nsort_groups, st);
            counts_clear = false;  // the keys kernel counts into them
            if (!skeys_valid) {  // the previous keys, from the group starts of the last sort
                uint32_t* sk = (skeys == skb0.p) ? skb1.p : skb0.p;
                group_keys(bstart, brick_count(pa), 1 << R, sk, st);
                skeys = sk;
                skeys_valid = true;
            }
            return FxKeyArgs{skeys, knew.p, nullptr, cnt3.p, cnt3.p + nsort_groups};
        }
        return FxKeyArgs{nullptr, k0.p, v0.p, nullptr, nullptr};
    }
    void fx_sort() {
        keys_ready = false;
        if (resort_ok) {
            uint32_t* bnew = (bstart == row_start.p) ? bs2.p : row_start.p;
            resort_counts(bstart, cnt3.p, cnt3.p + nsort_groups, cnt3.p + 2 * nsort_groups, nsort_groups, stay.p,
                          tot.p, st);
            counts_clear = true;
            exclusive_scan(tot.p, bnew, nsort_groups + 1, scan_scr.p, st);
            resort_place(knew.p, bstart, bnew, stay.p, cnt3.p + 2 * nsort_groups, brick_count(pa), 1 << R, Np, permb.p,
                         st);
            skeys_valid = false;
            spids = permb.p;
            bstart = bnew;
        } else {
            const int where = sorter.sort(k0.p, v0.p, k1.p, v1.p, st);
            skeys = where ? k1.p : k0.p;
            skeys_valid = true;
            spids = where ? v1.p : v0.p;
            row_bounds(skeys, Np, 0, nsort_groups, row_start.p, st);
            bstart = row_start.p;
        }
        resort_ok = false;  // until a gather moves the particles into this order
    }
    // Keys of the current positions (unless a fused gather produced them), then fx_sort.
    void fx_sort_current() {
        if (!keys_ready) {
            const FxKeyArgs ka = fx_key_args();
            if constexpr (FX) {
                keys_fx(R, reinterpret_cast<const uint32_t*>(X.p), Np, lgNf, pa, ka.prevkeys, ka.keys, ka.pids,
                        ka.in_count, ka.out_count, st);
            } else {
                keys_f64_brick(R, reinterpret_cast<const double*>(X.p), Np, lgNf, pa, ka.prevkeys, ka.keys, ka.pids,
                               ka.in_count, ka.out_count, st);
            }
        }
        fx_sort();
    }

    // partials: also the FP64 mesh-sum partials (row_partial), for the
    // callers that sum the mesh; the step's Poisson solve does not.
    void do_deposit(Mesh m, bool partials = true) {
        prof.begin(0, st);
        if (bricks) {
            fx_sort_current();
        } else {
            ParticleOps<PP>::keys(R, X.p, Np, lgNf, pa, k0.p, v0.p, st);
            const int where = sorter.sort(k0.p, v0.p, k1.p, v1.p, st);
            skeys = where ? k1.p : k0.p;
            spids = where ? v1.p : v0.p;
            row_bounds(skeys, Np, lgNf, nsort_groups, row_start.p, st);
            bstart = row_start.p;
        }
        prof.end(0, st);
        prof.begin(1, st);
        if constexpr (FX) {
            deposit_fx_brick(R, bstart, spids, X.p, Np, lgNf, pa, m, mesh.p, partials ? row_partial.p : nullptr, st);
            npartials = partials ? brick_count(pa) : 0;
        } else if (bricks) {
            deposit_f64_brick(R, bstart, spids, reinterpret_cast<const double*>(X.p), Np, lgNf, pa, (double)m,
                              reinterpret_cast<double*>(mesh.p), partials ? row_partial.p : nullptr, st);
            npartials = partials ? brick_count(pa) : 0;
        } else {
            ParticleOps<PP>::deposit(R, row_start.p, skeys, spids, X.p, Np, lgNf, pa, m, mesh.p, row_partial.p,
                                     nrows, frac.p, xsort.p, inv.p, st);
            npartials = nrows;
        }
        prof.end(1, st);
        have_deposit = true;
    }

    Mesh* phi_dst() { return phi_out ? phi_out : mesh.p; }
    // mesh (rho) -> phi_dst() (phi) at scale factor a.  mean_ready (FX): the
    // mesh is a caller's density, whose FP64 mean is in scal[0] and is
    // subtracted instead of 1.
    void do_poisson(double a, bool mean_ready = false) {
        const double s = fft.M / cos.G * a;  // real_part(size / G * a), src/run/nbody.hh:110
        const double kf = 2 * M_PI / cfg.L;
        // delta = rho - 1 (src/run/nbody.hh:103-104), in both forms.  The
        // FP32 deposit's mean is 1 to about 1e-11 (particle mass (Nf/Nm)^dim;
        // each weight rounded to 2^-40), so its FP32 value is 1 exactly and
        // no mesh sum is needed for it.
        const double* mean_ptr = (FX && mean_ready) ? scal.p : nullptr;
        const Mesh shift = (Mesh)1.0;
        for (int stage = 0; stage < 5; ++stage) {
            prof.begin(2 + stage, st);
            fft.poisson_stage(stage, 0, mesh.p, phi_dst(), spec.p, nyq.p, shift, mean_ptr, kf, s, st);
            prof.end(2 + stage, st);
        }
    }

    void drift(double a, double da) override {
        need_state();
        prof.begin(8, st);
        ParticleOps<PP>::drift(X.p, P.p, (long)R * Np, drift_factor(a, da), st);
        prof.end(8, st);
        keys_ready = false;
        CPM_HIP(hipStreamSynchronize(st));
        last_apos = a;
    }

    void kick_impl(double a, double da, bool fuse, double a_next, double da_next) {
        do_deposit((Mesh)mass, false);
        do_poisson(a_eval(a, da));
        const double adot = cos.adot(a_eval(a, da));
        const double ck = da / adot;  // src/run/nbody.hh:112-116
        const double cd = fuse ? drift_factor(a_next, da_next) : 0.0;
        prof.begin(7, st);
        if constexpr (FX) {
            // kick (+ drift) and move the particles into the deposit's brick
            // order; with the drift fused, also the keys of the next re-sort
            if (fuse && !counts_clear) zero_u32(cnt3.p, 3 * nsort_groups, st);
            if (fuse) counts_clear = false;
            gather_kick_fx_brick(R, bstart, spids, reinterpret_cast<const uint32_t*>(X.p), P.p, id.p,
                                 reinterpret_cast<uint32_t*>(X2.p), P2.p, id2.p, Np,
                                 reinterpret_cast<const float*>(phi_dst()), lgNf, pa, ck, pq, cd, fuse ? 1 : 0,
                                 fuse ? knew.p : nullptr, cnt3.p, cnt3.p + nsort_groups, defer.p, st);
            std::swap(X.p, X2.p);
            std::swap(P.p, P2.p);
            std::swap(id.p, id2.p);
            resort_ok = true;
            keys_ready = fuse;
        } else if (bricks) {
            // the same with the FP64 tile and the reference's arithmetic
            if (fuse && !counts_clear) zero_u32(cnt3.p, 3 * nsort_groups, st);
            if (fuse) counts_clear = false;
            gather_kick_f64_brick(R, bstart, spids, reinterpret_cast<const double*>(X.p), P.p, id.p,
                                  reinterpret_cast<double*>(X2.p), P2.p, id2.p, Np,
                                  reinterpret_cast<const double*>(phi_dst()), lgNf, pa, ck, cd, fuse ? 1 : 0,
                                  fuse ? knew.p : nullptr, cnt3.p, cnt3.p + nsort_groups, defer.p, st);
            std::swap(X.p, X2.p);
            std::swap(P.p, P2.p);
            std::swap(id.p, id2.p);
            resort_ok = true;
            keys_ready = fuse;
        } else {
            // serial-order deposit: forces in its cell order, then the update in storage order
            ParticleOps<PP>::gather_kick_sorted(R, X.p, P.p, Np, phi_dst(), lgNf, pa, ck, cd, fuse ? 1 : 0, xsort.p,
                                                fsort.p, inv.p, st);
        }
        prof.end(7, st);
        last_avel = a;
        if (fuse) last_apos = a_next;
    }

    void kick(double a, double da) override {
        need_state();
        kick_impl(a, da, false, 0, 0);
        CPM_HIP(hipStreamSynchronize(st));
    }

    // LeapFrog::run's loop body n times (src/run/leapfrog.hh:38-44).
    double steps(double a, double da, int n) override {
        need_state();
        if (n <= 0) return a;
        prof.begin(8, st);
        ParticleOps<PP>::drift(X.p, P.p, (long)R * Np, drift_factor(a, da), st);
        prof.end(8, st);
        keys_ready = false;
        last_apos = a;
        for (int i = 0; i < n; ++i) {
            const double ak = a + da / 2;
            const bool more = i + 1 < n;
            const double anext = a + da;
            kick_impl(ak, da, more, anext, da);
            a += da;
        }
        CPM_HIP(hipStreamSynchronize(st));
        return a;
    }

    void get_state(double* Xh, double* Ph) override {
        need_state();
        std::vector<double> tmp((size_t)R * Np);
        const std::vector<uint32_t> ids = host_ids();
        if (Xh) {
            if (FX) {
                DevBuf<double> d((size_t)R * Np);
                fx_to_f64(reinterpret_cast<const uint32_t*>(X.p), (long)R * Np, cfg.L, d.p, st);
                CPM_HIP(hipMemcpyAsync(tmp.data(), d.p, sizeof(double) * R * Np, hipMemcpyDeviceToHost, st));
                CPM_HIP(hipStreamSynchronize(st));
            } else {
                CPM_HIP(hipMemcpyAsync(tmp.data(), X.p, sizeof(double) * R * Np, hipMemcpyDeviceToHost, st));
                CPM_HIP(hipStreamSynchronize(st));
            }
            soa_to_aos(tmp, ids, Xh);
        }
        if (Ph) {
            CPM_HIP(hipMemcpyAsync(tmp.data(), P.p, sizeof(double) * R * Np, hipMemcpyDeviceToHost, st));
            CPM_HIP(hipStreamSynchronize(st));
            soa_to_aos(tmp, ids, Ph);
        }
    }

    void upload_positions_soa(const double* soa_host_or_dev, bool device) {
        if (FX) {
            DevBuf<double> d;
            const double* src = soa_host_or_dev;
            if (!device) {
                d.alloc((size_t)R * Np);
                CPM_HIP(hipMemcpyAsync(d.p, soa_host_or_dev, sizeof(double) * R * Np, hipMemcpyHostToDevice, st));
                src = d.p;
            }
            f64_to_fx(src, (long)R * Np, cfg.L, reinterpret_cast<uint32_t*>(X.p), st);
            CPM_HIP(hipStreamSynchronize(st));
        } else {
            CPM_HIP(hipMemcpyAsync(X.p, soa_host_or_dev, sizeof(double) * R * Np,
                                   device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
            CPM_HIP(hipStreamSynchronize(st));
        }
    }

    void set_state(const double* Xh, const double* Ph) override {
        std::vector<double> tmp((size_t)R * Np);
        const bool had_state = have_state;
        if (Xh) check_finite(Xh, (size_t)R * Np, "positions");
        if (Ph) check_finite(Ph, (size_t)R * Np, "momenta");
        if (Xh) {
            if (bricks && !Ph && have_state) {
                // the kept momenta follow the storage order of the last step;
                // the new positions come in mass-grid order, which reset_ids
                // makes the storage order: put the momenta there first
                DevBuf<double> byid((size_t)R * Np);
                scatter_by_id<double>(P.p, id.p, Np, R, byid.p, st);
                CPM_HIP(hipMemcpyAsync(P.p, byid.p, sizeof(double) * R * Np, hipMemcpyDeviceToDevice, st));
                CPM_HIP(hipStreamSynchronize(st));
            }
            for (long i = 0; i < Np; ++i)
                for (int k = 0; k < R; ++k) tmp[(size_t)k * Np + i] = Xh[i * R + k];
            upload_positions_soa(tmp.data(), false);
            reset_ids();
        }
        if (Ph) {
            const std::vector<uint32_t> ids = host_ids();  // momenta follow the stored order
            for (long s = 0; s < Np; ++s)
                for (int k = 0; k < R; ++k) tmp[(size_t)k * Np + s] = Ph[(size_t)ids[s] * R + k];
            CPM_HIP(hipMemcpyAsync(P.p, tmp.data(), sizeof(double) * R * Np, hipMemcpyHostToDevice, st));
            CPM_HIP(hipStreamSynchronize(st));
        } else if (!have_state) {
            CPM_HIP(hipMemsetAsync(P.p, 0, sizeof(double) * R * Np, st));
            CPM_HIP(hipStreamSynchronize(st));
        }
        if (Xh) have_state = true;
        have_deposit = false;
        // new momenta (or zeros): choose their grid; kept momenta stay on theirs
        if (Ph || !had_state) fx_quantize_momenta();
    }

    void set_positions_device(const double* Xsoa) override {
        upload_positions_soa(Xsoa, true);
        reset_ids();
        CPM_HIP(hipMemsetAsync(P.p, 0, sizeof(double) * R * Np, st));
        CPM_HIP(hipStreamSynchronize(st));
        fx_quantize_momenta();
        have_state = true;
    }

    void mesh_to_host(double* out) {
        if (FX) {
            std::vector<float> t(M);
            CPM_HIP(hipMemcpyAsync(t.data(), mesh.p, sizeof(float) * M, hipMemcpyDeviceToHost, st));
            CPM_HIP(hipStreamSynchronize(st));
            for (long i = 0; i < M; ++i) out[i] = t[i];
        } else {
            CPM_HIP(hipMemcpyAsync(out, mesh.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
            CPM_HIP(hipStreamSynchronize(st));
        }
    }
    void mesh_from_host(const double* in) {
        if (FX) {
            std::vector<float> t(M);
            for (long i = 0; i < M; ++i) t[i] = (float)in[i];
            CPM_HIP(hipMemcpyAsync(mesh.p, t.data(), sizeof(float) * M, hipMemcpyHostToDevice, st));
            CPM_HIP(hipStreamSynchronize(st));
        } else {
            CPM_HIP(hipMemcpyAsync(mesh.p, in, sizeof(double) * M, hipMemcpyHostToDevice, st));
            CPM_HIP(hipStreamSynchronize(st));
        }
    }

    void deposit(double* rho, int32_t* keys, double mass_override) override {
        need_state();
        do_deposit((Mesh)(mass_override > 0 ? mass_override : mass));
        if (rho) mesh_to_host(rho);
        if (keys) {
            PosArgs pk = pa;
            pk.brick = 0;  // the full cell index idx(origin)
            ParticleOps<PP>::keys(R, X.p, Np, lgNf, pk, k1.p, v1.p, st);
            std::vector<uint32_t> k(Np);
            CPM_HIP(hipMemcpyAsync(k.data(), k1.p, sizeof(uint32_t) * Np, hipMemcpyDeviceToHost, st));
            CPM_HIP(hipStreamSynchronize(st));
            const std::vector<uint32_t> ids = host_ids();
            for (long s = 0; s < Np; ++s) keys[ids[s]] = (int32_t)k[s];
        }
        CPM_HIP(hipStreamSynchronize(st));
    }

    void potential(double a, double* phi) override {
        need_state();
        do_deposit((Mesh)mass);
        do_poisson(a);
        mesh_to_host(phi);
    }

    void forces(double a, double* F) override {
        need_state();
        do_deposit((Mesh)mass);
        do_poisson(a);
        gather_to_host(F);
    }

    void gather_mesh(const double* phi, double* F) override {
        need_state();
        check_finite(phi, M, "potential");
        mesh_from_host(phi);
        have_deposit = false;
        gather_to_host(F);
    }

    void gather_to_host(double* F) {
        DevBuf<double> d((size_t)R * Np);
        ParticleOps<PP>::gather_kick(R, X.p, P.p, Np, mesh.p, lgNf, pa, 0.0, 0.0, 0, d.p, Reorder<PP>(), st);
        std::vector<double> tmp((size_t)R * Np);
        CPM_HIP(hipMemcpyAsync(tmp.data(), d.p, sizeof(double) * R * Np, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
        soa_to_aos(tmp, host_ids(), F);
    }

    void poisson_mesh(const double* rho, double a, double* phi) override {
        check_finite(rho, M, "density");
        mesh_from_host(rho);
        if (FX) reduce_sum(mesh.p, M, 0, 1.0 / (double)M, red.p, 4096, scal.p, st);
        do_poisson(a, true);
        mesh_to_host(phi);
    }

    void power_spectrum(int nbins, const double* edges, double* kmean, double* pk, double* cnt) override {
        need_state();
        if (nbins < 1) throw Error(PM_EINVAL, "nbins must be positive");
        do_deposit((Mesh)mass);
        std::vector<double> e(nbins + 1);
        const double kf = 2 * M_PI / cfg.L, kn = M_PI * cfg.Nf / cfg.L;
        if (edges) {
            for (int b = 0; b <= nbins; ++b) e[b] = edges[b];
        } else {
            for (int b = 0; b <= nbins; ++b) e[b] = kf * pow(kn / kf, (double)b / nbins);
        }
        const double* mean_ptr = nullptr;
        if (FX) {
            sum_partials(row_partial.p, npartials, 1.0 / (double)M, red.p, 4096, scal.p, st);
            mean_ptr = scal.p;
        }
        fft.forward(mesh.p, spec.p, nyq.p, (Mesh)1.0, mean_ptr, st);
        DevBuf<double> de(nbins + 1), part(pk_partial_elems(fft.spec_elems, fft.nyq_elems, nbins)), out(3 * nbins);
        CPM_HIP(hipMemcpyAsync(de.p, e.data(), sizeof(double) * (nbins + 1), hipMemcpyHostToDevice, st));
        const double norm = pow(cfg.L, R) / ((double)M * (double)M);
        cpm::power_spectrum<Mesh>(spec.p, nyq.p, (long)fft.spec_elems, (long)fft.nyq_elems, R, lgNf, kf, de.p, nbins,
                                  norm, part.p, out.p, out.p + nbins, out.p + 2 * nbins, st);
        std::vector<double> h(3 * nbins);
        CPM_HIP(hipMemcpyAsync(h.data(), out.p, sizeof(double) * 3 * nbins, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
        for (int b = 0; b < nbins; ++b) {
            if (kmean) kmean[b] = h[b];
            if (pk) pk[b] = h[nbins + b];
            if (cnt) cnt[b] = h[2 * nbins + b];
        }
        have_deposit = false;  // the mesh was consumed
    }

    // power_spectrum with no allocation of any kind: the handle's mesh and
    // spectrum, the estimator's fixed device scratch, host arrays on the stack.
    void power_spectrum_noalloc(int nbins, const double* edges, double* kmean, double* pk, double* cnt) override {
        need_state();
        if (nbins < 1 || nbins > PKS_MAXBINS)
            throw Error(PM_EINVAL, "pm_power_spectrum_noalloc: nbins must lie in [1, " + std::to_string(PKS_MAXBINS) +
                                       "]");
        double e[PKS_MAXBINS + 1], out[3 * PKS_MAXBINS];
        const double kf = 2 * M_PI / cfg.L, kn = M_PI * cfg.Nf / cfg.L;
        for (int b = 0; b <= nbins; ++b) e[b] = edges ? edges[b] : kf * pow(kn / kf, (double)b / nbins);
        do_deposit((Mesh)mass);
        const double* mean_ptr = nullptr;
        if (FX) {
            sum_partials(row_partial.p, npartials, 1.0 / (double)M, red.p, 4096, scal.p, st);
            mean_ptr = scal.p;
        }
        fft.forward(mesh.p, spec.p, nyq.p, (Mesh)1.0, mean_ptr, st);
        const double norm = pow(cfg.L, R) / ((double)M * (double)M);
        cpm::power_spectrum_noalloc<Mesh>(spec.p, nyq.p, (long)fft.spec_elems, (long)fft.nyq_elems, R, lgNf, kf, e,
                                          nbins, norm, out, st);
        for (int b = 0; b < nbins; ++b) {
            if (kmean) kmean[b] = out[b];
            if (pk) pk[b] = out[nbins + b];
            if (cnt) cnt[b] = out[2 * nbins + b];
        }
        have_deposit = false;  // the mesh was consumed
    }

    void diagnostics(pm_diag* d) override {
        need_state();
        std::memset(d, 0, sizeof(*d));
        double h[8] = {0};
        // the deposit of the current positions, its mass, and the potential
        // energy from its spectrum (the mesh is consumed)
        do_deposit((Mesh)mass);
        sum_partials(row_partial.p, npartials, 1.0, red.p, 4096, scal.p + 2, st);
        const double* mean_ptr = nullptr;
        if (FX) {
            sum_partials(row_partial.p, npartials, 1.0 / (double)M, red.p, 4096, scal.p, st);
            mean_ptr = scal.p;
        }
        fft.forward(mesh.p, spec.p, nyq.p, (Mesh)1.0, mean_ptr, st);
        have_deposit = false;
        const long ne = spectrum_energy_partials<Mesh>(spec.p, nyq.p, (long)fft.spec_elems, (long)fft.nyq_elems, R,
                                                      lgNf, 2 * M_PI / cfg.L, row_partial.p, nrows, st);
        sum_partials(row_partial.p, ne, 1.0, red.p, 4096, scal.p + 7, st);
        // The sums run in particle order: the FX storage order within a
        // brick group depends on the order in which the re-sort's atomics
        // placed the particles that changed group, and a sum in that order
        // would differ from run to run in its last bits.
        const double* Pd = P.p;
        DevBuf<double> byid;
        if (bricks) {
            byid.alloc((size_t)R * Np);
            scatter_by_id<double>(P.p, id.p, Np, R, byid.p, st);
            Pd = byid.p;
        }
        for (int k = 0; k < R; ++k) reduce_sum(Pd + (size_t)k * Np, Np, 0, 1.0, red.p, 4096, scal.p + 3 + k, st);
        reduce_sum(Pd, (long)R * Np, 1, 1.0, red.p, 4096, scal.p + 6, st);
        CPM_HIP(hipMemcpyAsync(h, scal.p + 2, sizeof(double) * 6, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
        d->mass = h[0];
        for (int k = 0; k < R; ++k) d->momentum[k] = h[1 + k];
        d->p2 = h[4];
        d->a_pos = last_apos;
        d->a_mom = last_avel;
        // the energies need the scale factors (none yet after a bare pm_set_state)
        if (last_avel > 0) d->kinetic = mass * d->p2 / (2 * last_avel * last_avel);
        if (last_apos > 0) d->potential = -cos.G / (2 * last_apos * (double)M) * h[5];
        for (double v : {d->mass, d->momentum[0], d->momentum[1], d->momentum[2], d->p2, d->kinetic, d->potential, h[5]})
            if (!std::isfinite(v))
                throw Error(PM_ENONFINITE, "the state holds non-finite values: a momentum, position or density sum "
                                           "overflowed");
    }
};

}  // namespace cpm
