// pm_api.hip -- the C ABI of include/conanpm.h over the PM solver handle
// (pm_impl.hpp), and the stand-alone ic / cic / FFT operations.
#include "pm_impl.hpp"


// ====================================================================== C ABI
struct pm_handle {
    std::unique_ptr<cpm::ImplBase> impl;
    pm_config cfg;
    std::string err;
};

static thread_local std::string g_err;

#define PM_TRY try {
#define PM_CATCH(H)                                           \
    }                                                         \
    catch (const cpm::Error& e) {                             \
        if (H) (H)->err = e.what(); else g_err = e.what();     \
        return e.code;                                        \
    }                                                         \
    catch (const std::exception& e) {                         \
        if (H) (H)->err = e.what(); else g_err = e.what();     \
        return PM_EDEVICE;                                    \
    }                                                         \
    return PM_OK;

#define PM_NEED(H)                                     \
    if (!(H)) {                                        \
        g_err = "null handle";                         \
        return PM_EINVAL;                              \
    }

namespace cpm {
template <typename T>
void fft_test(int dim, int N, const double* in, double* spec, bool forward) {
    cpm::FFTPlan<T> f(dim, N);
    hipStream_t st;
    CPM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    cpm::DevBuf<T> m(f.M);
    cpm::DevBuf<cpm::cx<T>> s(f.spec_elems + f.nyq_elems);
    const size_t ns = f.spec_elems + f.nyq_elems;
    if (forward) {
        std::vector<T> t(f.M);
        for (long i = 0; i < f.M; ++i) t[i] = (T)in[i];
        CPM_HIP(hipMemcpyAsync(m.p, t.data(), sizeof(T) * f.M, hipMemcpyHostToDevice, st));
        f.forward(m.p, s.p, s.p + f.spec_elems, (T)0, nullptr, st);
        std::vector<cpm::cx<T>> o(ns);
        CPM_HIP(hipMemcpyAsync(o.data(), s.p, sizeof(cpm::cx<T>) * ns, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < ns; ++i) {
            spec[2 * i] = o[i].re;
            spec[2 * i + 1] = o[i].im;
        }
    } else {
        std::vector<cpm::cx<T>> o(ns);
        for (size_t i = 0; i < ns; ++i) o[i] = cpm::mk<T>((T)in[2 * i], (T)in[2 * i + 1]);
        CPM_HIP(hipMemcpyAsync(s.p, o.data(), sizeof(cpm::cx<T>) * ns, hipMemcpyHostToDevice, st));
        f.backward(s.p, s.p + f.spec_elems, m.p, 1.0, st);
        std::vector<T> t(f.M);
        CPM_HIP(hipMemcpyAsync(t.data(), m.p, sizeof(T) * f.M, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
        for (long i = 0; i < f.M; ++i) spec[i] = t[i];
    }
    (void)hipStreamDestroy(st);
}

}  // namespace cpm

extern "C" {

const char* pm_last_error(const pm_handle* h) { return h ? h->err.c_str() : g_err.c_str(); }

static void validate(const pm_config* c) {
    if (!c) throw cpm::Error(PM_EINVAL, "null config");
    if (c->dim != 2 && c->dim != 3) throw cpm::Error(PM_EUNSUP, "dim must be 2 or 3");
    if (c->Nm < 1) throw cpm::Error(PM_EINVAL, "Nm must be positive");
    if (c->Nf < 4 || c->Nf > 1024 || (c->Nf & (c->Nf - 1)))
        throw cpm::Error(PM_EUNSUP, "Nf must be a power of two in [4, 1024]");
    if (!(c->L > 0) || !std::isfinite(c->L)) throw cpm::Error(PM_EINVAL, "L must be positive");
    if (c->precision != PM_FP64 && c->precision != PM_FP32) throw cpm::Error(PM_EINVAL, "unknown precision");
    if (c->flags & ~(PM_FLAG_WRAP_GATHER | PM_FLAG_MIDPOINT | PM_FLAG_SERIAL_DEPOSIT))
        throw cpm::Error(PM_EINVAL, "unknown flags");
    double np = std::pow((double)c->Nm, c->dim);
    if (np > 2147483647.0) throw cpm::Error(PM_EUNSUP, "too many particles");
    // FP32 deposit: 64-bit accumulators in units of 2^-40 of the mean cell
    // mass hold less than 2^23, and the quantum of a weight is formed in FP32
    // for weights below 2^23 (a particle's mass is (Nf / Nm)^dim)
    if (c->precision == PM_FP32 && std::pow((double)c->Nf / c->Nm, c->dim) >= 8388608.0)
        throw cpm::Error(PM_EUNSUP, "FP32 form: particle mass (Nf/Nm)^dim must be below 2^23");
    int dev = 0;
    hipError_t e = hipGetDeviceCount(&dev);
    if (e != hipSuccess || dev < 1) throw cpm::Error(PM_EDEVICE, "no HIP device available");
}

int pm_create(const pm_config* cfg, pm_handle** out) {
    pm_handle* h = nullptr;
    PM_TRY
    if (!out) throw cpm::Error(PM_EINVAL, "null output pointer");
    *out = nullptr;
    validate(cfg);
    std::unique_ptr<pm_handle> hh(new pm_handle);
    hh->cfg = *cfg;
    if (cfg->precision == PM_FP64)
        hh->impl.reset(new cpm::Impl<cpm::PosF64>(*cfg));
    else
        hh->impl.reset(new cpm::Impl<cpm::PosFX>(*cfg));
    *out = hh.release();
    PM_CATCH(h)
}

void pm_destroy(pm_handle* h) { delete h; }

int pm_init_zeldovich(pm_handle* h, const double* phi, double a_pos, double a_mom) {
    PM_NEED(h)
    PM_TRY
    if (!phi) throw cpm::Error(PM_EINVAL, "null potential");
    h->impl->init_zeldovich(phi, a_pos, a_mom);
    PM_CATCH(h)
}

// The scale factor must be positive and finite (adot(a) and 1/a enter every
// factor), the step finite.
static void check_times(double a, double da) {
    if (!std::isfinite(a) || !std::isfinite(da)) throw cpm::Error(PM_ENONFINITE, "non-finite scale factor or step");
    if (!(a > 0)) throw cpm::Error(PM_EINVAL, "the scale factor must be positive");
}

int pm_drift(pm_handle* h, double a, double da) {
    PM_NEED(h)
    PM_TRY
    check_times(a, da);
    h->impl->drift(a, da);
    PM_CATCH(h)
}

int pm_kick(pm_handle* h, double a, double da) {
    PM_NEED(h)
    PM_TRY
    check_times(a, da);
    h->impl->kick(a, da);
    PM_CATCH(h)
}

int pm_steps(pm_handle* h, double a, double da, int n, double* a_out) {
    PM_NEED(h)
    PM_TRY
    check_times(a, da);
    if (n > 0) check_times(a + (n - 1) * da, da);
    double r = h->impl->steps(a, da, n);
    if (a_out) *a_out = r;
    PM_CATCH(h)
}

int pm_get_state(pm_handle* h, double* X, double* P) {
    PM_NEED(h)
    PM_TRY
    h->impl->get_state(X, P);
    PM_CATCH(h)
}

int pm_set_state(pm_handle* h, const double* X, const double* P) {
    PM_NEED(h)
    PM_TRY
    if (!X && !h->impl->have_state) throw cpm::Error(PM_EINVAL, "positions required");
    h->impl->set_state(X, P);
    PM_CATCH(h)
}

int pm_get_times(const pm_handle* h, double* a_pos, double* a_mom) {
    if (!h) return PM_EINVAL;
    if (a_pos) *a_pos = h->impl->last_apos;
    if (a_mom) *a_mom = h->impl->last_avel;
    return PM_OK;
}

int pm_get_momentum_quantum(const pm_handle* h, double* pq) {
    if (!h || !pq) return PM_EINVAL;
    *pq = h->impl->momentum_quantum();
    return PM_OK;
}

int pm_set_momentum_quantum(pm_handle* h, double pq) {
    PM_NEED(h)
    PM_TRY
    h->impl->set_momentum_quantum(pq);
    PM_CATCH(h)
}

int pm_deposit(pm_handle* h, double* rho, int32_t* keys) {
    PM_NEED(h)
    PM_TRY
    h->impl->deposit(rho, keys, 0.0);
    PM_CATCH(h)
}

int pm_potential(pm_handle* h, double a, double* phi) {
    PM_NEED(h)
    PM_TRY
    if (!phi) throw cpm::Error(PM_EINVAL, "null output");
    h->impl->potential(a, phi);
    PM_CATCH(h)
}

int pm_forces(pm_handle* h, double a, double* F) {
    PM_NEED(h)
    PM_TRY
    if (!F) throw cpm::Error(PM_EINVAL, "null output");
    h->impl->forces(a, F);
    PM_CATCH(h)
}

int pm_gather_mesh(pm_handle* h, const double* phi, double* F) {
    PM_NEED(h)
    PM_TRY
    if (!phi || !F) throw cpm::Error(PM_EINVAL, "null argument");
    h->impl->gather_mesh(phi, F);
    PM_CATCH(h)
}

int pm_poisson_mesh(pm_handle* h, const double* rho, double a, double* phi) {
    PM_NEED(h)
    PM_TRY
    if (!rho || !phi) throw cpm::Error(PM_EINVAL, "null mesh");
    h->impl->poisson_mesh(rho, a, phi);
    PM_CATCH(h)
}

int pm_power_spectrum(pm_handle* h, int nbins, const double* edges, double* kmean, double* pk, double* count) {
    PM_NEED(h)
    PM_TRY
    h->impl->power_spectrum(nbins, edges, kmean, pk, count);
    PM_CATCH(h)
}

int pm_power_spectrum_noalloc(pm_handle* h, int nbins, const double* edges, double* kmean, double* pk,
                              double* count) {
    PM_NEED(h)
    PM_TRY
    h->impl->power_spectrum_noalloc(nbins, edges, kmean, pk, count);
    PM_CATCH(h)
}

int pm_diagnostics(pm_handle* h, pm_diag* d) {
    PM_NEED(h)
    PM_TRY
    if (!d) throw cpm::Error(PM_EINVAL, "null output");
    h->impl->diagnostics(d);
    PM_CATCH(h)
}

int pm_profile(pm_handle* h, int enable, double* ms, long long* counts) {
    PM_NEED(h)
    PM_TRY
    h->impl->prof.collect();
    if (ms)
        for (int i = 0; i < PM_NPHASES; ++i) ms[i] = h->impl->prof.ms[i];
    if (counts)
        for (int i = 0; i < PM_NPHASES; ++i) counts[i] = h->impl->prof.cnt[i];
    if (enable >= 0) h->impl->prof.enable(enable != 0);
    PM_CATCH(h)
}

int pm_sync(pm_handle* h) {
    PM_NEED(h)
    PM_TRY
    CPM_HIP(hipStreamSynchronize(h->impl->st));
    PM_CATCH(h)
}

// ---------------------------------------------------------------- ic
struct IcCtx {
    int dim, N;
    long M;
    cpm::FFTPlan<double> fft;
    cpm::DevBuf<double> mesh, red, scal;
    cpm::DevBuf<cpm::cx<double>> spec, nyq, spec2, nyq2;
    hipStream_t st = nullptr;
    IcCtx(int d, int mbits) : dim(d), N(1 << mbits), fft(d, 1 << mbits) {
        M = fft.M;
        mesh.alloc(M);
        red.alloc(4096);
        scal.alloc(4);
        spec.alloc(fft.spec_elems);
        nyq.alloc(fft.nyq_elems);
        CPM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    }
    ~IcCtx() {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }
    void upload(const double* a) {
        CPM_HIP(hipMemcpyAsync(mesh.p, a, sizeof(double) * M, hipMemcpyHostToDevice, st));
    }
    void download(double* a) {
        CPM_HIP(hipMemcpyAsync(a, mesh.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
        CPM_HIP(hipStreamSynchronize(st));
    }
    cpm::FilterArgs fargs(int kind, double kscale) const {
        cpm::FilterArgs f{};
        f.kind = kind;
        f.dim = dim;
        f.lgN = cpm::ilog2(N);
        f.kscale = kscale;
        return f;
    }
    // _compute_potential on the mesh (ic.cc:100-117)
    void potential(double L) {
        fft.forward(mesh.p, spec.p, nyq.p, 0.0, nullptr, st);
        cpm::apply_filter(spec.p, nyq.p, (long)fft.spec_elems, (long)fft.nyq_elems, fargs(cpm::FILT_POT, 2 * M_PI / L), st);
        fft.backward(spec.p, nyq.p, mesh.p, (double)M, st);
    }
};

static void ic_check(int dim, int mbits) {
    if (dim != 2 && dim != 3) throw cpm::Error(PM_EUNSUP, "only 2 and 3 dimensions supported.");
    if (mbits < 2 || mbits > 10) throw cpm::Error(PM_EUNSUP, "mbits must lie in [2, 10]");
    int dev = 0;
    if (hipGetDeviceCount(&dev) != hipSuccess || dev < 1) throw cpm::Error(PM_EDEVICE, "no HIP device available");
}

int pm_ic_scale_free(const pm_ic_params* p, const double* noise, double* density, double* potential) {
    pm_handle* h = nullptr;
    PM_TRY
    if (!p || !noise) throw cpm::Error(PM_EINVAL, "null argument");
    ic_check(p->dim, p->mbits);
    IcCtx c(p->dim, p->mbits);
    cpm::check_finite(noise, c.M, "noise");
    const size_t N = size_t(1) << p->mbits;
    c.spec2.alloc(c.fft.spec_elems);
    c.nyq2.alloc(c.fft.nyq_elems);
    c.upload(noise);
    c.fft.forward(c.mesh.p, c.spec.p, c.nyq.p, 0.0, nullptr, c.st);
    CPM_HIP(hipMemcpyAsync(c.spec2.p, c.spec.p, sizeof(cpm::cx<double>) * c.fft.spec_elems, hipMemcpyDeviceToDevice, c.st));
    CPM_HIP(hipMemcpyAsync(c.nyq2.p, c.nyq.p, sizeof(cpm::cx<double>) * c.fft.nyq_elems, hipMemcpyDeviceToDevice, c.st));
    cpm::FilterArgs fa = c.fargs(cpm::FILT_PS, 2 * M_PI / (double)N);  // kspace<R>(N, N)
    fa.ns = p->ns;
    fa.t = p->scale * N / p->L;  // Fourier::scale<R>(sigma * N/L), ic.cc:63
    fa.kcut2 = p->nyquist_cutoff ? M_PI * M_PI : 0.0;  // grid units: |K| > pi
    cpm::apply_filter(c.spec.p, c.nyq.p, (long)c.fft.spec_elems, (long)c.fft.nyq_elems, fa, c.st);
    c.fft.backward(c.spec.p, c.nyq.p, c.mesh.p, (double)c.M, c.st);
    cpm::reduce_sum(c.mesh.p, c.M, 1, 1.0 / (double)c.M, c.red.p, 4096, c.scal.p, c.st);
    double var = 0;
    CPM_HIP(hipMemcpyAsync(&var, c.scal.p, sizeof(double), hipMemcpyDeviceToHost, c.st));
    CPM_HIP(hipStreamSynchronize(c.st));
    fa.kind = p->smooth ? cpm::FILT_PS : cpm::FILT_P;
    cpm::apply_filter(c.spec2.p, c.nyq2.p, (long)c.fft.spec_elems, (long)c.fft.nyq_elems, fa, c.st);
    const double s = c.M * sqrt(var) / p->sigma;  // real_part(size * sqrt(var) / norm), ic.cc:80
    c.fft.backward(c.spec2.p, c.nyq2.p, c.mesh.p, s, c.st);
    if (density) c.download(density);
    if (potential) {
        c.potential(p->L);
        c.download(potential);
    }
    CPM_HIP(hipStreamSynchronize(c.st));
    PM_CATCH(h)
}

int pm_ic_cdm(int dim, int mbits, double L, double ns, double Omega0, double H0, int nyquist_cutoff,
              const double* noise, double* density) {
    pm_handle* h = nullptr;
    PM_TRY
    if (!noise || !density) throw cpm::Error(PM_EINVAL, "null argument");
    ic_check(dim, mbits);
    IcCtx c(dim, mbits);
    cpm::check_finite(noise, c.M, "noise");
    const size_t N = size_t(1) << mbits;
    const double res = L / N;
    c.upload(noise);
    c.fft.forward(c.mesh.p, c.spec.p, c.nyq.p, 0.0, nullptr, c.st);
    cpm::FilterArgs fa = c.fargs(cpm::FILT_CDM, 2 * M_PI / L);  // kspace<R>(N, L), ic.cc:84
    fa.ns = ns;
    fa.omega0 = Omega0;
    fa.h = H0 / 100.0;
    fa.kcut2 = nyquist_cutoff ? (M_PI * N / L) * (M_PI * N / L) : 0.0;
    cpm::apply_filter(c.spec.p, c.nyq.p, (long)c.fft.spec_elems, (long)c.fft.nyq_elems, fa, c.st);
    c.fft.backward(c.spec.p, c.nyq.p, c.mesh.p, c.M * pow(res, 0.5 * dim), c.st);  // ic.cc:92
    c.download(density);
    PM_CATCH(h)
}

int pm_ic_potential(int dim, int mbits, double L, double* field) {
    pm_handle* h = nullptr;
    PM_TRY
    if (!field) throw cpm::Error(PM_EINVAL, "null argument");
    ic_check(dim, mbits);
    IcCtx c(dim, mbits);
    cpm::check_finite(field, c.M, "density");
    c.upload(field);
    c.potential(L);
    c.download(field);
    PM_CATCH(h)
}

int pm_ic_displacement(int dim, int mbits, double L, const double* potential, double* psi) {
    pm_handle* h = nullptr;
    PM_TRY
    if (!potential || !psi) throw cpm::Error(PM_EINVAL, "null argument");
    ic_check(dim, mbits);
    IcCtx c(dim, mbits);
    cpm::check_finite(potential, c.M, "potential");
    c.spec2.alloc(c.fft.spec_elems);
    c.nyq2.alloc(c.fft.nyq_elems);
    c.upload(potential);
    c.fft.forward(c.mesh.p, c.spec2.p, c.nyq2.p, 0.0, nullptr, c.st);
    const int N = 1 << mbits;
    std::vector<double> comp(c.M);
    for (int k = 0; k < dim; ++k) {
        CPM_HIP(hipMemcpyAsync(c.spec.p, c.spec2.p, sizeof(cpm::cx<double>) * c.fft.spec_elems, hipMemcpyDeviceToDevice, c.st));
        CPM_HIP(hipMemcpyAsync(c.nyq.p, c.nyq2.p, sizeof(cpm::cx<double>) * c.fft.nyq_elems, hipMemcpyDeviceToDevice, c.st));
        cpm::FilterArgs fa = c.fargs(cpm::FILT_DERIV, 2 * M_PI / (double)N);
        fa.axis = k;
        cpm::apply_filter(c.spec.p, c.nyq.p, (long)c.fft.spec_elems, (long)c.fft.nyq_elems, fa, c.st);
        c.fft.backward(c.spec.p, c.nyq.p, c.mesh.p, c.M / L * N, c.st);  // real_part(size / L * N), ic.cc:142
        c.download(comp.data());
        for (long i = 0; i < c.M; ++i) psi[i * dim + k] = comp[i];
    }
    PM_CATCH(h)
}

int pm_fft_r2c(int dim, int N, int precision, const double* in, double* spec) {
    pm_handle* h = nullptr;
    PM_TRY
    if (!in || !spec) throw cpm::Error(PM_EINVAL, "null argument");
    ic_check(dim, cpm::ilog2(N));
    if (precision == PM_FP64) cpm::fft_test<double>(dim, N, in, spec, true);
    else cpm::fft_test<float>(dim, N, in, spec, true);
    PM_CATCH(h)
}

int pm_fft_c2r(int dim, int N, int precision, const double* spec, double* out) {
    pm_handle* h = nullptr;
    PM_TRY
    if (!spec || !out) throw cpm::Error(PM_EINVAL, "null argument");
    ic_check(dim, cpm::ilog2(N));
    if (precision == PM_FP64) cpm::fft_test<double>(dim, N, spec, out, false);
    else cpm::fft_test<float>(dim, N, spec, out, false);
    PM_CATCH(h)
}

// ---------------------------------------------------------------- cic
int pm_cic(int dim, int Nm, int Nf, double L, int s, const double* X, double* rho) {
    pm_handle* h = nullptr;
    PM_TRY
    if (!X || !rho) throw cpm::Error(PM_EINVAL, "null argument");
    if (s < 1) throw cpm::Error(PM_EINVAL, "sub-sample must be >= 1");
    pm_config c{};
    c.dim = dim;
    c.Nm = Nm * s;
    c.Nf = Nf;
    c.L = L;
    c.H0 = 1;
    c.Omega_m = 1;
    c.precision = PM_FP64;
    c.flags = PM_FLAG_SERIAL_DEPOSIT;  // md_cic_T's serial order: bit-identical density
    validate(&c);
    cpm::Impl<cpm::PosF64> im(c);
    if (s == 1) {
        im.set_state(X, nullptr);
        im.deposit(rho, nullptr, 1.0);  // md_cic_T with unit mass (cic.hh:59-63)
    } else {
        long npm = 1;
        for (int k = 0; k < dim; ++k) npm *= Nm;
        std::vector<double> soa((size_t)dim * npm);
        cpm::check_finite(X, (size_t)dim * npm, "positions");
        for (long i = 0; i < npm; ++i)
            for (int k = 0; k < dim; ++k) soa[(size_t)k * npm + i] = X[i * dim + k];
        cpm::DevBuf<double> dx(soa.size()), dy((size_t)dim * im.Np);
        CPM_HIP(hipMemcpyAsync(dx.p, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice, im.st));
        cpm::subsample(dim, dx.p, Nm, s, L, dy.p, im.st);
        im.set_positions_device(dy.p);
        const double a = pow(double(s * Nm) / double(Nf), -float(dim));  // cic.hh:76
        im.deposit(rho, nullptr, a);
    }
    PM_CATCH(h)
}

}  // extern "C"
