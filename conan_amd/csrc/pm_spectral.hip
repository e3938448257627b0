// pm_spectral.hip -- element-wise k-space filters of the initial conditions and
// the binned power-spectrum estimator, on the half spectrum of pm_fft.hip.
//
// Filters (src/base/fourier.hh:59-128, applied by src/ic/ic.cc:64-115):
//   PS    sqrt(P(|k|)) * S(k), P = |k|^ns, S = prod_i exp(t (cos k_i - 1)), k in grid units
//   P     sqrt(P(|k|))
//   POT   -1/|k|^2, k physical
//   DERIV i sin(k_axis), k in grid units
//   CDM   sqrt(A k^ns T0(k)^2), the BBKS-like transfer function of CDM<R>
//         (ic.cc:14-35), k physical
// The first three zero the k = 0 mode as the reference does after filtering
// (`fft.in[0] = 0`, ic.cc:68,78,113).  With kcut2 > 0 (conan ic
// --nyquist-cutoff, no reference counterpart) PS, P and CDM also zero every
// mode beyond the Nyquist sphere |k| > pi N / L.
#include "pm_spectral.hpp"

#include <mutex>
#include <string>

namespace cpm {

struct ModeIdx {
    int ix, iy, iz;
};

// Element e of the half spectrum: main block first, then the Nyquist plane.
__device__ __forceinline__ ModeIdx mode_of(long e, long nmain, int dim, int lgN) {
    const int N = 1 << lgN, H = N / 2, msk = N - 1;
    ModeIdx m;
    if (e < nmain) {
        m.ix = (int)(e & (H - 1));
        const long r = e >> (lgN - 1);
        m.iy = (int)(r & msk);
        m.iz = dim == 3 ? (int)(r >> lgN) : 0;
    } else {
        const long r = e - nmain;
        m.ix = H;
        m.iy = (int)(r & msk);
        m.iz = dim == 3 ? (int)(r >> lgN) : 0;
    }
    return m;
}

__device__ __forceinline__ double kcomp(int n, int N, double kscale) {
    return (double)(n > N / 2 ? n - N : n) * kscale;
}

__global__ __launch_bounds__(256) void k_filter(cx<double>* __restrict__ spec, cx<double>* __restrict__ nyq,
                                                long nmain, long nnyq, FilterArgs fa) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nmain + nnyq) return;
    const int N = 1 << fa.lgN;
    const ModeIdx m = mode_of(e, nmain, fa.dim, fa.lgN);
    double K[3];
    K[0] = kcomp(m.ix, N, fa.kscale);
    K[1] = kcomp(m.iy, N, fa.kscale);
    K[2] = fa.dim == 3 ? kcomp(m.iz, N, fa.kscale) : 0.0;
    cx<double>* p = e < nmain ? spec + e : nyq + (e - nmain);
    cx<double> z = *p;
    const bool dc = (m.ix == 0 && m.iy == 0 && m.iz == 0);
    if (fa.kind == FILT_DERIV) {
        const double s = sin(K[fa.axis]);
        *p = mk<double>(-z.im * s, z.re * s);
        return;
    }
    if (dc) {
        *p = mk<double>(0, 0);
        return;
    }
    double k2 = 0;
    for (int k = 0; k < fa.dim; ++k) k2 += K[k] * K[k];
    if (fa.kind != FILT_POT && fa.kcut2 > 0 && k2 > fa.kcut2) {  // --nyquist-cutoff
        *p = mk<double>(0, 0);
        return;
    }
    double f;
    if (fa.kind == FILT_POT) {
        f = -1. / k2;
    } else if (fa.kind == FILT_CDM) {
        const double e = exp(1.), Theta = 2.7255 / 2.7, A = 1122670;
        const double k = sqrt(k2);
        const double q = k * pow(Theta, 2) / (fa.omega0 * fa.h);
        const double L0 = log(2 * e + 1.8 * q);
        const double C0 = 14.2 + 731.0 / (1 + 62.5 * q);
        const double T0 = L0 / (L0 + C0 * pow(q, 2));
        f = sqrt(A * pow(k, fa.ns) * pow(T0, 2));
    } else {
        f = sqrt(pow(sqrt(k2), fa.ns));
        if (fa.kind == FILT_PS) {
            double v = 1.0;
            for (int k = 0; k < fa.dim; ++k) v *= exp(fa.t * (cos(K[k]) - 1));
            f = f * v;
        }
    }
    *p = mk<double>(z.re * f, z.im * f);
}

void apply_filter(cx<double>* spec, cx<double>* nyq, long nmain, long nnyq, const FilterArgs& fa,
                  hipStream_t st) {
    const long n = nmain + nnyq;
    hipLaunchKernelGGL(k_filter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, spec, nyq, nmain, nnyq,
                       fa);
    CPM_LAUNCH_CHECK();
}

// ------------------------------------------------------------ power spectrum
// Each block stages PK_CH modes (bin, weighted |delta_k|^2, weighted |k|,
// weight) in LDS; thread b then sums bin b over the chunk in index order.
// Per-block partials are summed over blocks in block order: the result does
// not depend on scheduling.  Interior modes 0 < kx < N/2 stand for k and -k
// (weight 2); the kx = 0 and kx = N/2 planes are complete (weight 1).
static constexpr int PK_NT = 256;
static constexpr int PK_CH = 1024;

template <typename T>
__global__ __launch_bounds__(PK_NT) void k_pk_partial(const cx<T>* __restrict__ spec, const cx<T>* __restrict__ nyq,
                                                      long nmain, long nnyq, int dim, int lgN, double kf,
                                                      const double* __restrict__ edges, int nbins,
                                                      double* __restrict__ part) {
    __shared__ short sb[PK_CH];
    __shared__ double sp[PK_CH];
    __shared__ double sk[PK_CH];
    __shared__ float sw[PK_CH];
    const int N = 1 << lgN;
    const long n = nmain + nnyq;
    const long e0 = (long)blockIdx.x * PK_CH;
    for (int t = threadIdx.x; t < PK_CH; t += PK_NT) {
        const long e = e0 + t;
        short bin = -1;
        double p2 = 0, kk = 0;
        float w = 0;
        if (e < n) {
            const ModeIdx m = mode_of(e, nmain, dim, lgN);
            double K[3];
            K[0] = kcomp(m.ix, N, kf);
            K[1] = kcomp(m.iy, N, kf);
            K[2] = dim == 3 ? kcomp(m.iz, N, kf) : 0.0;
            double k2 = 0;
            for (int k = 0; k < dim; ++k) k2 += K[k] * K[k];
            const double k = sqrt(k2);
            const bool dc = (m.ix == 0 && m.iy == 0 && m.iz == 0);
            if (!dc && k >= edges[0] && k <= edges[nbins]) {
                int lo = 0, hi = nbins - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (edges[mid] <= k) lo = mid;
                    else hi = mid - 1;
                }
                bin = (short)lo;
                const cx<T> z = e < nmain ? spec[e] : nyq[e - nmain];
                w = (m.ix == 0 || m.ix == N / 2) ? 1.f : 2.f;
                const double zr = (double)z.re, zi = (double)z.im;
                p2 = (double)w * (zr * zr + zi * zi);
                kk = (double)w * k;
            }
        }
        sb[t] = bin;
        sp[t] = p2;
        sk[t] = kk;
        sw[t] = w;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += PK_NT) {
        double P = 0, Kk = 0, W = 0;
        for (int t = 0; t < PK_CH; ++t)
            if (sb[t] == b) {
                P += sp[t];
                Kk += sk[t];
                W += (double)sw[t];
            }
        double* o = part + ((long)blockIdx.x * nbins + b) * 3;
        o[0] = P;
        o[1] = Kk;
        o[2] = W;
    }
}

// One block per bin: fixed strided per-thread sums, then a fixed LDS tree.
__global__ __launch_bounds__(PK_NT) void k_pk_final(const double* __restrict__ part, long nblocks, int nbins,
                                                    double norm, double* __restrict__ kmean,
                                                    double* __restrict__ pk, double* __restrict__ cnt) {
    __shared__ double s[3][PK_NT];
    const int b = blockIdx.x;
    double P = 0, K = 0, W = 0;
    for (long i = threadIdx.x; i < nblocks; i += PK_NT) {
        const double* o = part + (i * nbins + b) * 3;
        P += o[0];
        K += o[1];
        W += o[2];
    }
    s[0][threadIdx.x] = P;
    s[1][threadIdx.x] = K;
    s[2][threadIdx.x] = W;
    __syncthreads();
    for (int off = PK_NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 3; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        P = s[0][0];
        K = s[1][0];
        W = s[2][0];
        cnt[b] = W;
        kmean[b] = W > 0 ? K / W : 0.0;
        pk[b] = W > 0 ? P / W * norm : 0.0;
    }
}

long pk_partial_elems(long nmain, long nnyq, int nbins) {
    const long nb = (nmain + nnyq + PK_CH - 1) / PK_CH;
    return nb * nbins * 3;
}

template <typename T>
void power_spectrum(const cx<T>* spec, const cx<T>* nyq, long nmain, long nnyq, int dim, int lgN, double kf,
                    const double* edges, int nbins, double norm, double* part, double* kmean, double* pk,
                    double* cnt, hipStream_t st) {
    if (nbins > 32767) throw Error(-2, "too many power-spectrum bins");
    const long nb = (nmain + nnyq + PK_CH - 1) / PK_CH;
    hipLaunchKernelGGL((k_pk_partial<T>), dim3((unsigned)nb), dim3(PK_NT), 0, st, spec, nyq, nmain, nnyq, dim,
                       lgN, kf, edges, nbins, part);
    CPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pk_final, dim3((unsigned)nbins), dim3(PK_NT), 0, st, (const double*)part,
                       nb, nbins, norm, kmean, pk, cnt);
    CPM_LAUNCH_CHECK();
}

// ------------------------------------------------------------ potential energy
// S = sum over modes k != 0 of w |z_k|^2 / |k|^2 (w = 2 for the interior modes
// 0 < kx < N/2 that stand for k and -k, 1 on the kx = 0 and N/2 planes), k
// physical (kf per index step), in FP64: block b sums the modes
// [b chunk, (b + 1) chunk) with a fixed per-thread stride and a fixed tree.
template <typename T>
__global__ __launch_bounds__(PK_NT) void k_energy_partial(const cx<T>* __restrict__ spec, const cx<T>* __restrict__ nyq,
                                                          long nmain, long nnyq, int dim, int lgN, double kf, long chunk,
                                                          double* __restrict__ part) {
    __shared__ double red[PK_NT];
    const int N = 1 << lgN;
    const long n = nmain + nnyq;
    const long e0 = (long)blockIdx.x * chunk, e1 = e0 + chunk < n ? e0 + chunk : n;
    double acc = 0;
    for (long e = e0 + threadIdx.x; e < e1; e += PK_NT) {
        const ModeIdx m = mode_of(e, nmain, dim, lgN);
        if (m.ix == 0 && m.iy == 0 && m.iz == 0) continue;
        const double kx = kcomp(m.ix, N, kf), ky = kcomp(m.iy, N, kf), kz = dim == 3 ? kcomp(m.iz, N, kf) : 0.0;
        const cx<T> z = e < nmain ? spec[e] : nyq[e - nmain];
        const double zr = (double)z.re, zi = (double)z.im;
        const double w = (m.ix == 0 || m.ix == N / 2) ? 1.0 : 2.0;
        acc += w * (zr * zr + zi * zi) / (kx * kx + ky * ky + kz * kz);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = PK_NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

template <typename T>
long spectrum_energy_partials(const cx<T>* spec, const cx<T>* nyq, long nmain, long nnyq, int dim, int lgN, double kf,
                              double* part, long max_partials, hipStream_t st) {
    const long n = nmain + nnyq;
    long nb = (n + PK_NT - 1) / PK_NT;
    if (nb > max_partials) nb = max_partials;
    const long chunk = (n + nb - 1) / nb;
    nb = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL((k_energy_partial<T>), dim3((unsigned)nb), dim3(PK_NT), 0, st, spec, nyq, nmain, nnyq, dim,
                       lgN, kf, chunk, part);
    CPM_LAUNCH_CHECK();
    return nb;
}
template long spectrum_energy_partials<float>(const cx<float>*, const cx<float>*, long, long, int, int, double, double*,
                                              long, hipStream_t);
template long spectrum_energy_partials<double>(const cx<double>*, const cx<double>*, long, long, int, int, double,
                                               double*, long, hipStream_t);

// ------------------------------------------------------------ P(k) without allocation
// pm_power_spectrum_noalloc keeps everything the estimator needs beyond the
// handle's spectrum in one file-scope device array of fixed size: the bin
// edges, the partial sums of a fixed grid of PKS_BLOCKS workgroups, and the
// per-bin results.  Each device holds its own copy; its address comes from
// hipGetSymbolAddress (pk_scratch), never from the array's name in host code.
static constexpr int PKS_BLOCKS = 512;
static constexpr long PKS_EDGES = 0;
static constexpr long PKS_PART = PKS_EDGES + PKS_MAXBINS + 1;
static constexpr long PKS_OUT = PKS_PART + (long)PKS_BLOCKS * PKS_MAXBINS * 3;
static constexpr long PKS_WORDS = PKS_OUT + 3L * PKS_MAXBINS;
static __device__ __align__(256) double scratch[PKS_WORDS];

double* pk_scratch() {
    static std::mutex mu;
    static double* ptr[64] = {nullptr};
    int dev = 0;
    CPM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw Error(-2, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(-3, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        ptr[dev] = static_cast<double*>(p);
    }
    return ptr[dev];
}

// As k_pk_partial, over a fixed grid: workgroup g takes the chunks g,
// g + gridDim.x, ... in order and sums each bin over them in that order.
template <typename T>
__global__ __launch_bounds__(PK_NT) void k_pk_partial_grid(const cx<T>* __restrict__ spec,
                                                           const cx<T>* __restrict__ nyq, long nmain, long nnyq,
                                                           int dim, int lgN, double kf,
                                                           const double* __restrict__ edges, int nbins,
                                                           double* __restrict__ part) {
    __shared__ short sb[PK_CH];
    __shared__ double sp[PK_CH];
    __shared__ double sk[PK_CH];
    __shared__ float sw[PK_CH];
    const int N = 1 << lgN;
    const long n = nmain + nnyq;
    const long nch = (n + PK_CH - 1) / PK_CH;
    double P = 0, Kk = 0, W = 0;  // bin threadIdx.x (nbins <= PK_NT)
    for (long c = blockIdx.x; c < nch; c += gridDim.x) {
        const long e0 = c * PK_CH;
        for (int t = threadIdx.x; t < PK_CH; t += PK_NT) {
            const long e = e0 + t;
            short bin = -1;
            double p2 = 0, kk = 0;
            float w = 0;
            if (e < n) {
                const ModeIdx m = mode_of(e, nmain, dim, lgN);
                double K[3];
                K[0] = kcomp(m.ix, N, kf);
                K[1] = kcomp(m.iy, N, kf);
                K[2] = dim == 3 ? kcomp(m.iz, N, kf) : 0.0;
                double k2 = 0;
                for (int k = 0; k < dim; ++k) k2 += K[k] * K[k];
                const double k = sqrt(k2);
                const bool dc = (m.ix == 0 && m.iy == 0 && m.iz == 0);
                if (!dc && k >= edges[0] && k <= edges[nbins]) {
                    int lo = 0, hi = nbins - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (edges[mid] <= k) lo = mid;
                        else hi = mid - 1;
                    }
                    bin = (short)lo;
                    const cx<T> z = e < nmain ? spec[e] : nyq[e - nmain];
                    w = (m.ix == 0 || m.ix == N / 2) ? 1.f : 2.f;
                    const double zr = (double)z.re, zi = (double)z.im;
                    p2 = (double)w * (zr * zr + zi * zi);
                    kk = (double)w * k;
                }
            }
            sb[t] = bin;
            sp[t] = p2;
            sk[t] = kk;
            sw[t] = w;
        }
        __syncthreads();
        const int b = threadIdx.x;
        if (b < nbins) {
            double cP = 0, cK = 0, cW = 0;
            for (int t = 0; t < PK_CH; ++t)
                if (sb[t] == b) {
                    cP += sp[t];
                    cK += sk[t];
                    cW += (double)sw[t];
                }
            P += cP;
            Kk += cK;
            W += cW;
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < nbins) {
        double* o = part + ((long)blockIdx.x * nbins + threadIdx.x) * 3;
        o[0] = P;
        o[1] = Kk;
        o[2] = W;
    }
}

template <typename T>
void power_spectrum_noalloc(const cx<T>* spec, const cx<T>* nyq, long nmain, long nnyq, int dim, int lgN, double kf,
                            const double* edges_host, int nbins, double norm, double* out_host, hipStream_t st) {
    if (nbins < 1 || nbins > PKS_MAXBINS) throw Error(-1, "pm_power_spectrum_noalloc: nbins must lie in [1, 128]");
    double* const scr = pk_scratch();
    double* const edges = scr + PKS_EDGES;
    double* const part = scr + PKS_PART;
    double* const out = scr + PKS_OUT;
    CPM_HIP(hipMemcpyAsync(edges, edges_host, sizeof(double) * (nbins + 1), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL((k_pk_partial_grid<T>), dim3(PKS_BLOCKS), dim3(PK_NT), 0, st, spec, nyq, nmain, nnyq, dim, lgN,
                       kf, (const double*)edges, nbins, part);
    CPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pk_final, dim3((unsigned)nbins), dim3(PK_NT), 0, st, (const double*)part, (long)PKS_BLOCKS,
                       nbins, norm, out, out + nbins, out + 2 * nbins);
    CPM_LAUNCH_CHECK();
    CPM_HIP(hipMemcpyAsync(out_host, out, sizeof(double) * 3 * nbins, hipMemcpyDeviceToHost, st));
    CPM_HIP(hipStreamSynchronize(st));
}
template void power_spectrum_noalloc<float>(const cx<float>*, const cx<float>*, long, long, int, int, double,
                                            const double*, int, double, double*, hipStream_t);
template void power_spectrum_noalloc<double>(const cx<double>*, const cx<double>*, long, long, int, int, double,
                                             const double*, int, double, double*, hipStream_t);

template void power_spectrum<float>(const cx<float>*, const cx<float>*, long, long, int, int, double,
                                    const double*, int, double, double*, double*, double*, double*, hipStream_t);
template void power_spectrum<double>(const cx<double>*, const cx<double>*, long, long, int, int, double,
                                     const double*, int, double, double*, double*, double*, double*, hipStream_t);

}  // namespace cpm
