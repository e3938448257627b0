// pm_oracle.cpp -- CPU FP64 restatement of conan's particle-mesh path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: the parity checker that
// the GPU library is compared against, and the CPU baseline that bench.py
// times beside it.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  The product (libconanpm.so, the conan CLI)
// never links, loads or calls it.
//
// Every function restates the reference's arithmetic in the reference's own
// operation order, so that with -ffp-contract=off (no FMA, as the reference's
// x86-64 baseline build has none: make:21) the deposit weights, cell keys, FD
// gradient, gather, kick and drift are bit-identical to what the reference
// computes.  The only non-reference piece is the FFT: the reference calls
// FFTW (src/base/fft.cc:15-38, version unpinned, not installed here); this
// file has its own radix-2 complex-to-complex transform with the same
// conventions (unnormalised, forward sign -1, out-of-place in -> out).
//
// Layout conventions (src/base/box.hh:51-53,75-81; src/base/mdrange.hh:106-115):
// axis 0 varies fastest; particle arrays are AoS dVector<R> (R doubles each).
//
// Parity pins (tests/test_oracle.py): DFT closed forms and numpy.fft for the
// FFT; exhaustive integer wraps and special cases for modulus/keys; partition
// of unity, mass conservation and brute force for the deposit; plane-wave
// closed forms for the Poisson solve and the FD gradient; exactness of the
// gather on linear data and deposit/gather adjointness; the constant-momentum
// drift, the zero-force lattice and single-plane-wave Zel'dovich growth for
// the leapfrog; the mt19937 known answer for the noise; brute-force
// connected components for the friends-of-friends groups.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <unordered_map>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cplx;

// ---------------------------------------------------------------- modulus
// src/base/misc.hh:135-139.  For a == -k*b the reference returns b (not 0),
// and the index built from it lies outside the array (undefined behaviour in
// the reference).  wrap_index() folds that one case to 0.
inline int modulus_i(int a, int b) { return a < 0 ? b + (a % b) : (a % b); }
inline int wrap_index(int a, int b) {
    int m = modulus_i(a, b);
    return m == b ? 0 : m;
}
// src/base/misc.hh:144-148
inline double modulus_d(double a, double b) {
    if (a < 0) return a - static_cast<int>(a / b - 1) * b;
    return a - static_cast<int>(a / b) * b;
}

// ---------------------------------------------------------------- Box<R>
// src/base/box.hh:26-82: N, L, res = L/N, stride (axis 0 = 1), idx() with
// periodic wrap; the corner table block[j] has bit k of j as the offset
// along axis k (box.hh:55-57 via mdrange.hh:106-115).
struct Box {
    int R;
    int N;
    double L, res;
    size_t size;
    size_t stride[4];
    Box(int R_, int N_, double L_) : R(R_), N(N_), L(L_), res(L_ / N_) {
        size = 1;
        for (int i = 0; i < R; ++i) {
            stride[i] = size;
            size *= (size_t)N;
        }
    }
    size_t idx(const int* x) const {
        size_t a = 0;
        for (int i = 0; i < R; ++i) a += (size_t)wrap_index(x[i], N) * stride[i];
        return a;
    }
    void decode(size_t i, int* x) const {  // MdRange::operator[] (mdrange.hh:106-115)
        for (int k = 0; k < R; ++k) {
            x[k] = (int)(i % (size_t)N);
            i = (i - (size_t)x[k]) / (size_t)N;
        }
    }
};

inline int corner_bit(int j, int k) { return (j >> k) & 1; }

// ---------------------------------------------------------------- FFT
// Unnormalised complex-to-complex DFT over an N^R array (axis 0 fastest),
// sign -1 forward, +1 backward: the FFTW conventions used by
// src/base/fft.cc:22-28.  Radix-2, iterative, twiddles from long double.
struct FFT1 {
    int N;
    int logN;
    std::vector<cplx> tw[2];  // [0]: sign -1, [1]: sign +1; tw[m] = e^{s 2 pi i m / N}
    std::vector<int> rev;
    explicit FFT1(int N_) : N(N_) {
        logN = 0;
        while ((1 << logN) < N) ++logN;
        for (int s = 0; s < 2; ++s) {
            tw[s].resize(N / 2 > 0 ? N / 2 : 1);
            for (int m = 0; m < N / 2; ++m) {
                long double ang = 2.0L * 3.141592653589793238462643383279502884L * m / N;
                long double c = cosl(ang), sn = sinl(ang);
                tw[s][m] = cplx((double)c, (double)(s == 0 ? -sn : sn));
            }
        }
        rev.resize(N);
        for (int i = 0; i < N; ++i) {
            int r = 0;
            for (int b = 0; b < logN; ++b) r |= ((i >> b) & 1) << (logN - 1 - b);
            rev[i] = r;
        }
    }
    void run(cplx* a, int sgn) const {
        const std::vector<cplx>& w = tw[sgn < 0 ? 0 : 1];
        for (int i = 0; i < N; ++i)
            if (rev[i] > i) std::swap(a[i], a[rev[i]]);
        for (int len = 2; len <= N; len <<= 1) {
            int half = len >> 1, step = N / len;
            for (int b = 0; b < N; b += len)
                for (int j = 0; j < half; ++j) {
                    cplx t = w[j * step] * a[b + j + half];
                    cplx u = a[b + j];
                    a[b + j] = u + t;
                    a[b + j + half] = u - t;
                }
        }
    }
};

void fft_nd(const Box& box, const cplx* in, cplx* out, int sgn, int nthreads) {
    const size_t S = box.size;
    const int N = box.N;
    if (out != in) std::copy(in, in + S, out);
    FFT1 f(N);
    for (int d = 0; d < box.R; ++d) {
        const size_t st = box.stride[d];
        const size_t npencil = S / (size_t)N;
        // pencils with consecutive low indices are copied out and back in
        // groups of G, so the strided axes are read in runs of G elements
        // (the same transform of every pencil)
        const size_t G = st >= 16 ? 16 : 1;
#pragma omp parallel num_threads(nthreads)
        {
            std::vector<cplx> buf((size_t)N * G);
#pragma omp for schedule(static)
            for (long long pg = 0; pg < (long long)(npencil / G); ++pg) {
                // pencils p = pg G + g: low = p % st, high = p / st
                const size_t p0 = (size_t)pg * G;
                const size_t lo = p0 % st, hi = p0 / st;
                const size_t base = lo + hi * st * (size_t)N;
                for (int n = 0; n < N; ++n)
                    for (size_t g = 0; g < G; ++g) buf[g * N + n] = out[base + (size_t)n * st + g];
                for (size_t g = 0; g < G; ++g) f.run(buf.data() + g * N, sgn);
                for (int n = 0; n < N; ++n)
                    for (size_t g = 0; g < G; ++g) out[base + (size_t)n * st + g] = buf[g * N + n];
            }
        }
    }
}

// ---------------------------------------------------------------- k-space
// src/base/fourier.hh:42-55: k_i = (n_i > N/2 ? n_i - N : n_i) * (2 pi / L).
inline double wave(int n, int N, double L) {
    long a = n;
    return (double)((unsigned)n > (unsigned)(N / 2) ? a - (long)N : a) * (2 * M_PI / L);
}
inline void kvec(const Box& b, size_t i, double Lk, double* K) {
    int x[4];
    b.decode(i, x);
    for (int k = 0; k < b.R; ++k) K[k] = wave(x[k], b.N, Lk);
}
inline double ksqr(const double* K, int R) {  // mVector::sqr (mvector.hh:128-134)
    double v = 0;
    for (int k = 0; k < R; ++k) v += K[k] * K[k];
    return v;
}
inline cplx mul_real(cplx z, double f) { return cplx(z.real() * f, z.imag() * f); }

// ---------------------------------------------------------------- deposit
// md_cic: src/run/numeric.hh:106-139 (mass m), md_cic_T: src/cic/cic.hh:8-47.
// Returns the cell key idx(origin) per particle when keys != nullptr.
void md_cic(const Box& box, double* rho, const double* X, size_t np, double m, int nthreads,
            int32_t* keys) {
    std::fill(rho, rho + box.size, 0.0);
    const int R = box.R;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long long ii = 0; ii < (long long)np; ++ii) {
        size_t i = (size_t)ii;
        double p0[4], A[2][4];
        int origin[4];
        for (int k = 0; k < R; ++k) p0[k] = modulus_d(X[i * R + k], box.L) / box.res;
        for (int k = 0; k < R; ++k) {
            origin[k] = (int)(p0[k]);
            A[1][k] = p0[k] - origin[k];
            A[0][k] = 1 - A[1][k];
        }
        if (keys) keys[i] = (int32_t)box.idx(origin);
        for (int j = 0; j < (1 << R); ++j) {
            double z = m;
            int c[4];
            for (int k = 0; k < R; ++k) {
                z *= A[corner_bit(j, k)][k];
                c[k] = origin[k] + corner_bit(j, k);
            }
            size_t idx = box.idx(c);
            if (nthreads > 1) {
#pragma omp critical
                { rho[idx] += z; }
            } else {
                rho[idx] += z;
            }
        }
    }
}

// ---------------------------------------------------------------- gradient
// Gradient<R>::operator[] / gradient_2nd_order: src/run/numeric.hh:48-70,86-102.
inline void fd_gradient(const Box& box, const double* f, size_t n, double* v) {
    int p0[4], q[4];
    box.decode(n, p0);
    for (int k = 0; k < box.R; ++k) {
        size_t ii[4];
        const int off[4] = {-2, -1, 1, 2};
        for (int s = 0; s < 4; ++s) {
            for (int d = 0; d < box.R; ++d) q[d] = p0[d] + (d == k ? off[s] : 0);
            ii[s] = box.idx(q);
        }
        v[k] = f[ii[0]] / 12. - 2 * f[ii[1]] / 3. + 2 * f[ii[2]] / 3. - f[ii[3]] / 12.;
    }
    for (int k = 0; k < box.R; ++k) v[k] = v[k] / box.res;
}

// LinearInterpolation<Gradient<R>,R>::operator(): src/run/numeric.hh:17-44.
// wrap == 0 is the reference (truncation of the unwrapped coordinate, which
// extrapolates for x < 0); wrap == 1 first folds x into [0, L) the way the
// deposit does (numeric.hh:114).
inline void gather_grad(const Box& box, const double* phi, const double* Xi, int wrap, double* v) {
    const int R = box.R;
    double x[4], A[2][4];
    int origin[4];
    for (int k = 0; k < R; ++k) x[k] = (wrap ? modulus_d(Xi[k], box.L) : Xi[k]) / box.res;
    for (int k = 0; k < R; ++k) {
        origin[k] = (int)(x[k]);
        A[1][k] = x[k] - origin[k];
        A[0][k] = 1 - A[1][k];
    }
    for (int k = 0; k < R; ++k) v[k] = 0;
    for (int i = 0; i < (1 << R); ++i) {
        double z = 1;
        int c[4];
        for (int k = 0; k < R; ++k) {
            z *= A[corner_bit(i, k)][k];
            c[k] = origin[k] + corner_bit(i, k);
        }
        double g[4];
        fd_gradient(box, phi, box.idx(c), g);
        for (int k = 0; k < R; ++k) v[k] += g[k] * z;
    }
}

// ---------------------------------------------------------------- cosmology
// src/run/cosmology.hh:6-28
struct Cosmology {
    double H0, Om, OL, Ok, G;
    Cosmology(double H0_, double m_, double L_)
        : H0(H0_), Om(m_), OL(L_), Ok(1 - m_ - L_), G(3. / 2 * m_ * H0_ * H0_) {}
    double adot(double a) const { return H0 * a * sqrt(OL + Om * pow(a, -3) + Ok * pow(a, -2)); }
};

// Poisson part of Gravity::kick (src/run/nbody.hh:103-110): delta holds the
// deposit on entry and the potential phi on exit.
void poisson(const Box& fbox, double* delta, const Cosmology& cos, double a, int nthreads,
             std::vector<cplx>& in, std::vector<cplx>& out) {
    const size_t S = fbox.size;
    for (size_t i = 0; i < S; ++i) in[i] = cplx(delta[i] - 1.0, 0.0);
    fft_nd(fbox, in.data(), out.data(), -1, nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long long ii = 0; ii < (long long)S; ++ii) {
        double K[4];
        kvec(fbox, (size_t)ii, fbox.L, K);
        in[ii] = mul_real(out[ii], -1. / ksqr(K, fbox.R));
    }
    in[0] = 0;
    fft_nd(fbox, in.data(), out.data(), +1, nthreads);
    const double s = fbox.size / cos.G * a;  // real_part(size / G * a), fourier.hh:31-37
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long long ii = 0; ii < (long long)S; ++ii) delta[ii] = out[ii].real() / s;
}

// ---------------------------------------------------------------- the solver
// Gravity<R> (src/run/nbody.hh:53-120) driven by LeapFrog::run
// (src/run/leapfrog.hh:33-46).
struct Sim {
    Box mbox, fbox;
    Cosmology cos;
    double mass;
    int nthreads, wrap;
    // the deposit on one thread (the reference's serial summation order, which
    // the FP64 GPU deposit reproduces) while FFT and particle loops use nthreads
    int serial_deposit = 0;
    std::vector<double> X, P, delta;
    std::vector<cplx> in, out;
    double a, da;
    long step;
    double last_apos, last_avel;
    // factors at the middle of each drift / kick interval instead of its start
    // (not in the reference, which takes the start: nbody.hh:110-129)
    int midpoint = 0;
    Sim(int R, int Nm, int Nf, double L, double H0, double Om, double OL, int nthreads_, int wrap_)
        : mbox(R, Nm, L), fbox(R, Nf, L), cos(H0, Om, OL), nthreads(nthreads_), wrap(wrap_),
          X(mbox.size * R), P(mbox.size * R), delta(fbox.size), in(fbox.size), out(fbox.size),
          a(0), da(0), step(0), last_apos(0), last_avel(0) {
        mass = pow(double(fbox.N) / mbox.N, R);
    }
    // Gravity::init + Zeldovich (nbody.hh:26-51,88-95)
    void init(const double* phi_m, double a_pos, double a_mom) {
        const int R = mbox.R;
        std::vector<double> u(mbox.size * R);
        for (size_t i = 0; i < mbox.size; ++i) fd_gradient(mbox, phi_m, i, &u[i * R]);
        int q[4];
        for (size_t i = 0; i < mbox.size; ++i) {
            mbox.decode(i, q);
            for (int k = 0; k < R; ++k) {
                double G = q[k] * (mbox.L / mbox.N);  // grid_space (box.hh:11-24)
                X[i * R + k] = G - u[i * R + k] * a_pos;
                P[i * R + k] = u[i * R + k] * -a_mom;
            }
        }
        last_apos = a_pos;
        last_avel = a_mom;
    }
    void kick(double ak, double dak) {
        const int R = mbox.R;
        md_cic(fbox, delta.data(), X.data(), mbox.size, mass, serial_deposit ? 1 : nthreads, nullptr);
        const double ae = midpoint ? ak + dak / 2 : ak;
        poisson(fbox, delta.data(), cos, ae, nthreads, in, out);
        double adot = cos.adot(ae);
        double c = dak / adot;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long long ii = 0; ii < (long long)mbox.size; ++ii) {
            double F[4];
            gather_grad(fbox, delta.data(), &X[ii * R], wrap, F);
            for (int k = 0; k < R; ++k) P[ii * R + k] = P[ii * R + k] - F[k] * c;
        }
        last_avel = ak;
    }
    void drift(double ad, double dad) {
        const int R = mbox.R;
        const double ae = midpoint ? ad + dad / 2 : ad;
        double adot = cos.adot(ae);
        double c = dad / (ae * ae * adot);
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long long ii = 0; ii < (long long)(mbox.size * R); ++ii) X[ii] = X[ii] + P[ii] * c;
        last_apos = ad;
    }
};

}  // namespace

// ---------------------------------------------------------------- fof
// Friends-of-friends of src/fof/fof.cc:157-214.  Particles are visited in
// index order; an untagged one starts a flood fill (fof.cc:17-40) that tags
// every particle reachable through the Proximity predicate with the next
// group number.  Proximity (fof.cc:93-118): per axis b = pos[i] - a, wrapped
// once (b > L/2: b - L; b < -L/2: b + L), accepted when the sum of squares,
// accumulated from 0 in axis order (mvector.hh:128-133), is <= r*r.  The
// reference finds the candidates with a kd-tree (fof.cc:171-173,206-208);
// here a periodic grid of cells no smaller than r (the occupied ones in a hash
// map) does, which yields the same neighbour set, since a pair within r lies
// in the same or adjacent cells.
inline bool fof_near(const double* a, const double* b, int R, double L, double rsqr) {
    double v = 0;
    for (int k = 0; k < R; ++k) {
        double t = b[k] - a[k], d;
        if (t > L / 2) d = t - L;
        else if (t < -L / 2) d = t + L;
        else d = t;
        v += d * d;
    }
    return v <= rsqr;
}

static void fof_tags(int R, long n, double L, double r, const double* pos, std::vector<long>& tags, long& ngroups) {
    // cells a little wider than r, so that rounding in a cell index cannot
    // separate a linked pair by two cells; an axis with fewer than 3 cells is
    // one cell (every cell along it is a neighbour anyway)
    uint64_t nc[3];
    double cs[3];
    for (int k = 0; k < 3; ++k) {
        double m = k < R ? std::floor(L / (r * (1 + 1e-9))) : 1.0;
        if (!(m >= 3.0)) m = 1.0;
        if (m > 1048576.0) m = 1048576.0;
        nc[k] = (uint64_t)m;
        cs[k] = L / m;
    }
    auto cell_coord = [&](long i, int k) -> uint64_t {
        if (k >= R) return 0;
        const double x = pos[i * R + k];
        const double w = x - std::floor(x / L) * L;
        long ck = (long)std::floor(w / cs[k]);
        const long m = (long)nc[k];
        return (uint64_t)(((ck % m) + m) % m);
    };
    // particles sorted by cell id; the occupied cells in a hash map
    std::vector<std::pair<uint64_t, long>> byc(n);
    for (long i = 0; i < n; ++i)
        byc[i] = {cell_coord(i, 0) + nc[0] * (cell_coord(i, 1) + nc[1] * cell_coord(i, 2)), i};
    std::sort(byc.begin(), byc.end());
    std::unordered_map<uint64_t, std::pair<long, long>> cells;
    cells.reserve(2 * n + 1);
    for (long q = 0; q < n;) {
        long e = q + 1;
        while (e < n && byc[e].first == byc[q].first) ++e;
        cells[byc[q].first] = {q, e};
        q = e;
    }
    const double rsqr = r * r;
    tags.assign(n, 0);
    long current = 1;
    std::vector<long> stack;
    for (long idx = 0; idx < n; ++idx) {
        if (tags[idx] != 0) continue;
        tags[idx] = current;
        stack.push_back(idx);
        while (!stack.empty()) {
            const long a = stack.back();
            stack.pop_back();
            const uint64_t ca[3] = {cell_coord(a, 0), cell_coord(a, 1), cell_coord(a, 2)};
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int d[3] = {dx, dy, dz};
                        bool skip = false;
                        uint64_t cc = 0, stride = 1;
                        for (int k = 0; k < 3; ++k) {
                            if (nc[k] == 1 && d[k] != 0) skip = true;
                            cc += ((ca[k] + nc[k] + d[k]) % nc[k]) * stride;
                            stride *= nc[k];
                        }
                        if (skip) continue;
                        auto it = cells.find(cc);
                        if (it == cells.end()) continue;
                        for (long q = it->second.first; q < it->second.second; ++q) {
                            const long j = byc[q].second;
                            if (tags[j] != current && fof_near(pos + a * R, pos + j * R, R, L, rsqr)) {
                                tags[j] = current;
                                stack.push_back(j);
                            }
                        }
                    }
        }
        ++current;
    }
    ngroups = current - 1;
}

// ====================================================================== C ABI
extern "C" {

int orc_modulus_i(int a, int b) { return modulus_i(a, b); }
double orc_modulus_d(double a, double b) { return modulus_d(a, b); }

// Gaussian_white_noise(seed) + generate (src/base/misc.hh:179-195, ic.cc:55-57)
void orc_white_noise(unsigned long seed, size_t n, double* out) {
    std::mt19937 rng(seed);
    std::normal_distribution<double> normal;
    for (size_t i = 0; i < n; ++i) out[i] = normal(rng);
}
uint32_t orc_mt19937_nth(unsigned long seed, size_t n) {
    std::mt19937 rng(seed);
    uint32_t v = 0;
    for (size_t i = 0; i < n; ++i) v = (uint32_t)rng();
    return v;
}

// Complex N^R DFT, interleaved (re, im) doubles; sign -1 forward, +1 backward.
void orc_fft(int R, int N, const double* in, double* out, int sign, int nthreads) {
    Box b(R, N, 1.0);
    fft_nd(b, reinterpret_cast<const cplx*>(in), reinterpret_cast<cplx*>(out), sign, nthreads);
}

void orc_cic_keys(int R, int N, double L, const double* X, size_t np, int32_t* keys) {
    Box b(R, N, L);
    std::vector<double> rho(b.size);
    md_cic(b, rho.data(), X, np, 1.0, 1, keys);
}

void orc_md_cic(int R, int N, double L, const double* X, size_t np, double m, double* rho,
                int nthreads) {
    Box b(R, N, L);
    md_cic(b, rho, X, np, m, nthreads, nullptr);
}

// Kick's Poisson solve: rho (deposit, mean 1) -> phi, at scale factor a.
void orc_poisson(int R, int N, double L, double H0, double Om, double OL, double a,
                 const double* rho, double* phi, int nthreads) {
    Box b(R, N, L);
    Cosmology c(H0, Om, OL);
    std::vector<cplx> in(b.size), out(b.size);
    std::copy(rho, rho + b.size, phi);
    poisson(b, phi, c, a, nthreads, in, out);
}

void orc_gradient_2nd_order(int R, int N, double L, const double* f, double* out) {
    Box b(R, N, L);
    for (size_t i = 0; i < b.size; ++i) fd_gradient(b, f, i, out + i * R);
}

void orc_gather(int R, int N, double L, const double* phi, const double* X, size_t np, int wrap,
                double* F) {
    Box b(R, N, L);
    for (size_t i = 0; i < np; ++i) gather_grad(b, phi, X + i * R, wrap, F + i * R);
}

// Zeldovich<R> x()/p() (src/run/nbody.hh:26-51)
void orc_zeldovich(int R, int Nm, double L, const double* phi, double a_pos, double a_mom,
                   double* X, double* P) {
    Sim s(R, Nm, Nm, L, 1, 1, 0, 1, 0);
    s.init(phi, a_pos, a_mom);
    std::copy(s.X.begin(), s.X.end(), X);
    std::copy(s.P.begin(), s.P.end(), P);
}

double orc_adot(double H0, double Om, double OL, double a) { return Cosmology(H0, Om, OL).adot(a); }

// ---- the leapfrog driver as a stateful object (LeapFrog::run, leapfrog.hh:33-46)
void* orc_sim_create(int R, int Nm, int Nf, double L, double H0, double Om, double OL,
                     int nthreads, int wrap) {
    return new Sim(R, Nm, Nf, L, H0, Om, OL, nthreads < 1 ? 1 : nthreads, wrap);
}
void orc_sim_destroy(void* s) { delete static_cast<Sim*>(s); }
void orc_sim_set_midpoint(void* s, int on) { static_cast<Sim*>(s)->midpoint = on; }
void orc_sim_set_serial_deposit(void* s, int on) { static_cast<Sim*>(s)->serial_deposit = on; }
void orc_sim_init(void* sp, const double* phi_m, double a0, double da) {
    Sim* s = static_cast<Sim*>(sp);
    s->init(phi_m, a0, a0 + da / 2);
    s->a = a0;
    s->da = da;
    s->step = 0;
}
// n leapfrog iterations; returns wall seconds of the loop (steady_clock).
double orc_sim_steps(void* sp, int n) {
    Sim* s = static_cast<Sim*>(sp);
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; ++i) {
        s->drift(s->a, s->da);
        s->kick(s->a + s->da / 2, s->da);
        s->a += s->da;
        ++s->step;
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}
void orc_sim_drift(void* sp, double a, double da) { static_cast<Sim*>(sp)->drift(a, da); }
void orc_sim_kick(void* sp, double a, double da) { static_cast<Sim*>(sp)->kick(a, da); }
void orc_sim_get(void* sp, double* X, double* P, double* a_pos_mom) {
    Sim* s = static_cast<Sim*>(sp);
    if (X) std::copy(s->X.begin(), s->X.end(), X);
    if (P) std::copy(s->P.begin(), s->P.end(), P);
    if (a_pos_mom) {
        a_pos_mom[0] = s->last_apos;
        a_pos_mom[1] = s->last_avel;
        a_pos_mom[2] = s->a;
    }
}
void orc_sim_set(void* sp, const double* X, const double* P) {
    Sim* s = static_cast<Sim*>(sp);
    if (X) std::copy(X, X + s->X.size(), s->X.begin());
    if (P) std::copy(P, P + s->P.size(), s->P.begin());
}
// The mesh left by the last kick: the potential phi (nbody.hh:110 stores it in delta).
void orc_sim_mesh(void* sp, double* out) {
    Sim* s = static_cast<Sim*>(sp);
    std::copy(s->delta.begin(), s->delta.end(), out);
}

// ---- initial conditions (src/ic/ic.cc)
// Planted mistakes (ORC_PLANT_*): compiled only into copies of the oracle
// that tests/test_oracle.py builds in a scratch folder to show that the
// closed-form checks of these functions fail on a plausible error.  The
// oracle itself is built without them.
//
// --nyquist-cutoff (no reference counterpart; `conan ic` of this build): the
// initial-condition filters are zeroed for |k| above the Nyquist wave number
// pi N / L (pi in grid units), i.e. outside the sphere inscribed in the cube
// of modes.  Without it, the reference's filters over every mode.

// _generate_random_field<R>, scale-free branch (ic.cc:37-83).
void orc_ic_scale_free(int R, int mbits, double L, double slope, int smooth, double sigma_scale,
                       double norm, int nyq_cut, const double* noise, double* dens, int nthreads) {
    size_t N = size_t(1) << mbits;
    Box b(R, (int)N, L);
    const size_t size = b.size;
    std::vector<cplx> in(size), out(size);
    for (size_t i = 0; i < size; ++i) in[i] = cplx(noise[i], 0.0);
#ifdef ORC_PLANT_SMOOTH_T
    const double t = sigma_scale * L / N;
#else
    const double t = sigma_scale * N / L;  // Fourier::scale<R>(sigma * N/L)
#endif
    auto filt = [&](size_t i, bool withS) {
        double K[4];
        kvec(b, i, (double)N, K);  // kspace<R>(N, N): grid units
        if (nyq_cut && ksqr(K, R) > M_PI * M_PI) return 0.0;
        double P = sqrt(pow(sqrt(ksqr(K, R)), slope));
        if (!withS) return P;
        double v = 1.0;
        for (int k = 0; k < R; ++k) v *= exp(t * (cos(K[k]) - 1));
        return P * v;
    };
    fft_nd(b, in.data(), out.data(), -1, nthreads);
    for (size_t i = 0; i < size; ++i) in[i] = mul_real(out[i], filt(i, true));
    in[0] = 0;
    fft_nd(b, in.data(), out.data(), +1, nthreads);
    for (size_t i = 0; i < size; ++i) in[i] = cplx(noise[i], 0.0);
    for (size_t i = 0; i < size; ++i) dens[i] = out[i].real() / size;
    double acc = 0;
    for (size_t i = 0; i < size; ++i) acc += dens[i] * dens[i];
    double var = acc / size;
    fft_nd(b, in.data(), out.data(), -1, nthreads);
    for (size_t i = 0; i < size; ++i) in[i] = mul_real(out[i], filt(i, smooth != 0));
    in[0] = 0;
    fft_nd(b, in.data(), out.data(), +1, nthreads);
    const double s = size * sqrt(var) / norm;
    for (size_t i = 0; i < size; ++i) dens[i] = out[i].real() / s;
}

// CDM<R> (ic.cc:14-35): sqrt(P(k)), P = A k^ns T0(q)^2 with the zero-baryon
// transfer function T0 = L0 / (L0 + C0 q^2), L0 = ln(2e + 1.8 q),
// C0 = 14.2 + 731 / (1 + 62.5 q), q = k Theta^2 / (Omega0 h),
// Theta = 2.7255 / 2.7, A = 1122670; k physical.
double cdm_amplitude(double k, double ns, double Omega0, double h) {
    const double e = exp(1), Theta_CMB = 2.7255 / 2.7, A = 1122670;
    const double q = k * pow(Theta_CMB, 2) / (Omega0 * h), L0 = log(2 * e + 1.8 * q),
                 C0 = 14.2 + 731.0 / (1 + 62.5 * q), T0 = L0 / (L0 + C0 * pow(q, 2));
    return sqrt(A * pow(k, ns) * pow(T0, 2));
}
extern "C" double orc_cdm_amplitude(double k, double ns, double Omega0, double h) {
    return cdm_amplitude(k, ns, Omega0, h);
}

// _generate_random_field<R>, CDM branch (ic.cc:84-97): the noise filtered by
// sqrt(P_CDM(|k|)) with physical k (kspace<R>(N, L)), the k = 0 mode zeroed,
// divided by size res^(R/2).
void orc_ic_cdm(int R, int mbits, double L, double ns, double Omega0, double H0, int nyq_cut, const double* noise,
                double* dens, int nthreads) {
    size_t N = size_t(1) << mbits;
    Box b(R, (int)N, L);
    const size_t size = b.size;
    const double res = L / N;
    std::vector<cplx> in(size), out(size);
    for (size_t i = 0; i < size; ++i) in[i] = cplx(noise[i], 0.0);
    fft_nd(b, in.data(), out.data(), -1, nthreads);
    const double kn = M_PI * N / L;
    for (size_t i = 0; i < size; ++i) {
        double K[4];
#ifdef ORC_PLANT_CDM_GRIDK
        kvec(b, i, (double)N, K);
#else
        kvec(b, i, L, K);  // kspace<R>(N, L)
#endif
        const double k2 = ksqr(K, R);
        const double f = (nyq_cut && k2 > kn * kn) ? 0.0 : cdm_amplitude(sqrt(k2), ns, Omega0, H0 / 100.0);
        in[i] = mul_real(out[i], f);
    }
    in[0] = 0;
    fft_nd(b, in.data(), out.data(), +1, nthreads);
#ifdef ORC_PLANT_CDM_NORM
    const double s = size;
#else
    const double s = size * pow(res, 0.5 * R);  // real_part(size * pow(res, 0.5*R)), ic.cc:92
#endif
    for (size_t i = 0; i < size; ++i) dens[i] = out[i].real() / s;
}

// _compute_potential<R> (ic.cc:100-117), in place.
void orc_compute_potential(int R, int mbits, double L, double* density, int nthreads) {
    int N = 1 << mbits;
    Box b(R, N, L);
    const size_t size = b.size;
    std::vector<cplx> in(size), out(size);
    for (size_t i = 0; i < size; ++i) in[i] = cplx(density[i], 0.0);
    fft_nd(b, in.data(), out.data(), -1, nthreads);
    for (size_t i = 0; i < size; ++i) {
        double K[4];
        kvec(b, i, L, K);
        in[i] = mul_real(out[i], -1. / ksqr(K, R));
    }
    in[0] = 0;
    fft_nd(b, in.data(), out.data(), +1, nthreads);
    for (size_t i = 0; i < size; ++i) density[i] = out[i].real() / size;
}

// _compute_displacement<R> (ic.cc:119-146): psi AoS (R doubles per node).
void orc_compute_displacement(int R, int mbits, double L, const double* potential, double* psi,
                              int nthreads) {
    int N = 1 << mbits;
    Box b(R, N, L);
    const size_t size = b.size;
    std::vector<cplx> in(size), out(size), phi_f(size);
    for (size_t i = 0; i < size; ++i) in[i] = cplx(potential[i], 0.0);
    fft_nd(b, in.data(), out.data(), -1, nthreads);
    phi_f = out;
    const cplx math_i(0, 1);
    for (int k = 0; k < R; ++k) {
        for (size_t i = 0; i < size; ++i) {
            double K[4];
            kvec(b, i, (double)N, K);
            in[i] = phi_f[i] * (math_i * sin(K[k]));
        }
        fft_nd(b, in.data(), out.data(), +1, nthreads);
        const double s = size / L * N;
        for (size_t i = 0; i < size; ++i) psi[i * R + k] = out[i].real() / s;
    }
}

// ---- conan cic (src/cic/cic.hh:49-81)
// The sub-sampled Lagrangian grid of cic_run<R> with s > 1 (cic.hh:63-77): the
// displacements D = X - G on the mass grid, interpolated (LinearInterpolation,
// numeric.hh:17-44) at the points y of the (s Nm)^R grid: Y = y + D(y), AoS.
void orc_cic_subsample(int R, int Nm, double L, int s, const double* X, double* Y) {
    Box bm(R, Nm, L);
    std::vector<double> D(bm.size * R);
    int q[4];
    for (size_t i = 0; i < bm.size; ++i) {
        bm.decode(i, q);
        for (int k = 0; k < R; ++k) D[i * R + k] = X[i * R + k] - q[k] * (bm.L / bm.N);
    }
    const int Ns = Nm * s;
    Box bs(R, Ns, L);
    for (size_t i = 0; i < bs.size; ++i) {
        bs.decode(i, q);
        double y[4], x[4], A[2][4], v[4] = {0, 0, 0, 0};
        int origin[4];
        for (int k = 0; k < R; ++k) y[k] = q[k] * (L / Ns);
        for (int k = 0; k < R; ++k) x[k] = y[k] / bm.res;
        for (int k = 0; k < R; ++k) {
            origin[k] = (int)(x[k]);
#ifdef ORC_PLANT_CIC_SWAP
            A[0][k] = x[k] - origin[k];
            A[1][k] = 1 - A[0][k];
#else
            A[1][k] = x[k] - origin[k];
            A[0][k] = 1 - A[1][k];
#endif
        }
        for (int j = 0; j < (1 << R); ++j) {
            double z = 1;
            int c[4];
            for (int k = 0; k < R; ++k) {
                z *= A[corner_bit(j, k)][k];
                c[k] = origin[k] + corner_bit(j, k);
            }
            size_t id = bm.idx(c);
            for (int k = 0; k < R; ++k) v[k] += D[id * R + k] * z;
        }
        for (int k = 0; k < R; ++k) Y[i * R + k] = y[k] + v[k];
    }
}

void orc_cic_tool(int R, int Nm, int Nf, double L, int s, const double* X, double* rho) {
    Box bm(R, Nm, L), bf(R, Nf, L);
    if (s == 1) {
        md_cic(bf, rho, X, bm.size, 1.0, 1, nullptr);
        return;
    }
    Box bs(R, Nm * s, L);
    std::vector<double> Y(bs.size * R);
    orc_cic_subsample(R, Nm, L, s, X, Y.data());
    double a = pow(double(s * Nm) / double(Nf), -float(R));
    md_cic(bf, rho, Y.data(), bs.size, a, 1, nullptr);
}

// ---- power spectrum estimator (no reference counterpart; the same estimator
// is implemented on the GPU).  delta: N^R real field.  Bin edges are given
// (nbins+1 ascending |k| values); a mode with 0 < |k| <= edges[nbins] and
// |k| >= edges[0] goes to the last bin b with edges[b] <= |k|.
// P(k) = L^R / N^(2R) * |FFT(delta)_k|^2, averaged per bin.
void orc_power_spectrum(int R, int N, double L, const double* delta, int nbins,
                        const double* edges, double* kmean, double* pk, int64_t* count,
                        int nthreads) {
    Box b(R, N, L);
    std::vector<cplx> in(b.size), out(b.size);
    for (size_t i = 0; i < b.size; ++i) in[i] = cplx(delta[i], 0.0);
    fft_nd(b, in.data(), out.data(), -1, nthreads);
    std::vector<double> ks(nbins, 0.0), ps(nbins, 0.0);
    std::vector<int64_t> cnt(nbins, 0);
    double norm = pow(L, R) / pow((double)b.size, 2);
    for (size_t i = 1; i < b.size; ++i) {
        double K[4];
        kvec(b, i, L, K);
        double k = sqrt(ksqr(K, R));
        if (k < edges[0] || k > edges[nbins]) continue;
        int lo = 0, hi = nbins - 1;
        while (lo < hi) {  // last b with edges[b] <= k
            int mid = (lo + hi + 1) / 2;
            if (edges[mid] <= k) lo = mid; else hi = mid - 1;
        }
        double p2 = std::norm(out[i]);
        ks[lo] += k;
        ps[lo] += p2;
        cnt[lo] += 1;
    }
    for (int bb = 0; bb < nbins; ++bb) {
        count[bb] = cnt[bb];
        kmean[bb] = cnt[bb] ? ks[bb] / cnt[bb] : 0.0;
        pk[bb] = cnt[bb] ? ps[bb] / cnt[bb] * norm : 0.0;
    }
}

// fof_run's grouping (src/fof/fof.cc:175-214): group[i] = tag - 1, numbered
// in the order of each group's lowest particle index; returns the count.
long orc_fof(int R, long n, double L, double r, const double* pos, long* group) {
    std::vector<long> tags;
    long ng = 0;
    fof_tags(R, n, L, r, pos, tags, ng);
    for (long i = 0; i < n; ++i) group[i] = tags[i] - 1;
    return ng;
}

}  // extern "C"
