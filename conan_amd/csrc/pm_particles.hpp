// pm_particles.hpp -- particle kernels of the PM step: host interface and the
// two position forms.
#pragma once
#include "pm_common.hpp"

#include <cmath>

namespace cpm {

// Parameters of the position -> (cell, fraction) maps.
struct PosArgs {
    double L, res;   // box size and mesh spacing (F64)
    double rinv;     // RN(1 / res)
    int fastdiv;     // 1: v / res as div_fma (divides_by_fma(res)); 0: a divide
    double Linv;     // RN(1 / L)
    int fastdivL;    // the same for v / L
    int wrap;        // gather: 0 = truncate the unwrapped coordinate (reference), 1 = fold first
    int fshift;      // FX: FX_BITS - log2 N (bits of the fraction within a cell)
    uint32_t fmask;  // FX: (1 << fshift) - 1
    float fscale;    // FX: 2^-fshift
    int cmask;       // FX: N - 1 (the cell bits above the fraction)
    int brick;       // sort keys: 0 = cell index idx(origin), 1 = brick key (brick_key)
    int lb[3];       // log2 brick size per axis
    int lnb[3];      // log2 bricks per axis
};

// Brick of the cell (x, y, z): axis 0 fastest.
__host__ __device__ __forceinline__ uint32_t brick_of(int x, int y, int z, const PosArgs& a) {
    return (uint32_t)(x >> a.lb[0]) | ((uint32_t)(y >> a.lb[1]) << a.lnb[0]) |
           ((uint32_t)(z >> a.lb[2]) << (a.lnb[0] + a.lnb[1]));
}
// Sort key of an FX particle: brick, then the face class (bit k set when the
// origin cell lies on the brick's upper face along axis k, i.e. when its CIC
// cloud reaches into the next brick along k).
__host__ __device__ __forceinline__ uint32_t brick_key(int x, int y, int z, int R, const PosArgs& a) {
    const uint32_t c = (uint32_t)(((x & ((1 << a.lb[0]) - 1)) == (1 << a.lb[0]) - 1)) |
                       ((uint32_t)(((y & ((1 << a.lb[1]) - 1)) == (1 << a.lb[1]) - 1)) << 1) |
                       ((uint32_t)(R == 3 && ((z & ((1 << a.lb[2]) - 1)) == (1 << a.lb[2]) - 1)) << 2);
    return (brick_of(x, y, z, a) << R) | c;
}


// The 4-point derivative of the reference (src/run/numeric.hh:98,101):
// ((f[-2]/12 - 2 f[-1]/3) + 2 f[+1]/3) - f[+2]/12, then / res.
__device__ __forceinline__ double fd4(double fm2, double fm1, double fp1, double fp2, double res) {
    double v = fm2 / 12. - 2 * fm1 / 3. + 2 * fp1 / 3. - fp2 / 12.;
    return v / res;
}
// x / 3 rounded to nearest, without a division: q0 = x RN(1/3) is within an
// ulp of x / 3, r = x - 3 q0 is exact, and q0 + r RN(1/3) lies within
// 2^-52 ulp of x / 3.  x / 3 is k / 3 ulps past a representable number (k =
// 0, 1, 2), never within 1/6 ulp of a midpoint between two, so rounding
// q0 + r RN(1/3) once gives RN(x / 3) for every x with |x| >= 2^-1018 (and
// zeros, whose sign copysign keeps).  x / 12 = RN((x / 4) / 3) and 2 x / 3 =
// RN((2 x) / 3): scaling by a power of two is exact.
__device__ __forceinline__ double div3(double x) {
    const double c = 0x1.5555555555555p-2;  // RN(1/3)
    const double q0 = x * c;
    const double r = fma(-q0, 3.0, x);
    return __builtin_copysign(fma(r, c, q0), x);
}
// v / b rounded to nearest, from rinv = RN(1/b), without a divide.  Exact
// for every divisor b = m 2^e whose odd significand m is below 2^50
// (divides_by_fma): for m = 1 the quotient is v rinv itself; for m >= 3,
// v / b is A 2^s / m ulps (A the significand of v, s >= 1), a fraction k / m
// of an ulp past a representable number, never within 1/(2m) > 2^-51 ulp of
// a midpoint; q0 = RN(v rinv) lies within 1.5 ulp of v / b, r = v - b q0 is
// formed with one rounding, and q0 + r rinv lies within 3 2^-53 ulp of v / b,
// so rounding it once gives RN(v / b) (normal v, v / b; zeros keep their sign).
__device__ __forceinline__ double div_fma(double v, double b, double rinv) {
    const double q0 = v * rinv;
    const double r = fma(-q0, b, v);
    return __builtin_copysign(fma(r, rinv, q0), v);
}
// Whether div_fma divides by b exactly (the odd part of b's significand
// below 2^50; b positive and normal).
inline bool divides_by_fma(double b) {
    if (!(b > 0) || !std::isnormal(b)) return false;
    int e = 0;
    unsigned long long m = (unsigned long long)std::ldexp(std::frexp(b, &e), 53);
    while (m && !(m & 1ull)) m >>= 1;
    return m < (1ull << 50);
}
// Same operator in FP32, arranged for fewer roundings: (f[-2] - f[+2] + 8 (f[+1] - f[-1])) / (12 res),
// the sum fused with the multiplication by 8 (explicitly, so every kernel that
// uses it rounds the same way whatever its contraction setting).
__device__ __forceinline__ float fd4(float fm2, float fm1, float fp1, float fp2, float inv12res) {
    return fmaf(8.f, fp1 - fm1, fm2 - fp2) * inv12res;
}

// FX kick increment F * ck rounded to a multiple of the momentum quantum pq
// (a power of two).  FX momenta are kept on that grid, so P - dp and P + dp
// are exact and a kick can be undone bit for bit.  ckq = ck / pq: scaling
// by a power of two is exact, so F * ckq is the rounded F * ck over pq.
__host__ __device__ __forceinline__ double fx_kick_increment(double f, double ckq, double pq) {
    return rint(f * ckq) * pq;
}
// Low 32 bits of rint(y) (round to nearest even), for |y| < 2^51: adding
// 1.5 * 2^52 leaves rint(y) + 2^51 in the low mantissa bits.  The FX drift
// adds this to a 32-bit position (y: the step in units of L / 2^30).
__device__ __forceinline__ uint32_t fx_round_u32(double y) {
    return (uint32_t)__double_as_longlong(y + 6755399441055744.0);
}

// ------------------------------------------------------------ position policies
struct PosF64 {
    using Pos = double;
    using Mesh = double;
    // x / L and x / res, by div_fma where that rounds like the divide
    static __device__ __forceinline__ double over_L(double x, const PosArgs& a) {
        return a.fastdivL ? div_fma(x, a.L, a.Linv) : x / a.L;
    }
    static __device__ __forceinline__ double over_res(double x, const PosArgs& a) {
        return a.fastdiv ? div_fma(x, a.res, a.rinv) : x / a.res;
    }
    // modulus(x, L) of src/base/misc.hh:144-148, its quotient from over_L
    static __device__ __forceinline__ double fold(double x, const PosArgs& a) {
        if (x < 0) return x - (double)(int)(over_L(x, a) - 1) * a.L;
        return x - (double)(int)(over_L(x, a)) * a.L;
    }
    // deposit coordinate: wrapped into [0, L) first (src/run/numeric.hh:114-122)
    static __device__ __forceinline__ void dep(Pos x, const PosArgs& a, int& o, double& f) {
        double p0 = over_res(fold(x, a), a);
        o = (int)(p0);
        f = p0 - o;
    }
    // gather coordinate: the reference truncates the unwrapped coordinate
    // (src/run/numeric.hh:19,26-28); wrap != 0 folds it like the deposit.
    static __device__ __forceinline__ void gat(Pos x, const PosArgs& a, int& o, double& f) {
        double p = over_res(a.wrap ? fold(x, a) : x, a);
        o = (int)(p);
        f = p - o;
    }
};

// FX positions: a signed 32-bit fixed-point coordinate of the *unwrapped*
// position in units of L / 2^FX_BITS, so one word spans [-2L, 2L): the low
// FX_BITS bits are the position in the box (cell bits above the fraction),
// the top two bits count box images mod 4.  The periodic wrap is exact
// integer arithmetic (cell = bits masked to the box), and the sign of the
// unwrapped coordinate, which the reference's gather depends on (Q1), is the
// word's sign bit.  A particle stays representable while it moves less than
// L from where it started in [0, L).
constexpr int FX_BITS = 30;
constexpr double FX_UNIT = 1073741824.0;  // 2^FX_BITS positions per box

// Unit coordinate t = x / L -> FX word: t folded into [-2, 2) (whole
// multiples of 4 boxes), then floor(t 2^30), except within 2^-19 of an
// integer, which is taken: an FX position written as a double (q L / 2^30)
// comes back exactly.
__host__ __device__ __forceinline__ uint32_t fx_word(double t) {
    t = t - 4.0 * floor((t + 2.0) * 0.25);
    const double u = t * FX_UNIT, r = rint(u);
    const double v = fabs(u - r) <= 0x1p-19 ? r : floor(u);
    return (uint32_t)(long long)v;  // 2^31 (t = 2 after rounding) wraps to -2^31, the same point
}
// FX word -> unwrapped coordinate in [-2L, 2L)
__host__ __device__ __forceinline__ double fx_coord(uint32_t q, double L) {
    return (double)(int32_t)q * (L / FX_UNIT);
}

struct PosFX {
    using Pos = uint32_t;
    using Mesh = float;
    // deposit: the cell of the folded coordinate and the fraction within it
    // (the reference wraps before the deposit, src/run/numeric.hh:114)
    static __device__ __forceinline__ void dep(Pos x, const PosArgs& a, int& o, float& f) {
        o = (int)(x >> a.fshift) & a.cmask;
        f = (float)(x & a.fmask) * a.fscale;
    }
    // gather: the reference truncates the unwrapped coordinate towards zero
    // (src/run/numeric.hh:19,26-28), so for a negative coordinate with a
    // nonzero fraction the origin is the next cell up and the weight of its
    // upper neighbour is fraction - 1 < 0 (an extrapolation).  wrap != 0
    // folds first, like the deposit.
    static __device__ __forceinline__ void gat(Pos x, const PosArgs& a, int& o, float& f) {
        const uint32_t fr = x & a.fmask;
        o = (int)(x >> a.fshift) & a.cmask;
        if (!a.wrap && (int32_t)x < 0 && fr != 0u) {
            o = (o + 1) & a.cmask;
            f = (float)((int32_t)fr - (int32_t)(a.fmask + 1u)) * a.fscale;  // fraction - 1, rounded once
        } else {
            f = (float)fr * a.fscale;
        }
    }
};

// Optional reordering of the gather + kick: read particle perm[s], write it
// (and its id) to position s of the output arrays.  perm == nullptr: in place.
template <typename PP>
struct Reorder {
    const uint32_t* perm = nullptr;
    typename PP::Pos* Xo = nullptr;
    double* Po = nullptr;
    const uint32_t* idi = nullptr;
    uint32_t* ido = nullptr;
    bool in_place = false;  // perm only orders the threads; particles are updated where they are
};

template <typename PP>
struct ParticleOps {
    using Pos = typename PP::Pos;
    using Mesh = typename PP::Mesh;
    static void keys(int R, const Pos* X, long np, int lgN, const PosArgs& pa, uint32_t* keys, uint32_t* pids,
                     hipStream_t st);
    static void deposit(int R, const uint32_t* row_start, const uint32_t* keys, const uint32_t* pids,
                        const Pos* X, long np, int lgN, const PosArgs& pa, Mesh mass, Mesh* mesh,
                        double* row_partial, long nrows, Mesh* frac, Pos* Xs, uint32_t* inv,
                        hipStream_t st);
    static void gather_kick(int R, Pos* X, double* P, long np, const Mesh* phi, int lgN, const PosArgs& pa,
                            double ck, double cd, int cd_on, double* Fout, const Reorder<PP>& ro, hipStream_t st);
    // FP64 kick (+ drift) in two passes: forces in the deposit's cell order
    // from Xs (positions in that order), then the update in storage order
    // through inv (the inverse order); both from the last deposit.
    static void gather_kick_sorted(int R, Pos* X, double* P, long np, const Mesh* phi, int lgN, const PosArgs& pa,
                                   double ck, double cd, int cd_on, const Pos* Xs, double* Fs, const uint32_t* inv,
                                   hipStream_t st);
    static void drift(Pos* X, const double* P, long n, double cd, hipStream_t st);
    static void zeldovich(int R, const double* phi, int lgNm, double L, double a_pos, double a_mom, Pos* X,
                          double* P, hipStream_t st);
};

void row_bounds(const uint32_t* keys, long np, int lgN, long nrows, uint32_t* row_start, hipStream_t st);
// FX brick keys in storage order: prevkeys == nullptr writes pids = iota for
// a full radix sort; otherwise movers per new (in_count) and old (out_count) group for the
// incremental re-sort against prevkeys (pm_sort.hpp).
void keys_fx(int R, const uint32_t* X, long np, int lgN, const PosArgs& pa, const uint32_t* prevkeys, uint32_t* keys,
             uint32_t* pids, uint32_t* in_count, uint32_t* out_count, hipStream_t st);
// FX brick order: bricks of FXB_X x FXB_Y x FXB_Z cells (2-D: FXB_X x FXB_X),
// clamped to the mesh.  brick_layout() fills pa.brick / lb / lnb.
#ifndef CPM_FXB_X
#define CPM_FXB_X 32
#endif
#ifndef CPM_FXB_Y
#define CPM_FXB_Y 16
#endif
#ifndef CPM_FXB_Z
#define CPM_FXB_Z 8
#endif
constexpr int FXB_X = CPM_FXB_X, FXB_Y = CPM_FXB_Y, FXB_Z = CPM_FXB_Z;
// FP64 brick order (the plain form's brick deposit and tile gather, unless
// PM_FLAG_SERIAL_DEPOSIT): bricks of F64B_X x F64B_Y x F64B_Z cells (2-D:
// F64B_X2 x F64B_X2), so that the gather's ring of FP64 potential planes
// (2 F64B_Z + 5 planes of (F64B_X + 5)(F64B_Y + 5) cells) stays below 64 KiB.
constexpr int F64B_X = 16, F64B_Y = 16, F64B_Z = 4, F64B_X2 = 32;
template <typename PP>
struct BrickDims {
    static constexpr int X = FXB_X, Y = FXB_Y, Z = FXB_Z, X2 = FXB_X;
};
template <>
struct BrickDims<PosF64> {
    static constexpr int X = F64B_X, Y = F64B_Y, Z = F64B_Z, X2 = F64B_X2;
};
// f64: the FP64 brick sizes instead of the FX ones
void brick_layout(int R, int lgN, PosArgs& pa, bool f64 = false);
long brick_count(const PosArgs& pa);
// FX deposit / gather+kick(+drift) over particles grouped by brick (perm:
// sorted position -> storage index; bstart: brick_count + 1 offsets).  The
// deposit writes one FP64 partial sum per brick; MT is float (the step) or
// double (the exact value of the fixed-point accumulators).  The gather
// writes particle s to position s; with keys != nullptr it also produces
// what keys_fx would for the new positions against the keys of the old ones
// (incremental re-sort: keys, in_count, out_count, all counters zeroed).
// defer: np words, all zero on entry and on return: a particle that
// overflows a workgroup's LDS queue of particles kicked from the mesh is
// flagged at its sorted position, and the workgroup kicks and clears its
// flagged particles when its segment is done.
template <typename MT>
void deposit_fx_brick(int R, const uint32_t* bstart, const uint32_t* perm, const uint32_t* X, long np, int lgN,
                      const PosArgs& pa, float mass, MT* mesh, double* partial, hipStream_t st);
void gather_kick_fx_brick(int R, const uint32_t* bstart, const uint32_t* perm, const uint32_t* X, const double* P,
                          const uint32_t* idi, uint32_t* Xo, double* Po, uint32_t* ido, long np, const float* phi,
                          int lgN, const PosArgs& pa, double ck, double pq, double cd, int cd_on, uint32_t* keys,
                          uint32_t* in_count, uint32_t* out_count, uint32_t* defer, hipStream_t st);
// The FP64 brick pipeline: keys_f64_brick is keys_fx for FP64 positions
// (the brick of the deposit's cell, PosF64::dep); deposit_f64_brick is the
// FP64 fixed-point deposit (128-bit accumulators, the row deposit's cell sums
// exactly) per brick, one FP64 partial sum per brick; gather_kick_f64_brick
// is gather_kick_fx_brick with the FP64 potential tile and the reference's
// force, kick and drift arithmetic (k_gather_kick's, operation for operation).
void keys_f64_brick(int R, const double* X, long np, int lgN, const PosArgs& pa, const uint32_t* prevkeys,
                    uint32_t* keys, uint32_t* pids, uint32_t* in_count, uint32_t* out_count, hipStream_t st);
void deposit_f64_brick(int R, const uint32_t* bstart, const uint32_t* perm, const double* X, long np, int lgN,
                       const PosArgs& pa, double mass, double* mesh, double* partial, hipStream_t st);
void gather_kick_f64_brick(int R, const uint32_t* bstart, const uint32_t* perm, const double* X, const double* P,
                           const uint32_t* idi, double* Xo, double* Po, uint32_t* ido, long np, const double* phi,
                           int lgN, const PosArgs& pa, double ck, double cd, int cd_on, uint32_t* keys,
                           uint32_t* in_count, uint32_t* out_count, uint32_t* defer, hipStream_t st);
// max |a[i]| (deterministic), and a[i] := rint(a[i] / q) * q.
void reduce_max_abs(const double* a, long n, double* partials, long max_partials, double* out, hipStream_t st);
void quantize(double* a, long n, double q, hipStream_t st);
template <typename T>
void reduce_sum(const T* a, long n, int square, double scale, double* partials, long max_partials,
                double* out, hipStream_t st);
// Deterministic two-level sum of n partials (scratch: scratch_n doubles, not aliasing).
void sum_partials(const double* partials, long n, double scale, double* scratch, long scratch_n, double* out,
                  hipStream_t st);
void subsample(int R, const double* X, int Nm, int s, double L, double* Y, hipStream_t st);
void fx_to_f64(const uint32_t* q, long n, double L, double* x, hipStream_t st);
void f64_to_fx(const double* x, long n, double L, uint32_t* q, hipStream_t st);

}  // namespace cpm
