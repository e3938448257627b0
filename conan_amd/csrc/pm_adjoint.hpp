// pm_adjoint.hpp -- host interface of the adjoint kernels of the FX PM step
// (pm_adjoint.hip) and of the small vector kernels of the differentiable
// driver.
#pragma once
#include "pm_particles.hpp"

namespace cpm {

// The adjoint state lY = dL/dY, lP = dL/dP is held in FP32 (each update
// computed in FP64 and rounded once).
//
// Reverse of the drift that a fused kick appends, and the drift's adjoint:
//   Y -= round(P * cd_fx)        (exact inverse of the forward's integer step)
//   lP += cd * lY                (lY: dL/dY in 1/length, cd: drift factor)
// then the FX sort keys of the new Y (as keys_fx: full sort when prevkeys is
// null, else incremental), and max |lP| (float bits) into *lmax (bmax:
// scratch of one word per 256 particles).  cd_fx == 0 leaves Y as it is.
void rev_drift_keys(int R, uint32_t* Y, const double* P, const float* lY, float* lP, long np, double cd_fx,
                    double cd, int lgN, const PosArgs& pa, const uint32_t* prevkeys, uint32_t* keys, uint32_t* pids,
                    uint32_t* in_count, uint32_t* out_count, uint32_t* bmax, uint32_t* lmax, hipStream_t st);

// s_k(c) = sum_p lP_k(p) * W(Y_p, c), k < R: the CIC deposit of the R
// components of lP onto R meshes of M cells (s + k * M), accumulated in
// 64-bit fixed point with a scale taken from *lmax (deterministic).
void deposit_fx_vec(int R, const uint32_t* bstart, const uint32_t* perm, const uint32_t* Y, const float* lP,
                    long np, int lgN, const PosArgs& pa, const uint32_t* lmax, float* s, hipStream_t st);

// deposit_fx_vec and, in the same launch, deposit_fx_brick's mass deposit of
// the same positions (mesh and per-brick partials bit-identical to it;
// partial may be null, and then no partials are formed).
void deposit_fx_vm(int R, const uint32_t* bstart, const uint32_t* perm, const uint32_t* Y, const float* lP, long np,
                   int lgN, const PosArgs& pa, const uint32_t* lmax, float* s, float mass, float* mesh,
                   double* partial, hipStream_t st);

// psi = -sum_k D_k s_k with D_k the FP32 4-point derivative of the gather:
// the adjoint of the FD gradient, applied to R meshes.
void neg_div_f32(int R, const float* s, int lgN, double res, float* psi, hipStream_t st);

// Coefficients of the fused adjoint gather (per particle, per component k):
//   P_k  += cF * F_k (on the pq grid)         F = interpolated FD gradient of phi
//   lY_k += cT * T_k + cU * U_k              T_k = sum_i lP_i dF_i/dY_k (phi fixed)
//                                             U_k = sum_c dW(Y, c)/dY_k * chi(c)
struct AdjCoef {
    double cF, cT, cU;
    double pq, ipq;  // momentum quantum of the forward kick and 1 / pq (fx_kick_increment)
};
// The previous step's reverse drift, fused into the adjoint gather's output
// (what rev_drift_keys does, on the reconstructed momentum and the updated
// lY): on != 0 undoes Y -= round(P cd_fx) and adds lP += cd lY.  keys !=
// nullptr writes the incremental re-sort's keys and mover counts (in_count,
// out_count zeroed) of the output positions against the input ones; bmax !=
// nullptr takes max |lP| per workgroup, reduced into *lmax.  hash !=
// nullptr: *hash = the state_hash checksum of (id, Y, P) after the kick's
// undo and before the drift's (the state a checkpoint holds).
struct AdjDrift {
    double cd_fx, cd;
    int on;
    uint32_t *keys, *in_count, *out_count, *bmax;
    unsigned long long* hash;
};
// One workgroup per brick; reads particle perm[s] and writes it, its id and
// its adjoints to position s.  phi == nullptr: F = T = 0; chi == nullptr: U = 0.
void adj_gather_fx_brick(int R, const uint32_t* bstart, const uint32_t* perm, const uint32_t* Y, const double* P,
                         const float* lY, const float* lP, const uint32_t* idi, uint32_t* Yo, double* Po,
                         float* lYo, float* lPo, uint32_t* ido, long np, const float* phi, const float* chi,
                         int lgN, const PosArgs& pa, AdjCoef c, AdjDrift dr, uint32_t* lmax, hipStream_t st);

// *sum = sum over particles of a 64-bit hash of (id, X, P bits), independent
// of the storage order; with Xc != nullptr the state is copied to (Xc, Pc,
// idc) in the same pass.
void state_hash(int R, const uint32_t* X, const double* P, const uint32_t* id, long np, uint32_t* Xc, double* Pc,
                uint32_t* idc, unsigned long long* sum, hipStream_t st);

// dst[k * np + id[s]] = src[k * np + s]  /  dst[k * np + s] = src[k * np + id[s]]
template <typename T>
void scatter_by_id(const T* src, const uint32_t* id, long np, int nc, T* dst, hipStream_t st);
template <typename T>
void gather_by_id(const T* src, const uint32_t* id, long np, int nc, T* dst, hipStream_t st);

// psi = -sum_k D_k u_k on an N^R grid (u: R SoA components) with the FP64
// 4-point derivative of the Zel'dovich start (src/run/numeric.hh:96-101).
void neg_div_f64(int R, const double* u, int lgN, double res, double* psi, hipStream_t st);

void axpby(long n, double a, const double* x, double b, const double* y, double* out, hipStream_t st);
// out = a x + b y in FP64 from FP32 x, y
void axpby_f32(long n, double a, const float* x, double b, const float* y, double* out, hipStream_t st);
// Deterministic FP64 dot product (fixed block tree): *out = sum a[i] * b[i].
void dot(const double* a, const double* b, long n, double* partials, long max_partials, double* out,
         hipStream_t st);
void f32_to_f64(const float* a, long n, double* b, hipStream_t st);
void f64_to_f32(const double* a, long n, float* b, hipStream_t st);
void iota_u32(uint32_t* p, long n, hipStream_t st);
void fill_u32(uint32_t* p, long n, uint32_t v, hipStream_t st);
// Per-block {count(Ya != Yb), max |Pa - Pb|, max |Pb|} over n elements (Y, P
// in the same order); nb blocks, partials[3 * nb].
void state_diff(const uint32_t* Ya, const uint32_t* Yb, const double* Pa, const double* Pb, long n,
                double* partials, int nb, hipStream_t st);

}  // namespace cpm
