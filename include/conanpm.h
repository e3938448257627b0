/* conanpm.h -- C ABI of the MI355X particle-mesh library (libconanpm.so).
 *
 * Drop-in boundary for conan's PM path.  The reference has no FFI: its seams
 * are the in-process Solver interface driven by the leapfrog integrator
 *   src/run/leapfrog.hh:6-21   Solver { init, kick, drift, save_snapshot }
 * and the ic / run / cic command lines with their .conan files.  The entry
 * points below replace the Solver implementation Gravity<R>
 * (src/run/nbody.hh:53-219) and the numeric kernels it calls; the conan CLI
 * of this package calls them.  Plain C types only: host pointers in, host
 * pointers out (the library copies; it never keeps a caller's pointer).
 *
 * Conventions: particle arrays are AoS of dim doubles per particle in the
 * reference's order (mass-grid index, axis 0 fastest: src/base/box.hh:51-53);
 * meshes are N^dim doubles, axis 0 fastest.  Every call returns 0 on success
 * and a negative code on failure; pm_last_error() then holds the message.
 * A handle owns one HIP stream and its device buffers; calls on one handle are
 * stream-ordered and must come from one host thread.
 */
#ifndef CONANPM_H
#define CONANPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define PM_OK 0
#define PM_EINVAL -1   /* bad argument or configuration */
#define PM_EUNSUP -2   /* unsupported size or mode */
#define PM_EDEVICE -3  /* HIP error (no device, allocation failure, kernel fault) */
#define PM_ESTATE -4   /* call out of order (e.g. kick before init) */
#define PM_ENONFINITE -5 /* non-finite values in an input */

/* precision forms */
#define PM_FP64 0 /* plain form: FP64 everywhere, reference arithmetic (the drop-in) */
#define PM_FP32 1 /* mixed form: 32-bit fixed-point positions (unwrapped, 2^30 units per box,
                     so [-2L, 2L) is representable), FP32 mesh/FFT/gather, FP64 momenta */

/* flags */
#define PM_FLAG_WRAP_GATHER 1u /* fold positions into [0, L) before the gather, as the
                                  deposit does; without it both forms truncate the unwrapped
                                  coordinate like the reference (src/run/numeric.hh:19,26-28) */
#define PM_FLAG_MIDPOINT 2u    /* drift and kick factors (and the 1/a of the potential) at the
                                  middle a + da/2 of the interval [a, a + da] each integrates
                                  over: second order in da.  Without it they are taken at the
                                  interval's start, as the reference does (src/run/nbody.hh:
                                  110-129, leapfrog.hh:40-41): first order.  No counterpart in
                                  the reference; for runs that want the accuracy of a proper
                                  leapfrog at the same step count. */

#define PM_FLAG_SERIAL_DEPOSIT 4u /* FP64: sum every cell of the deposit in ascending particle
                                    order, the order of the reference's serial loop (md_cic,
                                    src/run/numeric.hh:106-139), so that the density is bit-identical
                                    to it.  Without it the FP64 deposit adds the same weights in
                                    128-bit fixed point (2^-96 units): exact, order-independent and
                                    faster, within an ulp or two of the serial sum.  The FP32 form's
                                    deposit is fixed point in any case. */

/* Configuration of a run: Gravity<R>(mbox, fbox, cosmos) (src/run/nbody.hh:75-86),
 * Cosmology(H0, Omega_m, Omega_L) (src/run/cosmology.hh:11-14). */
typedef struct pm_config {
    int dim;         /* 2 or 3 */
    int Nm;          /* mass grid per side: particles = Nm^dim */
    int Nf;          /* force mesh per side (f * Nm in conan run) */
    double L;        /* box size */
    double H0;       /* Hubble parameter */
    double Omega_m;
    double Omega_L;
    int precision;   /* PM_FP64 or PM_FP32 */
    unsigned flags;  /* PM_FLAG_* */
} pm_config;

typedef struct pm_handle pm_handle;

/* Message of the last failed call on h; h == NULL: of the calling thread's
 * last failed handle-less call (pm_create, pm_ic_*). */
const char* pm_last_error(const pm_handle* h);

/* new Gravity<R> / ~Gravity (src/run/nbody.hh:75-86, 221-231) */
int pm_create(const pm_config* cfg, pm_handle** out);
void pm_destroy(pm_handle* h);

/* Solver::init: Zel'dovich start from the potential on the mass grid
 * (src/run/nbody.hh:88-95, 26-51).  phi: Nm^dim doubles.  LeapFrog passes
 * a_pos = a0 and a_mom = a0 + da/2 (src/run/leapfrog.hh:35). */
int pm_init_zeldovich(pm_handle* h, const double* phi, double a_pos, double a_mom);

/* Solver::drift: X += P * da / (a^2 adot(a))  (src/run/nbody.hh:122-133) */
int pm_drift(pm_handle* h, double a, double da);

/* Solver::kick: deposit, Poisson solve at a, gather, P -= F * da / adot(a)
 * (src/run/nbody.hh:97-120) */
int pm_kick(pm_handle* h, double a, double da);

/* n iterations of LeapFrog::run's loop body (src/run/leapfrog.hh:38-44)
 * starting at scale factor a: drift(a, da); kick(a + da/2, da); a += da.
 * Kick i and drift i+1 run as one fused kernel; results equal the unfused
 * sequence bit for bit.  a_out (nullable) receives the final a. */
int pm_steps(pm_handle* h, double a, double da, int n, double* a_out);

/* State in the reference layout (Array<dVector<R>> X, P; src/run/nbody.hh:65).
 * Positions are unwrapped, as the reference keeps them.  PM_FP32 holds them
 * in [-2L, 2L) (2^30 fixed-point units per box): pm_set_state folds a
 * coordinate outside that range by whole multiples of 4L, and a particle
 * that moves more than L from where it started in [0, L) wraps the same way.
 * Only the sign of a coordinate (the gather's truncation, Q1) depends on it. */
int pm_get_state(pm_handle* h, double* X, double* P);
int pm_set_state(pm_handle* h, const double* X, const double* P);

/* Scale factors of the last drift / kick (save_conan's a-pos / a-mom,
 * src/run/nbody.hh:174-175). */
int pm_get_times(const pm_handle* h, double* a_pos, double* a_mom);

/* PM_FP32: the grid (a power of two) that momenta and kicks are rounded to,
 * chosen from the momenta at pm_init_zeldovich / pm_set_state; it makes
 * kicks exactly reversible.  Setting it pins it: the momenta are rounded to
 * it (no change when they already lie on it), and pm_set_state keeps it
 * instead of choosing one, until pm_init_zeldovich.  A run resumed from a
 * snapshot (conan run --resume) sets it, then the state, and continues bit
 * for bit.  PM_FP64: the quantum is 0 and only 0 is accepted.  No
 * counterpart in the reference (its momenta are plain doubles); for callers
 * that resume runs. */
int pm_get_momentum_quantum(const pm_handle* h, double* pq);
int pm_set_momentum_quantum(pm_handle* h, double pq);

/* CIC deposit of the current positions with particle mass (Nf/Nm)^dim
 * (md_cic, src/run/numeric.hh:106-139): rho (Nf^dim doubles, nullable),
 * keys (Nm^dim cell indices idx(origin), nullable). */
int pm_deposit(pm_handle* h, double* rho, int32_t* keys);

/* Deposit + Poisson solve at scale factor a (src/run/nbody.hh:99-110):
 * the potential phi (Nf^dim doubles). */
int pm_potential(pm_handle* h, double a, double* phi);

/* Deposit + Poisson + gather at a: the interpolated FD gradient F at every
 * particle (LinearInterpolation<Gradient<R>>, src/run/numeric.hh:6-45,72-103),
 * AoS Nm^dim x dim doubles.  State is left unchanged. */
int pm_forces(pm_handle* h, double a, double* F);

/* Gather of a given potential mesh phi (Nf^dim doubles) at the current
 * positions: F as in pm_forces, with no deposit or Poisson solve.  Lets the
 * interpolation be compared with the reference independent of FFT rounding. */
int pm_gather_mesh(pm_handle* h, const double* phi, double* F);

/* Poisson solve of a given density mesh (rho - 1 -> phi at a). */
int pm_poisson_mesh(pm_handle* h, const double* rho, double a, double* phi);

/* Power spectrum of delta = rho - 1 of the current deposit.  edges: nbins+1
 * ascending |k| (physical, 2 pi / L units); NULL = nbins logarithmic bins
 * from 2 pi / L to pi Nf / L.  Outputs per bin: mean |k|, P(k) =
 * L^dim / Nf^(2 dim) <|FFT(delta)_k|^2>, and the mode count. */
int pm_power_spectrum(pm_handle* h, int nbins, const double* edges, double* kmean, double* pk,
                      double* count);

/* pm_power_spectrum without any allocation: no hipMalloc / hipFree, no
 * host heap.  No counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight serving or monitoring loops,
 * harnesses that forbid allocation) and want P(k) every step.  Its scratch
 * (bin edges, the partial sums of a fixed grid of 512 workgroups, results:
 * 1.6 MB) is one file-scope device array per GPU, held until the process
 * ends; so nbins <= 128 (PM_EINVAL beyond).  Same outputs as
 * pm_power_spectrum, summed in another fixed order (equal to ~1e-14).  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU, from
 * any handle or thread, race. */
int pm_power_spectrum_noalloc(pm_handle* h, int nbins, const double* edges, double* kmean, double* pk,
                              double* count);

/* FP64 reductions of the state (no counterpart in the reference), for
 * conservation checks.  m = (Nf/Nm)^dim is the particle mass in units of the
 * mean cell mass, a_pos / a_mom the scale factors of the positions and
 * momenta (pm_get_times).  The energies are those of the Layzer-Irvine
 * equation d[a (K + W)]/da = -K, with P = a^2 dx/dt and phi the peculiar
 * potential the kick uses (nabla^2 phi = G delta / a, src/run/nbody.hh:
 * 99-110): K = m sum_i |P_i|^2 / (2 a_mom^2), W = 1/2 sum_cells rho phi at
 * a_pos, evaluated in Fourier space as -(G / (2 a_pos M)) sum_k |rho_k|^2 /
 * |k|^2 (K, W = 0 while the handle has no scale factors: a state set with
 * pm_set_state before any step).  The call deposits the current positions
 * (one deposit and one forward FFT).  A sum that overflows returns
 * PM_ENONFINITE. */
typedef struct pm_diag {
    double mass;          /* sum of the deposit of the current positions (= Nm^dim * m) */
    double momentum[3];   /* sum_i P_i */
    double p2;            /* sum_i |P_i|^2 */
    double kinetic;       /* K */
    double potential;     /* W */
    double a_pos, a_mom;
} pm_diag;
int pm_diagnostics(pm_handle* h, pm_diag* d);

/* Per-phase device time (HIP events on the handle's stream), accumulated
 * while enabled.  Phases: 0 keys+sort, 1 deposit, 2 FFT rows forward,
 * 3 FFT y forward, 4 Green pass, 5 FFT y backward, 6 FFT rows backward,
 * 7 gather+kick(+drift), 8 drift, 9 reserved (the kx = N/2 plane's column
 * passes now run inside phases 3-5).  Phases 1-7 are one kernel launch each. */
#define PM_NPHASES 10
int pm_profile(pm_handle* h, int enable, double* ms, long long* counts);

/* Synchronise the handle's stream. */
int pm_sync(pm_handle* h);

/* ---------------------------------------------------------------- conan ic
 * _generate_random_field<R> scale-free branch + _compute_potential
 * (src/ic/ic.cc:37-83, 100-117), on the GPU in FP64.  noise: N^dim white
 * noise in linear order (Gaussian_white_noise, src/base/misc.hh:179-188).
 * density (nullable), potential (nullable): N^dim doubles. */
typedef struct pm_ic_params {
    int dim;
    int mbits;        /* N = 2^mbits */
    double L;         /* "size" */
    double ns;        /* power-law slope */
    int smooth;       /* apply the smoothing to the final field too */
    double scale;     /* smoothing scale (Mpc/h) */
    double sigma;     /* normalisation */
    int nyquist_cutoff; /* zero the modes with |k| > pi N / L (beyond the Nyquist sphere) in
                           the filters; no counterpart in the reference (conan ic
                           --nyquist-cutoff); 0: every mode, as the reference */
} pm_ic_params;
int pm_ic_scale_free(const pm_ic_params* p, const double* noise, double* density, double* potential);

/* _generate_random_field<R>, CDM branch (src/ic/ic.cc:84-97 with CDM<R>,
 * ic.cc:14-35): noise filtered by sqrt(P_CDM(k)), physical k, divided by
 * N^dim * res^(dim/2).  density: N^dim doubles.  nyquist_cutoff: as in
 * pm_ic_params. */
int pm_ic_cdm(int dim, int mbits, double L, double ns, double Omega0, double H0, int nyquist_cutoff,
              const double* noise, double* density);

/* _compute_potential<R> alone (src/ic/ic.cc:100-117): in-place. */
int pm_ic_potential(int dim, int mbits, double L, double* field);

/* _compute_displacement<R> (src/ic/ic.cc:119-146): psi AoS N^dim x dim. */
int pm_ic_displacement(int dim, int mbits, double L, const double* potential, double* psi);

/* Unnormalised real-to-complex forward FFT of an N^dim mesh in the library's
 * half-spectrum layout, as interleaved doubles: main block (N/2 x N^(dim-1)
 * modes, kx fastest) followed by the kx = N/2 plane (N^(dim-1) modes).
 * precision: PM_FP64 or PM_FP32 (computed in that precision). */
int pm_fft_r2c(int dim, int N, int precision, const double* in, double* spec);
/* Inverse of pm_fft_r2c, unnormalised (output = N^dim * input field). */
int pm_fft_c2r(int dim, int N, int precision, const double* spec, double* out);

/* ---------------------------------------------------------------- conan cic
 * cic_run<R> (src/cic/cic.hh:49-81): unit-mass deposit of snapshot positions
 * X (Nm^dim particles) on the Nf mesh; s > 1 first sub-samples the Lagrangian
 * grid s times per axis by interpolating displacements.  rho: Nf^dim doubles. */
int pm_cic(int dim, int Nm, int Nf, double L, int s, const double* X, double* rho);

/* ---------------------------------------------------------------- conan fof
 * The grouping of fof_run<R> (src/fof/fof.cc:157-214): friends-of-friends
 * groups of n positions pos (n x dim doubles, as a snapshot's "pos" record
 * holds them, unwrapped), linked when the Proximity predicate holds
 * (fof.cc:93-118: per-axis difference wrapped once by L, squared length <=
 * ll^2).  group[i] (n long longs) = the group of particle i, numbered like
 * the reference's tags minus one: in the order of each group's lowest
 * particle index.  *ngroups = the number of groups.  n < 2^31.
 * pm_fof is the one-shot form of pm_fof_create + pm_fof_run (host memory);
 * its errors are read with pm_fof_last_error(NULL).
 * A pm_fof_plan handle keeps its device buffers between runs (grown on demand) and
 * owns one HIP stream.  device != 0: pos and group are device memory of the
 * current HIP device, complete before the call (the run is ordered on the
 * handle's stream and synchronised before it returns); device == 0: host
 * memory. */
typedef struct pm_fof_plan pm_fof_plan;
typedef struct {
    double device_ms;      /* device time of the last run: from its first kernel to its last */
    long long particles;
    long long ngroups;
    long long candidates;  /* pairs the predicate was evaluated on */
    long long links;       /* pairs it accepted */
    int cells[3];          /* cells per axis of the neighbour grid */
    int table_bits;        /* log2 of the cell-key table size */
} pm_fof_stats;
int pm_fof_create(int dim, double L, double ll, pm_fof_plan** out);
void pm_fof_destroy(pm_fof_plan* f);
const char* pm_fof_last_error(const pm_fof_plan* f);
int pm_fof_run(pm_fof_plan* f, long long n, const double* pos, long long* group, long long* ngroups, int device);
int pm_fof_get_stats(const pm_fof_plan* f, pm_fof_stats* s);
int pm_fof(int dim, long long n, double L, double ll, const double* pos, long long* group, long long* ngroups);

/* ---------------------------------------------------------------- differentiable PM
 * No counterpart in the reference (it has no gradients).  The map
 *   white noise w (N^dim, the input of `conan ic --scale-free`)
 *     -> density and potential (pm_ic_scale_free) -> Zel'dovich start at a0
 *        (pm_init_zeldovich with a_mom = a0 + da/2, as LeapFrog::run,
 *        src/run/leapfrog.hh:36-37) -> nsteps leapfrog steps of da (pm_steps)
 *     -> final CIC density rho on the force mesh (pm_deposit)
 * run in the FP32 form with the gather on folded coordinates
 * (PM_FLAG_WRAP_GATHER: the force is then a smooth function of the position
 * on the torus), and its adjoint: given dL/drho, dL/dw by reverse-mode
 * differentiation with hand-written adjoint kernels (the particles' adjoint
 * state in FP32, each update computed in FP64; the IC part in FP64).  The
 * backward pass integrates the leapfrog backwards (the drift is undone exactly in the
 * fixed-point positions, the kick with the recomputed force) and restores the
 * exact state from the checkpoint stored every ckpt_every steps of the
 * forward pass, instead of storing the trajectory.
 * pm_diff_forward: noise (N^dim doubles) -> rho ((f N)^dim doubles).
 * pm_diff_backward: grad_rho ((f N)^dim) -> grad_noise (N^dim), for the
 *   state of the last pm_diff_forward on this handle.
 * device != 0: the pointers are device memory of the current HIP device
 * (the caller orders its own work before the call; the call returns after
 * the handle's stream is synchronised); device == 0: host memory. */
typedef struct pm_diff_params {
    int dim;          /* 2 or 3 */
    int mbits;        /* particles (and noise) per side N = 2^mbits */
    int f;            /* force mesh per side = f * N */
    double L;         /* box size */
    double ns;        /* scale-free power-law slope */
    int smooth;       /* ic --smooth */
    double scale;     /* ic --scale: smoothing scale */
    double sigma;     /* ic --sigma: normalisation */
    double H0, Omega_m, Omega_L;
    double a0, da;    /* Zel'dovich start and step in a */
    int nsteps;       /* leapfrog steps */
    int ckpt_every;   /* checkpoint interval in steps (>= 1) */
    int precision;    /* PM_FP32 (the only form with an adjoint) */
    int flags;        /* PM_DIFF_* below, or 0 */
} pm_diff_params;

/* pm_diff_params.flags */
/* Keep the potential of every step from the forward pass (nsteps * Nf^dim
   floats of device memory) so that the backward pass does not recompute it. */
#define PM_DIFF_STORE_POTENTIALS 1
/* At every checkpoint, compare the reversed state with the stored one value
   by value (pm_diff_stats.recon_*) and continue from the stored state.
   Without it the states are compared by an order-independent 64-bit
   checksum, and only a mismatch triggers the full comparison and restore. */
#define PM_DIFF_VERIFY_CHECKPOINTS 2

typedef struct pm_diff pm_diff;

typedef struct pm_diff_stats {
    double forward_ms;      /* device time of the last pm_diff_forward (IC, steps, final deposit) */
    double backward_ms;     /* device time of the last pm_diff_backward */
    double device_bytes;    /* device memory held by the handle */
    int ncheckpoints;       /* stored states */
    long long recon_pos;    /* position components of the reversed states that differed from the checkpoints */
    double recon_mom;       /* max |P_reversed - P_checkpoint| / max |P| over the checkpoints */
} pm_diff_stats;

int pm_diff_create(const pm_diff_params* p, pm_diff** out);
void pm_diff_destroy(pm_diff* d);
const char* pm_diff_last_error(const pm_diff* d);
int pm_diff_forward(pm_diff* d, const double* noise, double* rho, int device);
int pm_diff_backward(pm_diff* d, const double* grad_rho, double* grad_noise, int device);
int pm_diff_get_stats(const pm_diff* d, pm_diff_stats* s);

#ifdef __cplusplus
}
#endif

#endif /* CONANPM_H */
