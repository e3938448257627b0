// pm_spectral.hpp -- host interface of the k-space filters and the P(k) estimator.
#pragma once
#include "pm_common.hpp"

namespace cpm {

enum FilterKind { FILT_PS = 0, FILT_P = 1, FILT_POT = 2, FILT_DERIV = 3, FILT_CDM = 4 };

struct FilterArgs {
    int kind;
    int dim;
    int lgN;
    double kscale;  // 2 pi / N (grid units) or 2 pi / L (physical)
    double ns;      // power-law slope
    double t;       // smoothing: scale * N / L
    int axis;       // FILT_DERIV axis
    double omega0;  // FILT_CDM: matter density
    double h;       // FILT_CDM: H0 / 100
    double kcut2;   // PS, P, CDM: zero modes with |k|^2 > kcut2 (kscale units); 0: no cutoff
};

void apply_filter(cx<double>* spec, cx<double>* nyq, long nmain, long nnyq, const FilterArgs& fa,
                  hipStream_t st);

long pk_partial_elems(long nmain, long nnyq, int nbins);
// Partial sums (at most max_partials, returned: how many) of sum_k w |z_k|^2 /
// |k|^2 over the half spectrum, k != 0, physical k (kf per index step).
template <typename T>
long spectrum_energy_partials(const cx<T>* spec, const cx<T>* nyq, long nmain, long nnyq, int dim, int lgN, double kf,
                              double* part, long max_partials, hipStream_t st);
// The estimator with no device allocation: its scratch (edges, partials,
// results) is a fixed file-scope device array, so at most PKS_MAXBINS bins.
// edges_host: nbins + 1 host doubles; out_host: 3 nbins host doubles (mean
// |k|, P(k), mode count per bin).  Calls that overlap on one device share
// the array and race.
constexpr int PKS_MAXBINS = 128;
template <typename T>
void power_spectrum_noalloc(const cx<T>* spec, const cx<T>* nyq, long nmain, long nnyq, int dim, int lgN, double kf,
                            const double* edges_host, int nbins, double norm, double* out_host, hipStream_t st);
template <typename T>
void power_spectrum(const cx<T>* spec, const cx<T>* nyq, long nmain, long nnyq, int dim, int lgN, double kf,
                    const double* edges, int nbins, double norm, double* part, double* kmean, double* pk,
                    double* cnt, hipStream_t st);

}  // namespace cpm
