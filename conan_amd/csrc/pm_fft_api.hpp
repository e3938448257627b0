// pm_fft_api.hpp -- host-side interface of the LDS-pencil FFT (pm_fft.hip).
#pragma once
#include "pm_common.hpp"

namespace cpm {

// A batch of strided pencils: pencil (c, b) starts at c*cs + b*bs and its
// element n lies n*es further; c in [0, nc), b in [0, nb).
struct ColGeom {
    long es;
    long cs;
    long nc;
    long bs;
    int nb;
};

// Which block of the half spectrum a Green's-function pass covers.
struct GreenGeom {
    int dim;
    int nyq;    // 1: the kx = N/2 plane, 0: the main block
    double kf;  // 2 pi / L
};

// One part of a column-pass launch (pm_fft.hip): the pencils of geometry g in
// data, with the Green's-function geometry gg for the Poisson pass; gx
// workgroups per batch, nblk in all.  A launch covers two parts, the main
// block of the half spectrum and the kx = N/2 plane.
template <typename T>
struct ColPart {
    cx<T>* data;
    ColGeom g;
    GreenGeom gg;
    long gx;
    long nblk;
};

template <typename T>
struct FFTPlan {
    int dim, N, H;
    long nrows;         // N^(dim-1)
    long M;             // N^dim
    size_t spec_elems;  // H * nrows
    size_t nyq_elems;   // nrows
    cx<T>* tw;          // tw[m] = exp(-2 pi i m / N), device

    FFTPlan(int dim, int N);
    ~FFTPlan();
    FFTPlan(const FFTPlan&) = delete;
    FFTPlan& operator=(const FFTPlan&) = delete;

    void geoms(ColGeom* ymain, ColGeom* ynyq, ColGeom* zmain, ColGeom* znyq) const;
    // Unnormalised forward transform of (in - shift); mean_ptr, when given,
    // supplies the shift from device memory.
    void forward(const T* in, cx<T>* spec, cx<T>* nyq, T shift, const double* mean_ptr,
                 hipStream_t st) const;
    // Unnormalised backward transform, divided by s; destroys spec/nyq.
    void backward(cx<T>* spec, cx<T>* nyq, T* out, double s, hipStream_t st) const;
    // phi = IFFT[ FFT(rho - shift) * (-1/|k|^2) ] / s, in 2*dim - 1 sweeps.
    // One sweep of poisson(): 0 rows r2c, 1 y forward (3-D), 2 last axis with
    // the Green's function, 3 y backward (3-D), 4 rows c2r; part 0 launches the
    // sweep (main block and kx = N/2 plane together), part 1 launches nothing.
    void poisson_stage(int stage, int part, const T* rho, T* phi, cx<T>* spec, cx<T>* nyq, T shift,
                       const double* mean_ptr, double kf, double s, hipStream_t st) const;
    void poisson(const T* rho, T* phi, cx<T>* spec, cx<T>* nyq, T shift, const double* mean_ptr,
                 double kf, double s, hipStream_t st) const;
};

}  // namespace cpm
