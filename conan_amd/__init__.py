"""conan_amd: conan's particle-mesh step (deposit, FFT Poisson solve, FD-gradient
gather, leapfrog in scale factor, scale-free Zel'dovich initial conditions)
rebuilt for MI355X (gfx950) as hand-written HIP kernels behind a C ABI
(include/conanpm.h, libconanpm.so), with a drop-in `conan` CLI for the
ic / run / cic commands."""

__all__ = ["pm"]
