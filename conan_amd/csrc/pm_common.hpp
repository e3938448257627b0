// pm_common.hpp -- shared device/host definitions of the MI355X particle-mesh library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace cpm {

// Complex number in interleaved (re, im) layout, aligned for one vector load.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T r, T i) {
    cx<T> z;
    z.re = r;
    z.im = i;
    return z;
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) {
    return mk(a.re + b.re, a.im + b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) {
    return mk(a.re - b.re, a.im - b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> conj(cx<T> a) {
    return mk(a.re, -a.im);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> scale(cx<T> a, T s) {
    return mk(a.re * s, a.im * s);
}

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }

inline int ilog2(long long n) {
    int l = 0;
    while ((1LL << l) < n) ++l;
    return l;
}

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CPM_HIP(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::cpm::Error(-3, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define CPM_LAUNCH_CHECK() CPM_HIP(hipGetLastError())

// Compute units of the current device (cached per device), for grid sizing.
inline long device_cus() {
    static int cus[64] = {0};
    int dev = 0;
    CPM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) dev = 0;
    if (!cus[dev]) CPM_HIP(hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev));
    return cus[dev];
}

}  // namespace cpm
