// pm_common.hpp -- shared device/host definitions of the MI355X particle-mesh library.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
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

// Device allocations of the library go through device_alloc / device_free.
// The environment variable CONANPM_MAX_DEVICE_BYTES, when set, caps the
// bytes the library holds at once: a request beyond it fails like a
// hipMalloc that runs out of memory (PM_EDEVICE), which lets callers and
// tests exercise that path without exhausting the GPU.
inline std::atomic<unsigned long long>& device_bytes_held() {
    static std::atomic<unsigned long long> held{0};
    return held;
}
inline void device_alloc(void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return;
    if (const char* cap = std::getenv("CONANPM_MAX_DEVICE_BYTES")) {
        const unsigned long long lim = std::strtoull(cap, nullptr, 10);
        if (device_bytes_held().load() + bytes > lim)
            throw Error(-3, "device allocation of " + std::to_string(bytes) + " bytes exceeds CONANPM_MAX_DEVICE_BYTES = " +
                                std::to_string(lim) + " (" + std::to_string(device_bytes_held().load()) + " held)");
    }
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        throw Error(-3, "hipMalloc of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    }
    device_bytes_held() += bytes;
}
inline void device_free(void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    device_bytes_held() -= bytes;
}

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
