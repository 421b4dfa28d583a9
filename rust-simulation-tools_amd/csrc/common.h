// common.h -- shared device helpers for the MM-PBSA engine (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace mmpbsa {

struct DeviceError : std::runtime_error {
    int code;
    DeviceError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIP_OK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::mmpbsa::DeviceError(e_ == hipErrorOutOfMemory ? 4 : 3,                \
                                        std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_OK()                                                                       \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess)                                                             \
            throw ::mmpbsa::DeviceError(3, std::string("kernel launch: ") +               \
                                               hipGetErrorString(e_));                    \
    } while (0)

constexpr double kCoulomb = 332.0522;
constexpr double kPi = 3.14159265358979323846;
constexpr int kWave = 64;

// Two-sum based double-double accumulation (error-free transform; no FMA
// contraction is allowed in this library, see the build flags).
struct dd {
    double s, c;
};
__host__ __device__ inline void dd_add(dd &a, double v)
{
    const double t = a.s + v;
    const double bp = t - a.s;
    const double err = (a.s - (t - bp)) + (v - bp);
    a.s = t;
    a.c += err;
}
__host__ __device__ inline void dd_add(dd &a, const dd &b)
{
    dd_add(a, b.s);
    a.c += b.c;
}
__host__ __device__ inline double dd_val(const dd &a) { return a.s + a.c; }

template <typename T>
__device__ inline T wave_sum(T v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline dd wave_sum_dd(dd v)
{
    for (int off = 32; off > 0; off >>= 1) {
        dd o;
        o.s = __shfl_xor(v.s, off, 64);
        o.c = __shfl_xor(v.c, off, 64);
        dd_add(v, o);
    }
    return v;
}

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// A pointer loaded from a descriptor is generic to the compiler, and generic
// (flat) loads also count against the LDS counter that every barrier drains,
// so a prefetch issued before a barrier would be waited for at it.  to_global
// tells the compiler the data is in device memory (global loads).
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *to_global(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// MI355X dispatches workgroups to its 8 XCDs round-robin (flat id mod 8), and
// each XCD has its own L2.  xcd_remap turns the flat workgroup id into a work
// id such that every XCD walks one contiguous range of work ids, so tiles that
// share halo data run on the same L2.  A bijection on [0, total).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_remap(uint32_t flat, uint32_t total)
{
    const uint32_t q = total / kXcds, r = total % kXcds;
    const uint32_t x = flat % kXcds, k = flat / kXcds;
    return x * q + min(x, r) + k;
}

// Unsigned 32-bit division by a runtime-constant divisor with a multiply and a
// shift (Granlund-Montgomery): q = (umulhi(n, m) + n) >> s, exact for every
// n < 2^32.  Grid index arithmetic uses it instead of integer division.
struct FastDiv {
    uint32_t d = 1, m = 0, s = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t div) : d(div)
    {
        s = 0;
        while ((1ull << s) < div) ++s;
        m = (uint32_t)(((1ull << 32) * ((1ull << s) - div)) / div + 1);
    }
    __host__ __device__ inline uint32_t div(uint32_t n) const
    {
#ifdef __HIP_DEVICE_COMPILE__
        const uint32_t hi = __umulhi(n, m);
#else
        const uint32_t hi = (uint32_t)(((uint64_t)n * m) >> 32);
#endif
        return (uint32_t)(((uint64_t)hi + n) >> s);
    }
};

}  // namespace mmpbsa
