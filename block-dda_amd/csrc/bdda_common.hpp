// bdda_common.hpp -- shared types and helpers for the gfx950 solve path.
//
// Complex values are stored interleaved (re, im) exactly like
// std::complex<T>, so host buffers taken from Eigen::VectorXcd::data() or
// numpy complex128/complex64 arrays can be copied to the device unchanged.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include "../../include/bdda.h"

namespace bdda {

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
    R x, y;
};

using c32 = cplx<float>;
using c64 = cplx<double>;

template <typename R>
__host__ __device__ __forceinline__ cplx<R> mk(R a, R b) { cplx<R> c; c.x = a; c.y = b; return c; }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator-(cplx<R> a) { return mk(-a.x, -a.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator*(cplx<R> a, cplx<R> b)
{
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator*(R s, cplx<R> a) { return mk(s * a.x, s * a.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> conj(cplx<R> a) { return mk(a.x, -a.y); }
// a * (-i) and a * (+i)
template <typename R>
__host__ __device__ __forceinline__ cplx<R> mul_mi(cplx<R> a) { return mk(a.y, -a.x); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> mul_pi(cplx<R> a) { return mk(-a.y, a.x); }
// a += b * c
template <typename R>
__host__ __device__ __forceinline__ void fma_c(cplx<R>& a, cplx<R> b, cplx<R> c)
{
    a.x = a.x + b.x * c.x - b.y * c.y;
    a.y = a.y + b.x * c.y + b.y * c.x;
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> czero() { return mk(R(0), R(0)); }

template <typename Dst, typename Src>
__host__ __device__ __forceinline__ cplx<Dst> cvt(cplx<Src> a) { return mk(Dst(a.x), Dst(a.y)); }

// ---------------------------------------------------------------------
// compile-time trigonometry for butterfly constants (Taylor series after
// reduction to [-pi, pi]); accurate to the last bit of a double for the
// small rational multiples of pi used by radices 3, 5, 7 and 8.
// ---------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double ct_sin(double x)
{
    while (x > kPi) x -= 2 * kPi;
    while (x < -kPi) x += 2 * kPi;
    double term = x, sum = x;
    for (int n = 1; n < 30; ++n) {
        term *= -x * x / ((2 * n) * (2 * n + 1));
        sum += term;
    }
    return sum;
}
constexpr double ct_cos(double x) { return ct_sin(x + kPi / 2); }

// status codes (BDDA_OK, BDDA_EINVAL, ...) are the C ABI's, include/bdda.h
void set_error(const std::string& msg);
const char* last_error();

struct HipError {
    hipError_t code;
    std::string where;
};

#define BDDA_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) throw ::bdda::HipError{e_, std::string(#call)}; \
    } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace bdda
