// builtin_kernels.hpp -- device functors for the FMMA_KERNEL_* built-ins.
//
// Each follows the formula of the same id in oracle/fmm_oracle.c and
// include/fmma_hip.h.  1/r uses the hardware reciprocal square root
// (relative error below 5.3e-8 in FP64, measured on gfx950) refined by one
// Newton step y = y0 (1 + e/2), e = 1 - r2 y0^2, which leaves at most 4.1e-15
// (measured over 1e-8 <= r2 <= 1e2).
//
// `symmetry` is +1 when f(y,x) = f(x,y) and -1 when f(y,x) = -f(x,y), exactly
// in floating point too (x - y and y - x differ only in sign); P2P uses it
// to evaluate each pair of distinct neighbour boxes once.
//
// grad(x, y, g) writes grad_x f(x, y), the near-field part of the FMM
// gradient; grad_far() is its form without the r = 0 case where there is one.
//
// Kernels that special-case r = 0 also provide far(): the same function
// without that test, which P2P uses for the 3^D - 1 neighbour boxes other
// than the target's own box (points in different boxes never coincide).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace fmma {
namespace lib {

__device__ __forceinline__ double rsqrt_refined(double r2)
{
    const double y = __builtin_amdgcn_rsq(r2);
    const double h = r2 * y;
    const double e = __builtin_fma(-h, y, 1.0);
    return __builtin_fma(y * e, 0.5, y);
}

__device__ __forceinline__ float rsqrt_refined(float r2)
{
    const float y = __builtin_amdgcn_rsqf(r2);
    const float h = r2 * y;
    const float e = __builtin_fmaf(-h, y, 1.0f);
    return __builtin_fmaf(y * e, 0.5f, y);
}

// 2/sqrt(r2): the refined value of rsqrt_refined times 2, as
// y0 (3 - r2 y0^2), one instruction fewer; the pair loops divide the
// weights by the factor 2 (pair_scale).
__device__ __forceinline__ double rsqrt2_refined(double r2)
{
    const double y = __builtin_amdgcn_rsq(r2);
    const double h = r2 * y;
    return y * __builtin_fma(-h, y, 3.0);
}

__device__ __forceinline__ float rsqrt2_refined(float r2)
{
    const float y = __builtin_amdgcn_rsqf(r2);
    const float h = r2 * y;
    return y * __builtin_fmaf(-h, y, 3.0f);
}

// exp(x) in FP64 without the library call (about 22 instructions instead of
// about 40): Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split
// into a high part of 32 significant bits (k * high exact for |k| < 2^21)
// and the low remainder; exp(r) by its Taylor series to degree 13 (the
// remainder r^14/14! < 4e-18 relative), evaluated by Horner's rule; then
// 2^k by ldexp.  x is clamped to [-746, 710] first, so k fits an int and
// the result underflows to 0 below about -745.1 and overflows to inf above
// about 709.8; NaN passes through.  Accurate to about 2 ulp.
__device__ __forceinline__ double exp_fast(double x)
{
    const double xc = __builtin_fmin(__builtin_fmax(x, -746.0), 710.0);
    const double k = __builtin_rint(xc * 1.4426950408889634);  // 1 / ln2
    double r = __builtin_fma(-k, 6.93147180369123816490e-01, xc);
    r = __builtin_fma(-k, 1.90821492927058770002e-10, r);
    double p = 1.0 / 6227020800.0;  // 1/13!
    p = __builtin_fma(p, r, 1.0 / 479001600.0);
    p = __builtin_fma(p, r, 1.0 / 39916800.0);
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    const double y = __builtin_ldexp(p, (int)k);
    return x == x ? y : x;
}

// exp(x) in FP32 as the hardware 2^t of t = x log2(e): relative error about
// |x| * 1e-7 from the rounding of t (below 1e-6 where the Gaussian and
// Yukawa kernels are not negligible), far inside the FP32 tolerances.
__device__ __forceinline__ float exp_fast(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }

// ln(x) / 2 for finite x > 0 in FP64 without the library call (about 35
// instead of about 100 instructions).  x = 2^k m with m in [sqrt(1/2),
// sqrt(2)); ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, by the
// series 2 (s + s^3/3 + ... + s^19/19) (the remainder s^21/21 is below
// 3e-17 relative); k ln2 is added with ln 2 split as in exp_fast (both parts
// halved, which is exact).  The quotient s comes from the hardware
// reciprocal refined by two Newton steps and one correction of the quotient.
// Within 3 ulp of libm's log(x) / 2 on 2e7 samples over [1e-300, 1e300].
__device__ __forceinline__ double half_log_fast(double x)
{
    int k = __builtin_amdgcn_frexp_exp(x);  // x = m 2^k, m in [1/2, 1)
    double m = __builtin_amdgcn_frexp_mant(x);
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        --k;
    }
    const double a = m - 1.0, b = m + 1.0;
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    double q = a * r;
    q = __builtin_fma(__builtin_fma(-b, q, a), r, q);
    const double z = q * q;
    double p = 1.0 / 19;
    p = __builtin_fma(p, z, 1.0 / 17);
    p = __builtin_fma(p, z, 1.0 / 15);
    p = __builtin_fma(p, z, 1.0 / 13);
    p = __builtin_fma(p, z, 1.0 / 11);
    p = __builtin_fma(p, z, 1.0 / 9);
    p = __builtin_fma(p, z, 1.0 / 7);
    p = __builtin_fma(p, z, 1.0 / 5);
    p = __builtin_fma(p, z, 1.0 / 3);
    p = __builtin_fma(p, z, 1.0);
    const double kd = (double)k;
    return __builtin_fma(kd, 6.93147180369123816490e-01 / 2,
                         __builtin_fma(kd, 1.90821492927058770002e-10 / 2, q * p));
}

__device__ __forceinline__ float half_log_fast(float x) { return 0.5f * std::log(x); }

template <typename T, int D>
__device__ __forceinline__ T dist2(const T *x, const T *y)
{
    T r2 = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const T t = x[d] - y[d];
        r2 = __builtin_fma(t, t, r2);
    }
    return r2;
}

struct KInvR {  // reference default: 1/|x-y|, inf at r = 0 (include/fmma/fmma.hpp:55-63)
    static constexpr bool translation_invariant = true;
    static constexpr int symmetry = 1;
    double p0, p1;
    template <typename T, int D>
    __device__ __forceinline__ T operator()(const T *x, const T *y) const
    {
        const T r2 = dist2<T, D>(x, y);
        return r2 == T(0) ? T(INFINITY) : rsqrt_refined(r2);
    }
    template <typename T, int D>
    __device__ __forceinline__ T far(const T *x, const T *y) const
    {
        return rsqrt_refined(dist2<T, D>(x, y));
    }
    // grad_x f = -(x - y) / r^3; taken as 0 at r = 0, where f is infinite
    template <typename T, int D>
    __device__ __forceinline__ void grad(const T *x, const T *y, T *g) const
    {
        const T r2 = dist2<T, D>(x, y);
        const T inv = rsqrt_refined(r2);
        const T c = r2 == T(0) ? T(0) : -(inv * inv * inv);
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = c * (x[d] - y[d]);
    }
    template <typename T, int D>
    __device__ __forceinline__ void grad_far(const T *x, const T *y, T *g) const
    {
        const T inv = rsqrt_refined(dist2<T, D>(x, y));
        const T c = -(inv * inv * inv);
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = c * (x[d] - y[d]);
    }
};

struct KInvRReg {  // 1/|x-y|, 0 at r = 0
    static constexpr bool translation_invariant = true;
    static constexpr int symmetry = 1;
    static constexpr double pair_scale = 2.0;  // the pair loops evaluate 2/r
    double p0, p1;
    template <typename T, int D>
    __device__ __forceinline__ T operator()(const T *x, const T *y) const
    {
        const T r2 = dist2<T, D>(x, y);
        const T v = rsqrt_refined(r2);
        return r2 == T(0) ? T(0) : v;
    }
    template <typename T, int D>
    __device__ __forceinline__ T near_scaled(const T *x, const T *y) const
    {
        const T r2 = dist2<T, D>(x, y);
        const T v = rsqrt2_refined(r2);
        return r2 == T(0) ? T(0) : v;
    }
    template <typename T, int D>
    __device__ __forceinline__ T far_scaled(const T *x, const T *y) const
    {
        return rsqrt2_refined(dist2<T, D>(x, y));
    }
    template <typename T, int D>
    __device__ __forceinline__ T far(const T *x, const T *y) const
    {
        return rsqrt_refined(dist2<T, D>(x, y));
    }
    template <typename T, int D>
    __device__ __forceinline__ void grad(const T *x, const T *y, T *g) const
    {
        const T r2 = dist2<T, D>(x, y);
        const T inv = rsqrt_refined(r2);
        const T c = r2 == T(0) ? T(0) : -(inv * inv * inv);
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = c * (x[d] - y[d]);
    }
    template <typename T, int D>
    __device__ __forceinline__ void grad_far(const T *x, const T *y, T *g) const
    {
        const T inv = rsqrt_refined(dist2<T, D>(x, y));
        const T c = -(inv * inv * inv);
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = c * (x[d] - y[d]);
    }
};

struct KInvDiff {  // 1/(x_0 - y_0)
    static constexpr bool translation_invariant = true;
    static constexpr int symmetry = -1;
    double p0, p1;
    template <typename T, int D>
    __device__ __forceinline__ T operator()(const T *x, const T *y) const
    {
        return T(1) / (x[0] - y[0]);
    }
    // d/dx_0 = -1 / (x_0 - y_0)^2, the other components 0
    template <typename T, int D>
    __device__ __forceinline__ void grad(const T *x, const T *y, T *g) const
    {
        const T inv = T(1) / (x[0] - y[0]);
        g[0] = -(inv * inv);
#pragma unroll
        for (int d = 1; d < D; ++d) g[d] = T(0);
    }
};

struct KLogR {  // log|x-y| = 0.5 log r^2, 0 at r = 0
    static constexpr bool translation_invariant = true;
    static constexpr int symmetry = 1;
    double p0, p1;
    template <typename T, int D>
    __device__ __forceinline__ T operator()(const T *x, const T *y) const
    {
        const T r2 = dist2<T, D>(x, y);
        return r2 == T(0) ? T(0) : half_log_fast(r2);
    }
    template <typename T, int D>
    __device__ __forceinline__ T far(const T *x, const T *y) const
    {
        return half_log_fast(dist2<T, D>(x, y));
    }
    // grad_x f = (x - y) / r^2, 0 at r = 0
    template <typename T, int D>
    __device__ __forceinline__ void grad(const T *x, const T *y, T *g) const
    {
        const T r2 = dist2<T, D>(x, y);
        const T c = r2 == T(0) ? T(0) : T(1) / r2;
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = c * (x[d] - y[d]);
    }
};

struct KGauss {  // exp(-r^2 / (2 h^2)); p1 = 1 / (2 h^2)
    static constexpr bool translation_invariant = true;
    static constexpr int symmetry = 1;
    double p0, p1;
    template <typename T, int D>
    __device__ __forceinline__ T operator()(const T *x, const T *y) const
    {
        const T r2 = dist2<T, D>(x, y);
        return exp_fast(-r2 * (T)p1);
    }
    // grad_x f = -2 p1 (x - y) f
    template <typename T, int D>
    __device__ __forceinline__ void grad(const T *x, const T *y, T *g) const
    {
        const T r2 = dist2<T, D>(x, y);
        const T c = T(-2) * (T)p1 * exp_fast(-r2 * (T)p1);
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = c * (x[d] - y[d]);
    }
};

struct KYukawa {  // exp(-kappa r) / r, 0 at r = 0; p0 = kappa
    static constexpr bool translation_invariant = true;
    static constexpr int symmetry = 1;
    double p0, p1;
    template <typename T, int D>
    __device__ __forceinline__ T operator()(const T *x, const T *y) const
    {
        const T r2 = dist2<T, D>(x, y);
        const T inv = rsqrt_refined(r2);
        const T v = exp_fast(-(T)p0 * (r2 * inv)) * inv;
        return r2 == T(0) ? T(0) : v;
    }
    template <typename T, int D>
    __device__ __forceinline__ T far(const T *x, const T *y) const
    {
        const T r2 = dist2<T, D>(x, y);
        const T inv = rsqrt_refined(r2);
        return exp_fast(-(T)p0 * (r2 * inv)) * inv;
    }
    // grad_x f = -(1 + kappa r) exp(-kappa r) (x - y) / r^3, 0 at r = 0
    template <typename T, int D>
    __device__ __forceinline__ void grad(const T *x, const T *y, T *g) const
    {
        const T r2 = dist2<T, D>(x, y);
        const T c = r2 == T(0) ? T(0) : coef(r2);
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = c * (x[d] - y[d]);
    }
    template <typename T, int D>
    __device__ __forceinline__ void grad_far(const T *x, const T *y, T *g) const
    {
        const T c = coef(dist2<T, D>(x, y));
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = c * (x[d] - y[d]);
    }
    template <typename T>
    __device__ __forceinline__ T coef(T r2) const
    {
        const T inv = rsqrt_refined(r2);
        const T kr = (T)p0 * (r2 * inv);
        return -(T(1) + kr) * exp_fast(-kr) * (inv * inv * inv);
    }
};

struct KSumR2 {  // sum_d (x_d - y_d) / r^2
    static constexpr bool translation_invariant = true;
    static constexpr int symmetry = -1;
    double p0, p1;
    template <typename T, int D>
    __device__ __forceinline__ T operator()(const T *x, const T *y) const
    {
        T r2 = T(0), s = T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T t = x[d] - y[d];
            r2 = __builtin_fma(t, t, r2);
            s += t;
        }
        return s / r2;
    }
    // d/dx_d = 1 / r^2 - 2 s (x_d - y_d) / r^4, s = sum_e (x_e - y_e)
    template <typename T, int D>
    __device__ __forceinline__ void grad(const T *x, const T *y, T *g) const
    {
        T r2 = T(0), s = T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T t = x[d] - y[d];
            r2 = __builtin_fma(t, t, r2);
            s += t;
        }
        const T i2 = T(1) / r2, c = T(-2) * s * i2 * i2;
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = i2 + c * (x[d] - y[d]);
    }
};

struct KSum2R2 {  // sum_d (x_d + y_d)^2 / r^2: depends on x and y, not only x - y
    static constexpr bool translation_invariant = false;
    static constexpr int symmetry = 1;
    double p0, p1;
    template <typename T, int D>
    __device__ __forceinline__ T operator()(const T *x, const T *y) const
    {
        T r2 = T(0), s2 = T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T t = x[d] - y[d], u = x[d] + y[d];
            r2 = __builtin_fma(t, t, r2);
            s2 = __builtin_fma(u, u, s2);
        }
        return s2 / r2;
    }
    // d/dx_d = 2 (x_d + y_d) / r^2 - 2 S (x_d - y_d) / r^4, S = sum_e (x_e + y_e)^2
    template <typename T, int D>
    __device__ __forceinline__ void grad(const T *x, const T *y, T *g) const
    {
        T r2 = T(0), s2 = T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T t = x[d] - y[d], u = x[d] + y[d];
            r2 = __builtin_fma(t, t, r2);
            s2 = __builtin_fma(u, u, s2);
        }
        const T i2 = T(1) / r2, c = T(-2) * s2 * i2 * i2;
#pragma unroll
        for (int d = 0; d < D; ++d) g[d] = T(2) * (x[d] + y[d]) * i2 + c * (x[d] - y[d]);
    }
};

}  // namespace lib
}  // namespace fmma
