// common.hpp -- internal definitions of libfmma_hip (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/fmma_hip.h"
#include "../../include/fmma/detail/morton.hpp"

namespace fmma {
namespace lib {

using detail::morton_decode;
using detail::morton_encode;

constexpr int kMaxNodes = 16;     // nodes per dimension (poly_ord <= 15)
constexpr int kMaxP = 512;        // coefficients per box handled by the box kernels

void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define FMMA_HIP(call)                                   \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return ::fmma::lib::hip_fail(e_, #call); \
    } while (0)

#define FMMA_LAUNCHED(what)                                   \
    do {                                                      \
        hipError_t e_ = hipGetLastError();                    \
        if (e_ != hipSuccess) return ::fmma::lib::hip_fail(e_, what);\
    } while (0)

// Bounding cube and leaf grid, computed on the device in TYPE arithmetic
// (src/fmm_core.cpp:160-173).
template <typename T>
struct Geom {
    T origin[3];
    T Len;
    T len_leaf;  // Len / (TYPE)N_L, N_L = 1 << leaf_level (src/fmm.cpp:58,129)
    int status;  // kGeomOk, or why the cube cannot be binned
};

// Geom::status.  The reference divides by Len without a check
// (src/fmm_core.cpp:261): with every point at one position, or a non-finite
// coordinate, its box indices are undefined.  Here the solve is rejected.
constexpr int kGeomOk = 0;
constexpr int kGeomNonFinite = 1;  // a coordinate is NaN or infinite
constexpr int kGeomZeroLen = 2;    // Len == 0: every point at one position

// Leaf coordinate along one dimension: min((int)((x-o)/len), N-1) with IEEE
// subtraction and division (src/fmm_core.cpp:261, 595; src/fmm.cpp:130).
template <typename T>
__device__ __forceinline__ int leaf_coord(T x, T o, T len, int N)
{
#pragma clang fp contract(off)
    const T q = (x - o) / len;
    int b = (int)q;
    b = b < N - 1 ? b : N - 1;
    return b > 0 ? b : 0;  // q >= -1 for any finite cube; keeps garbage geometry in range
}

// The same, plus the clamped relative coordinate 2((x-o)/len - b) - 1
// (src/fmm_core.cpp:264, 596).  2*t is exact, so evaluating in TYPE gives the
// same value as the reference's double-literal expression.
template <typename T>
__device__ __forceinline__ int leaf_coord_rel(T x, T o, T len, int N, T &rel)
{
#pragma clang fp contract(off)
    const T q = (x - o) / len;
    int b = (int)q;
    b = b < N - 1 ? b : N - 1;
    b = b > 0 ? b : 0;
    T r = T(2) * (q - (T)b) - T(1);
    r = r < T(1) ? r : T(1);
    r = r > T(-1) ? r : T(-1);
    rel = r;
    return b;
}

// Chebyshev interpolation weights S_n(c_k, r) for k = 0..n-1
// (src/math.cpp:31-50): coef[k][j] = (2/n) T_j(c_k) for j >= 1, 1/n for
// j = 0, so S_n(c_k, r) = sum_j coef[k][j] T_j(r), with T_j(r) from the
// three-term recurrence instead of cos(j acos r).  `out` (stride `ostride`)
// is usually LDS; `scale` multiplies every weight (a source weight in P2M).
template <typename T>
__device__ __forceinline__ void cheb_weights(const T *coef, int n, T r, T scale, T *out, int ostride)
{
    for (int k = 0; k < n; ++k) out[k * ostride] = coef[k * n];
    T tp = T(1), t = r;
    for (int j = 1; j < n; ++j) {
        for (int k = 0; k < n; ++k) out[k * ostride] += coef[k * n + j] * t;
        const T tn = T(2) * r * t - tp;
        tp = t;
        t = tn;
    }
    for (int k = 0; k < n; ++k) out[k * ostride] *= scale;
}

}  // namespace lib
}  // namespace fmma
