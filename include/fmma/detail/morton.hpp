// morton.hpp -- box indexing shared by the library and user-functor launchers.
//
// Boxes are stored in Morton (Z-order) at every level: bit i of coordinate d
// goes to bit  DIM*i + (DIM-1-d)  of the index, so that the low DIM bits of a
// leaf index are the child position inside its parent with dimension 0 as the
// most significant bit -- the reference's child selector `sind`
// (src/fmm_core.cpp:343-349) -- and the parent of box k is k >> DIM.  The
// reference's own box id is row-major (src/fmm_core.cpp:205-212); the two
// orders hold the same boxes, only the storage order differs.
#pragma once
#include <hip/hip_runtime.h>

namespace fmma {
namespace detail {

__host__ __device__ __forceinline__ unsigned spread3(unsigned v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__host__ __device__ __forceinline__ unsigned compact3(unsigned v)
{
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030c30c3u;
    v = (v | (v >> 4)) & 0x0300f00fu;
    v = (v | (v >> 8)) & 0x030000ffu;
    v = (v | (v >> 16)) & 0x000003ffu;
    return v;
}

__host__ __device__ __forceinline__ unsigned spread2(unsigned v)
{
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__host__ __device__ __forceinline__ unsigned compact2(unsigned v)
{
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0f0f0f0fu;
    v = (v | (v >> 4)) & 0x00ff00ffu;
    v = (v | (v >> 8)) & 0x0000ffffu;
    return v;
}

template <int D>
__host__ __device__ __forceinline__ unsigned morton_encode(const int *b)
{
    if constexpr (D == 1) {
        return (unsigned)b[0];
    } else if constexpr (D == 2) {
        return (spread2((unsigned)b[0]) << 1) | spread2((unsigned)b[1]);
    } else {
        return (spread3((unsigned)b[0]) << 2) | (spread3((unsigned)b[1]) << 1) | spread3((unsigned)b[2]);
    }
}

template <int D>
__host__ __device__ __forceinline__ void morton_decode(unsigned k, int *b)
{
    if constexpr (D == 1) {
        b[0] = (int)k;
    } else if constexpr (D == 2) {
        b[0] = (int)compact2(k >> 1);
        b[1] = (int)compact2(k);
    } else {
        b[0] = (int)compact3(k >> 2);
        b[1] = (int)compact3(k >> 1);
        b[2] = (int)compact3(k);
    }
}

}  // namespace detail
}  // namespace fmma
