// batch.hip -- tree build for a batch of independent problems.
//
// A batch is a forest: problem p keeps its own bounding cube
// (src/fmm_core.cpp:127-173) and leaf grid, and its leaf boxes occupy a
// contiguous range [box_base[p], box_base[p+1]) of one global box index
// space, Morton order inside.  One stable radix sort over the global keys
// then sorts every problem at once without mixing them, since the key
// ranges of the problems follow their order.
#include "batch.hpp"

namespace fmma {
namespace lib {

// One workgroup per problem: exact min/max over its points, then the
// single-problem formulas (tree.hip bbox_final).
template <typename T, int D>
__global__ __launch_bounds__(256) void batch_bbox_kernel(const T *t, const long *t_off, const T *s, const long *s_off,
                                                         const int *levels, Geom<T> *geoms)
{
    const int p = blockIdx.x;
    bool bad = false;
    T lo[D], hi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lo[d] = T(INFINITY);
        hi[d] = T(-INFINITY);
    }
    for (int set = 0; set < 2; ++set) {
        const T *x = set == 0 ? t : s;
        if (!x) continue;
        const long *off = set == 0 ? t_off : s_off;
        for (long i = off[p] + threadIdx.x; i < off[p + 1]; i += 256)
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const T v = x[i * D + d];
                lo[d] = v < lo[d] ? v : lo[d];
                hi[d] = hi[d] < v ? v : hi[d];
                bad = bad || !__builtin_isfinite(v);
            }
    }
    __shared__ int s_bad;
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    if (bad) s_bad = 1;
#pragma unroll
    for (int d = 0; d < D; ++d)
        for (int o = 32; o > 0; o >>= 1) {
            const T l2 = __shfl_xor(lo[d], o), h2 = __shfl_xor(hi[d], o);
            lo[d] = l2 < lo[d] ? l2 : lo[d];
            hi[d] = hi[d] < h2 ? h2 : hi[d];
        }
    __shared__ T sm[4][2 * D];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            sm[wave][d] = lo[d];
            sm[wave][D + d] = hi[d];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma clang fp contract(off)
        for (int w = 1; w < 4; ++w)
#pragma unroll
            for (int d = 0; d < D; ++d) {
                sm[0][d] = sm[w][d] < sm[0][d] ? sm[w][d] : sm[0][d];
                sm[0][D + d] = sm[0][D + d] < sm[w][D + d] ? sm[w][D + d] : sm[0][D + d];
            }
        T Len = T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T e = sm[0][D + d] - sm[0][d];
            Len = Len < e ? e : Len;
        }
        Geom<T> *g = geoms + p;
#pragma unroll
        for (int d = 0; d < D; ++d) g->origin[d] = (sm[0][D + d] + sm[0][d]) / 2 - Len / 2;
        for (int d = D; d < 3; ++d) g->origin[d] = T(0);
        g->Len = Len;
        g->len_leaf = Len / (T)(1u << levels[p]);
        g->status = s_bad ? kGeomNonFinite : Len == T(0) ? kGeomZeroLen : kGeomOk;
    }
}

template <typename T, int D>
int batch_bbox(const T *t, const long *t_off, const T *s, const long *s_off, const int *levels, int nprob,
               Geom<T> *geoms, hipStream_t st)
{
    if (nprob <= 0) return FMMA_OK;
    hipLaunchKernelGGL((batch_bbox_kernel<T, D>), dim3((unsigned)nprob), dim3(256), 0, st, t, t_off, s, s_off, levels,
                       geoms);
    FMMA_LAUNCHED("batch_bbox");
    return FMMA_OK;
}

template <typename T, int D>
__global__ __launch_bounds__(256) void batch_keys_kernel(const T *pts, long n, const long *off, int nprob,
                                                         const Geom<T> *geoms, const int *levels,
                                                         const unsigned *box_base, unsigned *keys)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nprob;  // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    const Geom<T> *g = geoms + lo;
    const int NL = 1 << levels[lo];
    int b[D];
#pragma unroll
    for (int d = 0; d < D; ++d) b[d] = leaf_coord(pts[i * D + d], g->origin[d], g->len_leaf, NL);
    keys[i] = box_base[lo] + morton_encode<D>(b);
}

template <typename T, int D>
int batch_keys(const T *pts, long n, const long *off, int nprob, const Geom<T> *geoms, const int *levels,
               const unsigned *box_base, unsigned *keys, hipStream_t st)
{
    if (n <= 0) return FMMA_OK;
    hipLaunchKernelGGL((batch_keys_kernel<T, D>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, pts, n, off,
                       nprob, geoms, levels, box_base, keys);
    FMMA_LAUNCHED("batch_keys");
    return FMMA_OK;
}

#define FMMA_BATCH_INST(T, D)                                                                                   \
    template int batch_bbox<T, D>(const T *, const long *, const T *, const long *, const int *, int, Geom<T> *, \
                                  hipStream_t);                                                                 \
    template int batch_keys<T, D>(const T *, long, const long *, int, const Geom<T> *, const int *,            \
                                  const unsigned *, unsigned *, hipStream_t);
FMMA_BATCH_INST(double, 1)
FMMA_BATCH_INST(double, 2)
FMMA_BATCH_INST(double, 3)
FMMA_BATCH_INST(float, 1)
FMMA_BATCH_INST(float, 2)
FMMA_BATCH_INST(float, 3)

}  // namespace lib
}  // namespace fmma
