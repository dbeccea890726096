// pair_kernels.hpp -- the FMM stages whose arithmetic depends on the kernel f.
//
// Templated on a device functor K with
//     static constexpr bool translation_invariant;
//     template <typename T, int D> __device__ T operator()(const T *x, const T *y) const;
// The library instantiates them for its built-in kernels; a caller compiled
// with hipcc instantiates them for its own functor (include/fmma/fmma_device.hpp)
// and hands the launchers to a plan through fmma_kernel_ops.
//
// Reference loops these replace:
//   P2P          src/fmm.cpp:122-141 (+ neighbour list src/fmm_core.cpp:81-115)
//   exact        src/fmma.cpp:79-94
//   M2L operator src/fmm_core.cpp:388-418 (translation-symmetric branch)
//   M2L general  src/fmm_core.cpp:456-492 (kernel evaluated per box pair)
#pragma once
#include <hip/hip_runtime.h>

#include "../../fmma_hip.h"
#include "morton.hpp"

namespace fmma {
namespace detail {

// ----------------------------------------------------------------- P2P
//
// One workgroup of 256 threads per target leaf box (Morton order).  Wave w
// takes targets [start + 64w, +64), then +256, ...; every lane keeps one
// target in registers and the wave streams the sources of the 3^D clipped
// neighbour boxes in the reference's order (row-major over the block, each
// box in ascending source index).  The source index is wave-uniform, so the
// coordinates and weights arrive through scalar loads, with no LDS and no
// barrier.  The result is written once, in the caller's target order.
template <typename T, int D, class K>
__global__ __launch_bounds__(256) void p2p_kernel(fmma_p2p_args a, K kern)
{
    const unsigned box = blockIdx.x;
    const unsigned t0 = a.t_start[box], t1 = a.t_start[box + 1];
    const unsigned wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    unsigned tb = t0 + wave * 64u;
    if (tb >= t1) return;

    int bc[D], lo[D], hi[D];
    morton_decode<D>(box, bc);
    const int NL = 1 << a.leaf_level;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lo[d] = bc[d] > 0 ? bc[d] - 1 : bc[d];
        hi[d] = bc[d] + 1 < NL ? bc[d] + 1 : bc[d];
    }
    const T *tp = static_cast<const T *>(a.t_pos);
    const T *sp = static_cast<const T *>(a.s_pos);
    const T *sw = static_cast<const T *>(a.s_w);
    const long nt = a.n_t, ns = a.n_s;

    for (; tb < t1; tb += 256u) {
        const unsigned ti = tb + lane;
        const bool valid = ti < t1;
        const unsigned tc = valid ? ti : t1 - 1;
        T x[D];
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = tp[d * nt + tc];
        T acc = T(0);
        int nb[D];
#pragma unroll
        for (int d = 0; d < D; ++d) nb[d] = lo[d];
        while (true) {
            const unsigned m = morton_encode<D>(nb);
            const unsigned s0 = a.s_start[m], s1 = a.s_start[m + 1];
#pragma unroll 4
            for (unsigned j = s0; j < s1; ++j) {
                T y[D];
#pragma unroll
                for (int d = 0; d < D; ++d) y[d] = sp[d * ns + j];
                acc += sw[j] * kern.template operator()<T, D>(x, y);
            }
            // row-major odometer over the neighbour block, last dimension fastest
            int d = D - 1;
            while (d >= 0 && nb[d] == hi[d]) { nb[d] = lo[d]; --d; }
            if (d < 0) break;
            ++nb[d];
        }
        if (valid) {
            T *ans = static_cast<T *>(a.ans);
            const unsigned oi = a.t_perm[ti];
            if (a.base)
                ans[oi] = static_cast<const T *>(a.base)[ti] + acc;
            else if (a.accumulate)
                ans[oi] += acc;
            else
                ans[oi] = acc;
        }
    }
}

template <typename T, int D, class K>
int launch_p2p(const fmma_p2p_args *a, const K &k, hipStream_t s)
{
    if (a->n_target_boxes <= 0) return FMMA_OK;
    hipLaunchKernelGGL((p2p_kernel<T, D, K>), dim3((unsigned)a->n_target_boxes), dim3(256), 0, s, *a, k);
    return hipGetLastError() == hipSuccess ? FMMA_OK : FMMA_ERR_HIP;
}

// --------------------------------------------------------------- exact
//
// All pairs.  One thread per target; sources are wave-uniform (scalar loads).
// With split == 1 each target sums its sources in ascending order, the
// reference's order (src/fmma.cpp:88-92).  With split > 1 (few targets, many
// sources) block column s sums the s-th contiguous slice of sources and
// exact_reduce adds the slices in increasing s.
template <typename T, int D, class K>
__global__ __launch_bounds__(256) void exact_kernel(fmma_exact_args a, K kern)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < a.n_t;
    const long ic = valid ? i : a.n_t - 1;
    const T *tg = static_cast<const T *>(a.targets);
    const T *sr = static_cast<const T *>(a.sources);
    const T *w = static_cast<const T *>(a.weights);
    T x[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = tg[ic * D + d];
    const long s = blockIdx.y;
    const long j0 = a.n_s * s / a.split, j1 = a.n_s * (s + 1) / a.split;
    T acc = T(0);
    if (w) {
#pragma unroll 4
        for (long j = j0; j < j1; ++j) {
            T y[D];
#pragma unroll
            for (int d = 0; d < D; ++d) y[d] = sr[j * D + d];
            acc += w[j] * kern.template operator()<T, D>(x, y);
        }
    } else {
#pragma unroll 4
        for (long j = j0; j < j1; ++j) {
            T y[D];
#pragma unroll
            for (int d = 0; d < D; ++d) y[d] = sr[j * D + d];
            acc += kern.template operator()<T, D>(x, y);
        }
    }
    if (!valid) return;
    if (a.split == 1)
        static_cast<T *>(a.ans)[i] = acc;
    else
        static_cast<T *>(a.partial)[s * a.n_t + i] = acc;
}

template <typename T>
__global__ __launch_bounds__(256) void exact_reduce(fmma_exact_args a)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n_t) return;
    const T *p = static_cast<const T *>(a.partial);
    T acc = p[i];
    for (int s = 1; s < a.split; ++s) acc += p[(long)s * a.n_t + i];
    static_cast<T *>(a.ans)[i] = acc;
}

template <typename T, int D, class K>
int launch_exact(const fmma_exact_args *a, const K &k, hipStream_t s)
{
    if (a->n_t <= 0) return FMMA_OK;
    const unsigned nb = (unsigned)((a->n_t + 255) / 256);
    hipLaunchKernelGGL((exact_kernel<T, D, K>), dim3(nb, (unsigned)a->split), dim3(256), 0, s, *a, k);
    if (a->split > 1) hipLaunchKernelGGL((exact_reduce<T>), dim3(nb), dim3(256), 0, s, *a);
    return hipGetLastError() == hipSuccess ? FMMA_OK : FMMA_ERR_HIP;
}

// --------------------------------------------------------- M2L operators
//
// K[o][kl][km] = f(X_kl, Y_km), X = len*(c_kl/2), Y = len*(o + c_km/2) in
// TYPE (src/fmm_core.cpp:405-414).  Multi-index digits: dimension 0 most
// significant (src/fmm_core.cpp:246-253).
template <typename T, int D, class K>
__global__ __launch_bounds__(256) void m2l_op_kernel(fmma_m2l_op_args a, K kern)
{
    const int oi = blockIdx.y;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const long PP = (long)a.P * a.P;
    if (e >= PP) return;
    const int kl = (int)(e / a.P), km = (int)(e % a.P);
    const T len = *static_cast<const T *>(a.Len) / (T)(1u << a.level);
    T X[D], Y[D];
    int dl = kl, dm = km;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
        const T cl = (T)a.nodes[dl % a.n_nodes], cm = (T)a.nodes[dm % a.n_nodes];
        dl /= a.n_nodes;
        dm /= a.n_nodes;
        X[d] = len * (cl / 2);
        Y[d] = len * ((T)a.offsets[oi * D + d] + cm / 2);
    }
    static_cast<T *>(a.K)[oi * PP + e] = kern.template operator()<T, D>(X, Y);
}

template <typename T, int D, class K>
int launch_m2l_op(const fmma_m2l_op_args *a, const K &k, hipStream_t s)
{
    const long PP = (long)a->P * a->P;
    hipLaunchKernelGGL((m2l_op_kernel<T, D, K>), dim3((unsigned)((PP + 255) / 256), (unsigned)a->n_off), dim3(256),
                       0, s, *a, k);
    return hipGetLastError() == hipSuccess ? FMMA_OK : FMMA_ERR_HIP;
}

// ----------------------------------------------------------- M2L general
//
// For kernels that are not functions of x-y: the operator of every box pair
// is evaluated on the fly at absolute node positions
// X = len*(b + (1+c)/2) + origin (src/fmm_core.cpp:471,480), as the reference
// does when trans_sym_flag is 0.  One workgroup per target box, one thread
// per row kl, the interaction list walked in the reference's row-major order.
template <typename T, int D, class K>
__global__ __launch_bounds__(64) void m2l_general_kernel(fmma_m2l_general_args a, K kern)
{
    const unsigned box = blockIdx.x;
    const int N = 1 << a.level, n = a.n_nodes, P = a.P;
    const double len = (double)(*static_cast<const T *>(a.Len) / (T)(1u << a.level));
    int bc[D], lo[D], hi[D];
    morton_decode<D>(box, bc);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int p = bc[d] / 2;
        lo[d] = (p > 0 ? p - 1 : p) * 2;
        hi[d] = (p + 1 < N / 2 ? p + 1 : p) * 2 + 1;
    }
    const T *org = static_cast<const T *>(a.origin);
    const T *Wm = static_cast<const T *>(a.Wm);
    T *Wl = static_cast<T *>(a.Wl);
    for (int kl = threadIdx.x; kl < P; kl += 64) {
        T X[D];
        int dl = kl;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            X[d] = (T)(len * ((double)(T)bc[d] + (1.0 + (double)(T)a.nodes[dl % n]) / 2) + (double)org[d]);
            dl /= n;
        }
        T acc = T(0);
        int m[D];
#pragma unroll
        for (int d = 0; d < D; ++d) m[d] = lo[d];
        while (true) {
            int far = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int dd = m[d] > bc[d] ? m[d] - bc[d] : bc[d] - m[d];
                far = dd > far ? dd : far;
            }
            if (far > 1) {
                const T *w = Wm + (long)morton_encode<D>(m) * P;
                for (int km = 0; km < P; ++km) {
                    T Y[D];
                    int dm = km;
#pragma unroll
                    for (int d = D - 1; d >= 0; --d) {
                        Y[d] = (T)(len * ((double)(T)m[d] + (1.0 + (double)(T)a.nodes[dm % n]) / 2) +
                                   (double)org[d]);
                        dm /= n;
                    }
                    acc += kern.template operator()<T, D>(X, Y) * w[km];
                }
            }
            int d = D - 1;
            while (d >= 0 && m[d] == hi[d]) { m[d] = lo[d]; --d; }
            if (d < 0) break;
            ++m[d];
        }
        Wl[(long)box * P + kl] = acc;
    }
}

template <typename T, int D, class K>
int launch_m2l_general(const fmma_m2l_general_args *a, const K &k, hipStream_t s)
{
    const unsigned nbox = 1u << (a->level * D);
    hipLaunchKernelGGL((m2l_general_kernel<T, D, K>), dim3(nbox), dim3(64), 0, s, *a, k);
    return hipGetLastError() == hipSuccess ? FMMA_OK : FMMA_ERR_HIP;
}

// ------------------------------------------------- launcher table builder
//
// Fills an fmma_kernel_ops for functor K whose state lives at `ctx`
// (a K object that must outlive the plan's use of the table).
template <typename T, int D, class K>
struct ops_for {
    static int p2p(const fmma_p2p_args *a, void *ctx, void *s)
    {
        return launch_p2p<T, D, K>(a, *static_cast<const K *>(ctx), (hipStream_t)s);
    }
    static int exact(const fmma_exact_args *a, void *ctx, void *s)
    {
        return launch_exact<T, D, K>(a, *static_cast<const K *>(ctx), (hipStream_t)s);
    }
    static int m2l_op(const fmma_m2l_op_args *a, void *ctx, void *s)
    {
        return launch_m2l_op<T, D, K>(a, *static_cast<const K *>(ctx), (hipStream_t)s);
    }
    static int m2l_general(const fmma_m2l_general_args *a, void *ctx, void *s)
    {
        return launch_m2l_general<T, D, K>(a, *static_cast<const K *>(ctx), (hipStream_t)s);
    }
    static fmma_kernel_ops make(const K *ctx)
    {
        fmma_kernel_ops o;
        o.dim = D;
        o.dtype = sizeof(T) == 8 ? FMMA_F64 : FMMA_F32;
        o.translation_invariant = K::translation_invariant ? 1 : 0;
        o.ctx = const_cast<K *>(ctx);
        o.p2p = &p2p;
        o.exact = &exact;
        o.m2l_op = &m2l_op;
        o.m2l_general = &m2l_general;
        return o;
    }
};

}  // namespace detail
}  // namespace fmma
