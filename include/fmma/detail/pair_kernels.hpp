// pair_kernels.hpp -- the FMM stages whose arithmetic depends on the kernel f.
//
// Templated on a device functor K with
//     static constexpr bool translation_invariant;
//     template <typename T, int D> __device__ T operator()(const T *x, const T *y) const;
// and optionally far(), grad(), grad_far() and `static constexpr int symmetry`.
// The library instantiates them for its built-in kernels; a caller compiled
// with hipcc instantiates them for its own functor (include/fmma/fmma_device.hpp)
// and hands the launchers to a plan through fmma_kernel_ops.
//
// Reference loops these replace:
//   P2P          src/fmm.cpp:122-141 (+ neighbour list src/fmm_core.cpp:81-115)
//   exact        src/fmma.cpp:79-94
//   M2L operator src/fmm_core.cpp:388-418 (translation-symmetric branch)
//   M2L general  src/fmm_core.cpp:456-492 (kernel evaluated per box pair)
//   M2P          src/fmm_core.cpp:625-661, src/nrnmm.cpp:79-91 (tree, nrnmm)
// and, new, the near field and all-pairs forms of the gradient with respect
// to the target coordinates.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "../../fmma_hip.h"
#include "m2l_layout.hpp"
#include "morton.hpp"

namespace fmma {
namespace detail {

// A functor may provide far(x, y): f for points known to be distinct (in
// different leaf boxes), without its r = 0 special case.
template <class K, typename T, int D, typename = void>
struct has_far : std::false_type {};
template <class K, typename T, int D>
struct has_far<K, T, D,
               std::void_t<decltype(std::declval<const K &>().template far<T, D>((const T *)nullptr,
                                                                                  (const T *)nullptr))>>
    : std::true_type {};

// A functor may declare `static constexpr int symmetry`: +1 when f(y,x) =
// f(x,y), -1 when f(y,x) = -f(x,y).  Without it P2P does not assume either.
template <class K, typename = void>
struct kernel_symmetry {
    static constexpr int value = 0;
};
template <class K>
struct kernel_symmetry<K, std::void_t<decltype(K::symmetry)>> {
    static constexpr int value = K::symmetry;
};

// A functor may evaluate c*f instead of f in the pair loops (P2P, M2P) when
// that is cheaper, declaring `static constexpr double pair_scale = c` (a power
// of two) and near_scaled() / far_scaled(); the loops then carry every weight
// divided by c, which is exact, so sum w f = sum (w / c)(c f).  1/r uses it to
// fold the Newton step's factor 1/2 into the weights.
template <class K, typename = void>
struct pair_scale {
    static constexpr double value = 1.0;
};
template <class K>
struct pair_scale<K, std::void_t<decltype(K::pair_scale)>> {
    static constexpr double value = K::pair_scale;
};

// The weight as the pair loops carry it.
template <class K, typename T>
__device__ __forceinline__ T pair_weight(T w)
{
    if constexpr (pair_scale<K>::value != 1.0)
        return w * T(1.0 / pair_scale<K>::value);
    else
        return w;
}

template <bool FAR, typename T, int D, class K>
__device__ __forceinline__ T pair_eval(const K &k, const T *x, const T *y)
{
    if constexpr (pair_scale<K>::value != 1.0) {
        if constexpr (FAR)
            return k.template far_scaled<T, D>(x, y);
        else
            return k.template near_scaled<T, D>(x, y);
    } else if constexpr (FAR && has_far<K, T, D>::value) {
        return k.template far<T, D>(x, y);
    } else {
        return k.template operator()<T, D>(x, y);
    }
}

// A functor may provide grad(x, y, g): g[d] = d/dx_d f(x, y) (the near
// field of the gradient), and grad_far() without its r = 0 case.
template <class K, typename T, int D, typename = void>
struct has_grad : std::false_type {};
template <class K, typename T, int D>
struct has_grad<K, T, D,
                std::void_t<decltype(std::declval<const K &>().template grad<T, D>(
                    (const T *)nullptr, (const T *)nullptr, (T *)nullptr))>> : std::true_type {};
template <class K, typename T, int D, typename = void>
struct has_grad_far : std::false_type {};
template <class K, typename T, int D>
struct has_grad_far<K, T, D,
                    std::void_t<decltype(std::declval<const K &>().template grad_far<T, D>(
                        (const T *)nullptr, (const T *)nullptr, (T *)nullptr))>> : std::true_type {};

template <bool FAR, typename T, int D, class K>
__device__ __forceinline__ void pair_grad(const K &k, const T *x, const T *y, T *g)
{
    if constexpr (FAR && has_grad_far<K, T, D>::value)
        k.template grad_far<T, D>(x, y, g);
    else
        k.template grad<T, D>(x, y, g);
}

// f^T(x, y) = f(y, x): the kernel of the transposed solve.
template <class K>
struct Transposed {
    K k;
    static constexpr bool translation_invariant = K::translation_invariant;
    template <typename T, int D>
    __device__ __forceinline__ T operator()(const T *x, const T *y) const
    {
        return k.template operator()<T, D>(y, x);
    }
};

// The leaf grid of a box: the plan's, or in a batch of independent problems
// that of the box's problem, whose boxes start at gbase (Morton order).
template <class A>
__device__ __forceinline__ int leaf_grid(const A &a, unsigned box, unsigned &gbase)
{
    if (a.box_problem) {
        const unsigned p = a.box_problem[box];
        gbase = a.problem_box_base[p];
        return a.problem_leaf_level[p];
    }
    gbase = 0;
    return a.leaf_level;
}

// Grid coordinates of a leaf box, and the boxes per side: a Morton tree (or
// a batch's problem tree), or a row-major grid of grid_n per side (nrnmm).
template <int D, class A>
__device__ __forceinline__ int box_coords_of(const A &a, unsigned box, int (&bc)[D], unsigned &gbase)
{
    if (a.grid_n > 0) {
        gbase = 0;
        unsigned r = box;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            bc[d] = (int)(r % (unsigned)a.grid_n);
            r /= (unsigned)a.grid_n;
        }
        return a.grid_n;
    }
    const int L = leaf_grid(a, box, gbase);
    morton_decode<D>(box - gbase, bc);
    return 1 << L;
}

template <int D, class A>
__device__ __forceinline__ unsigned box_index_of(const A &a, const int (&m)[D], unsigned gbase)
{
    if (a.grid_n > 0) {
        unsigned id = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) id = id * (unsigned)a.grid_n + (unsigned)m[d];
        return id;
    }
    return gbase + morton_encode<D>(m);
}

// ----------------------------------------------------------------- P2P
//
// One wave (64 threads) per target leaf box (Morton order).  Lane l is
// (slot s = l % S, slice q = l / S) with S = 4..64 slots chosen per box
// (the fewest that keep at most kP2PTargets targets per lane): slot s owns
// the box's targets s, s+S, s+2S, ... in registers and slice q takes every
// (64/S)-th source of the neighbourhood.  The sources of the 3^D clipped
// neighbour boxes (src/fmm_core.cpp:81-115), the box itself first, are
// staged through LDS in blocks of kP2PBlock (x.., w) records; every LDS read
// feeds one pair evaluation per target of the lane.  The 64/S slice partial
// sums of a target are added by xor shuffles in a fixed order, so results are
// deterministic.  Records of the target's own box may coincide with a target
// (f's r = 0 case); the other boxes use the functor's far() form if it has one.
constexpr int kP2PTargets = 12;  // targets per lane per pass
constexpr int kP2PBlock = 256;   // sources per LDS stage

template <int D>
struct p2p_rec {  // LDS record width in elements: coordinates + weight, padded to even
    static constexpr int W = (D + 2) & ~1;
};

// The pair loop over staged records [j0, j1) of slice q (indices = q mod
// nslice) for a compile-time number of targets.
template <int TPL, bool FAR, typename T, int D, int NT, class K>
__device__ __forceinline__ void p2p_range(const T *src, unsigned j0, unsigned j1, unsigned q, unsigned nslice,
                                          const T (&x)[NT][D], T (&acc)[NT], const K &kern)
{
    constexpr int RW = p2p_rec<D>::W;
    const unsigned r0 = j0 % nslice;
    for (unsigned j = j0 + (q >= r0 ? q - r0 : q + nslice - r0); j < j1; j += nslice) {
        const T *r = src + j * RW;
        T y[D];
#pragma unroll
        for (int d = 0; d < D; ++d) y[d] = r[d];
        const T wj = r[D];
#pragma unroll
        for (int k = 0; k < TPL; ++k) acc[k] += wj * pair_eval<FAR, T, D>(kern, x[k], y);
    }
}

template <int TPL, typename T, int D, class K>
__device__ __forceinline__ void p2p_block(const T *src, unsigned cnt, unsigned nself, unsigned q, unsigned nslice,
                                          const T (&x)[kP2PTargets][D], T (&acc)[kP2PTargets], const K &kern)
{
    const unsigned m = nself < cnt ? nself : cnt;
    if (m > 0) p2p_range<TPL, false>(src, 0, m, q, nslice, x, acc, kern);
    p2p_range<TPL, true>(src, m, cnt, q, nslice, x, acc, kern);
}

// The work of one workgroup of the direct P2P kernels: a whole target box,
// or one item (box, pass, chunk) of a box the library split (balanced
// form); false when there is nothing to do.
template <class A>
__device__ __forceinline__ bool p2p_work_item(const A &a, unsigned &box, unsigned &t0, unsigned &t1, unsigned &chunk,
                                              unsigned &nchunk)
{
    chunk = 0;
    nchunk = 1;
    if (a.split_items) {
        const unsigned *it = a.split_items + 3 * (size_t)blockIdx.x;
        box = it[0];
        chunk = it[2];
        nchunk = a.split_parts[box];
        const unsigned tb0 = a.t_start[box], tb1 = a.t_start[box + 1];
        t0 = tb0 + it[1] * a.split_pass;
        t1 = t0 + a.split_pass < tb1 ? t0 + a.split_pass : tb1;
        return t0 < t1;
    }
    box = blockIdx.x;
    if (a.split_parts && a.split_parts[box]) return false;  // the item launch covers it
    t0 = a.t_start[box];
    t1 = a.t_start[box + 1];
    return t0 < t1;
}

// Records [r_lo, r_hi) of chunk c of nchunk equal chunks of total records.
__device__ __forceinline__ void chunk_range(unsigned total, unsigned c, unsigned nchunk, unsigned &r_lo,
                                            unsigned &r_hi)
{
    if (nchunk <= 1) {
        r_lo = 0;
        r_hi = total;
        return;
    }
    const unsigned cb = (total + nchunk - 1) / nchunk;
    r_lo = c * cb < total ? c * cb : total;
    r_hi = r_lo + cb < total ? r_lo + cb : total;
}

template <typename T, int D, class K>
__global__ __launch_bounds__(64) void p2p_kernel(fmma_p2p_args a, K kern)
{
    constexpr int NNB = D == 1 ? 3 : D == 2 ? 9 : 27;
    constexpr int RW = p2p_rec<D>::W;
    __shared__ __align__(16) T src[kP2PBlock * RW];
    __shared__ unsigned nb_first[NNB], nb_pre[NNB + 1];

    unsigned box, t0, t1, chunk, nchunk;
    if (!p2p_work_item(a, box, t0, t1, chunk, nchunk)) return;
    const unsigned lane = threadIdx.x;
    const T *tp = static_cast<const T *>(a.t_pos);
    const T *sp = static_cast<const T *>(a.s_pos);
    const T *sw = static_cast<const T *>(a.s_w);
    const long nt = a.n_t, ns = a.n_s;

    // neighbour ranges: the box itself first, then the other entries of the
    // row-major 3^D block (last dimension fastest); prefix sum by a wave scan
    unsigned cnt_nb = 0;
    if (lane < (unsigned)NNB) {
        int bc[D], m[D];
        unsigned gbase;
        const int NL = box_coords_of<D>(a, box, bc, gbase);
        bool ok = true;
        constexpr int centre = (NNB - 1) / 2;
        int k = lane == 0 ? centre : (int)lane <= centre ? (int)lane - 1 : (int)lane;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            m[d] = bc[d] + (k % 3) - 1;
            k /= 3;
            ok = ok && m[d] >= 0 && m[d] < NL;
        }
        unsigned first = 0;
        if (ok) {
            const unsigned mm = box_index_of<D>(a, m, gbase);
            first = a.s_start[mm];
            cnt_nb = a.s_start[mm + 1] - first;
        }
        nb_first[lane] = first;
    }
    unsigned incl = cnt_nb;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const unsigned y = __shfl_up(incl, o);
        if ((int)lane >= o) incl += y;
    }
    if (lane < (unsigned)NNB) nb_pre[lane + 1] = incl;
    if (lane == 0) nb_pre[0] = 0;
    __syncthreads();
    const unsigned total = nb_pre[NNB], nself = nb_pre[1];
    unsigned r_lo, r_hi;
    chunk_range(total, chunk, nchunk, r_lo, r_hi);

    // slots: the fewest of 4, 8, .., 64 with <= kP2PTargets targets per lane
    const unsigned ntb = t1 - t0;
    unsigned nslot = 4;
    while (nslot < 64 && nslot * kP2PTargets < ntb) nslot <<= 1;
    const unsigned nslice = 64 / nslot, s = lane % nslot, q = lane / nslot;

    for (unsigned pass0 = t0; pass0 < t1; pass0 += nslot * kP2PTargets) {
        const unsigned npass = t1 - pass0 < nslot * kP2PTargets ? t1 - pass0 : nslot * kP2PTargets;
        const int tpl = (int)((npass + nslot - 1) / nslot);
        T x[kP2PTargets][D], acc[kP2PTargets];
#pragma unroll
        for (int k = 0; k < kP2PTargets; ++k) {
            acc[k] = T(0);
            const unsigned i = pass0 + s + nslot * k;
            const unsigned ic = i < t1 ? i : t1 - 1;
#pragma unroll
            for (int d = 0; d < D; ++d) x[k][d] = k < tpl ? tp[d * nt + ic] : T(0);
        }
        for (unsigned base = r_lo; base < r_hi; base += kP2PBlock) {
            const unsigned cnt = r_hi - base < (unsigned)kP2PBlock ? r_hi - base : (unsigned)kP2PBlock;
            __syncthreads();  // previous block fully consumed
            int b = 0;
            for (unsigned e = lane; e < cnt; e += 64) {
                const unsigned p = base + e;
                while (nb_pre[b + 1] <= p) ++b;
                const long j = (long)nb_first[b] + (p - nb_pre[b]);
                T *r = src + e * RW;
#pragma unroll
                for (int d = 0; d < D; ++d) r[d] = sp[d * ns + j];
                r[D] = pair_weight<K>(sw[j]);
            }
            __syncthreads();
            const unsigned ns_blk = nself > base ? nself - base : 0u;
            switch (tpl) {
            case 1: p2p_block<1>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 2: p2p_block<2>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 3: p2p_block<3>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 4: p2p_block<4>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 5: p2p_block<5>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 6: p2p_block<6>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 7: p2p_block<7>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 8: p2p_block<8>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 9: p2p_block<9>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 10: p2p_block<10>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 11: p2p_block<11>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            default: p2p_block<12>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            }
        }
        // sum the slices of each target: lanes s, s + nslot, s + 2 nslot, ...
#pragma unroll
        for (int k = 0; k < kP2PTargets; ++k) {
            if (k < tpl) {
                T v = acc[k];
                for (unsigned o = nslot; o < 64; o <<= 1) v += __shfl_xor(v, (int)o);
                acc[k] = v;
            }
        }
        if (q == 0) {
#pragma unroll
            for (int k = 0; k < kP2PTargets; ++k) {
                const unsigned i = pass0 + s + nslot * k;
                if (k >= tpl || i >= t1) continue;
                if (nchunk > 1) {  // one chunk of a split box: its partial sum
                    static_cast<T *>(a.split_part)[(long)chunk * nt + i] = acc[k];
                    continue;
                }
                T *ans = static_cast<T *>(a.ans);
                const unsigned oi = a.t_perm[i];
                if (a.base)
                    ans[oi] = static_cast<const T *>(a.base)[i] + acc[k];
                else if (a.accumulate)
                    ans[oi] += acc[k];
                else
                    ans[oi] = acc[k];
            }
        }
    }
}

// ------------------------------------------------------ P2P, mutual form
//
// Targets are the sources (the same sorted points) and f(y,x) = sym f(x,y).
// One wave per box A.  The wave evaluates A against itself as the direct
// form does, then each neighbour B that follows A in the row-major 3^D block
// (the upper half, (3^D-1)/2 offsets) once for both directions: w_j f(x_t,y_j)
// goes to target t's sum and sym w_t f(x_t,y_j) to B's point j.
//
// Lane l is (slot s = l % S, slice q = l / S), S = 4 or 16 slots: slot s
// owns the box's targets s, s+S, ... (at most kP2PMutualTargets) in registers.
// B's staged records are taken in chunks of S; slice q takes every (64/S)-th
// chunk.  In step t of a chunk, lane s evaluates record (s + t) mod S of the
// chunk and adds to a running partial sum for that record, which then moves
// one lane down the slot ring (a DPP rotate).  After S steps each record's
// partial has visited every slot once and sits in the lane with its index:
// one DPP move per record and lane replaces a reduction per record.
//
// Output, sorted order: slots[u][j] is what point j of box B received from
// box B - o_u (written by that box's wave, or zero by B's own wave when that
// box is outside the grid or empty), slots[n_upper][i] point i's own sum.
// p2p_combine (boxops.hip) adds the n_upper + 1 values of each point.
constexpr int kP2PMutualTargets = 10;  // targets per lane per pass: one pass up to 160 points

template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int CTRL>
__device__ __forceinline__ float dpp_move(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Lane s of each ring of S lanes gets the value of lane (s + 1) mod S.
template <int S, typename T>
__device__ __forceinline__ T ring_next(T v)
{
    static_assert(S == 4 || S == 16, "rings of 4 (quad) or 16 (row) lanes");
    if constexpr (S == 4)
        return dpp_move<0x39>(v);  // quad_perm [1,2,3,0]
    else
        return dpp_move<0x12F>(v);  // row_ror:15
}

// Chunks c = q, q + nslice, ... of the mutual records starting at m.
template <int S, int TPL, typename T, int D, int NT, class K>
__device__ __forceinline__ void p2p_chunks(const T *src, T *rev, unsigned m, unsigned nch, unsigned s, unsigned q,
                                           const T (&x)[NT][D], const T (&wt)[NT], T (&acc)[NT], const K &kern)
{
    constexpr int RW = p2p_rec<D>::W;
    constexpr unsigned nslice = 64 / S;
    for (unsigned c = q; c < nch; c += nslice) {
        const unsigned rb = m + c * S;
        T part = T(0);
#pragma unroll
        for (unsigned t = 0; t < (unsigned)S; ++t) {
            const T *r = src + (rb + ((s + t) & (S - 1))) * RW;
            T y[D];
#pragma unroll
            for (int d = 0; d < D; ++d) y[d] = r[d];
            const T wj = r[D];
#pragma unroll
            for (int k = 0; k < TPL; ++k) {
                const T f = pair_eval<true, T, D>(kern, x[k], y);
                acc[k] += wj * f;
                part += wt[k] * f;
            }
            part = ring_next<S>(part);
        }
        rev[rb + s] = part;
    }
}

template <int S, int TPL, typename T, int D, int NT, class K>
__device__ __forceinline__ void p2p_block_mutual(const T *src, T *rev, unsigned m, unsigned nch, unsigned s,
                                                 unsigned q, const T (&x)[NT][D], const T (&wt)[NT], T (&acc)[NT],
                                                 const K &kern)
{
    if (m > 0) p2p_range<TPL, false>(src, 0, m, q, 64 / S, x, acc, kern);
    p2p_chunks<S, TPL>(src, rev, m, nch, s, q, x, wt, acc, kern);
}

template <int S, typename T, int D, int NT, class K>
__device__ __forceinline__ void p2p_block_mutual_tpl(int tpl, const T *src, T *rev, unsigned m, unsigned nch,
                                                     unsigned s, unsigned q, const T (&x)[NT][D], const T (&wt)[NT],
                                                     T (&acc)[NT], const K &kern)
{
    static_assert(NT == 10, "the switch covers 1..10 targets per lane");
    switch (tpl) {
    case 1: p2p_block_mutual<S, 1>(src, rev, m, nch, s, q, x, wt, acc, kern); break;
    case 2: p2p_block_mutual<S, 2>(src, rev, m, nch, s, q, x, wt, acc, kern); break;
    case 3: p2p_block_mutual<S, 3>(src, rev, m, nch, s, q, x, wt, acc, kern); break;
    case 4: p2p_block_mutual<S, 4>(src, rev, m, nch, s, q, x, wt, acc, kern); break;
    case 5: p2p_block_mutual<S, 5>(src, rev, m, nch, s, q, x, wt, acc, kern); break;
    case 6: p2p_block_mutual<S, 6>(src, rev, m, nch, s, q, x, wt, acc, kern); break;
    case 7: p2p_block_mutual<S, 7>(src, rev, m, nch, s, q, x, wt, acc, kern); break;
    case 8: p2p_block_mutual<S, 8>(src, rev, m, nch, s, q, x, wt, acc, kern); break;
    case 9: p2p_block_mutual<S, 9>(src, rev, m, nch, s, q, x, wt, acc, kern); break;
    default: p2p_block_mutual<S, 10>(src, rev, m, nch, s, q, x, wt, acc, kern); break;
    }
}

// Two waves per SIMD: with ten targets per lane the 1/r kernel then keeps
// every value in registers (at three waves, 168 VGPRs, it spills), measured
// 4.00-4.06 vs 4.06-4.14 ms at 128 points per box, N = 4M.
template <typename T, int D, class K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void p2p_mutual_kernel(fmma_p2p_args a,
                                                                                                 K kern)
{
    constexpr int NNB = D == 1 ? 3 : D == 2 ? 9 : 27;
    constexpr int NU = (NNB - 1) / 2;
    constexpr int RW = p2p_rec<D>::W;
    constexpr int MT = kP2PMutualTargets;
    constexpr unsigned kPad = 16;  // a chunk's padding past the block
    __shared__ __align__(16) T src[(kP2PBlock + kPad) * RW];
    __shared__ T rev[kP2PBlock + kPad];
    __shared__ unsigned nb_first[NU + 1], nb_pre[NU + 2];

    const unsigned box = blockIdx.x;
    const unsigned t0 = a.t_start[box], t1 = a.t_start[box + 1];
    if (t0 == t1) return;
    const unsigned lane = threadIdx.x;
    const T *pos = static_cast<const T *>(a.s_pos);
    const T *pw = static_cast<const T *>(a.s_w);
    T *slots = static_cast<T *>(a.slots);
    const long n = a.n_t;
    constexpr T sym = (T)kernel_symmetry<K>::value;

    // lane 0: the box itself; lane 1 + u (u < NU): the upper neighbour B =
    // box + o_u, o_u = entry NU + 1 + u of the row-major 3^D block (last
    // dimension fastest, centre = entry NU); lane NU + 1 + u: the lower
    // neighbour box - o_u, whose wave writes slot u of this box's points
    unsigned cnt_nb = 0;
    bool lower_missing = false;
    if (lane <= (unsigned)(2 * NU)) {
        const bool upper = lane <= (unsigned)NU;
        int bc[D], m[D];
        unsigned gbase;
        const int NL = box_coords_of<D>(a, box, bc, gbase);
        bool ok = true;
        int k = lane == 0 ? NU : upper ? NU + (int)lane : (int)lane;
        const int sgn = upper ? 1 : -1;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            m[d] = bc[d] + sgn * ((k % 3) - 1);
            k /= 3;
            ok = ok && m[d] >= 0 && m[d] < NL;
        }
        unsigned first = 0, cnt = 0;
        if (ok) {
            const unsigned mm = box_index_of<D>(a, m, gbase);
            first = a.s_start[mm];
            cnt = a.s_start[mm + 1] - first;
        }
        if (upper) {
            cnt_nb = cnt;
            nb_first[lane] = first;
        } else {
            lower_missing = cnt == 0;
        }
    }
    const unsigned long long missing = __ballot(lower_missing) >> (NU + 1);
    unsigned incl = cnt_nb;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const unsigned y = __shfl_up(incl, o);
        if ((int)lane >= o) incl += y;
    }
    if (lane <= (unsigned)NU) nb_pre[lane + 1] = incl;
    if (lane == 0) nb_pre[0] = 0;
    // no wave of a missing (or empty) lower neighbour writes these slots
    for (int u = 0; u < NU; ++u)
        if ((missing >> u) & 1ull)
            for (unsigned i = t0 + lane; i < t1; i += 64) slots[(long)u * n + i] = T(0);
    __syncthreads();
    const unsigned total = nb_pre[NU + 1], nself = nb_pre[1];

    // ring size: 4 slots for boxes of <= 4 MT points, else 16 with balanced passes
    const unsigned ntb = t1 - t0;
    const bool small = ntb <= 4u * MT;
    const unsigned S = small ? 4u : 16u, s = lane % S, q = lane / S;
    const unsigned npasses = small ? 1u : (ntb + 16u * MT - 1) / (16u * MT);
    const unsigned per = (ntb + npasses - 1) / npasses;

    for (unsigned pass = 0; pass < npasses; ++pass) {
        const unsigned p0 = t0 + pass * per, p1 = p0 + per < t1 ? p0 + per : t1;
        if (p0 >= t1) break;
        const int tpl = (int)((p1 - p0 + S - 1) / S);
        T x[MT][D], wt[MT], acc[MT];
#pragma unroll
        for (int k = 0; k < MT; ++k) {
            acc[k] = T(0);
            const unsigned i = p0 + s + S * k;
            const unsigned ic = i < p1 ? i : p1 - 1;
#pragma unroll
            for (int d = 0; d < D; ++d) x[k][d] = k < tpl ? pos[d * n + ic] : T(0);
            wt[k] = k < tpl && i < p1 ? pair_weight<K>(pw[i]) : T(0);  // padding targets add nothing to B
        }
        for (unsigned base = 0; base < total; base += kP2PBlock) {
            const unsigned cnt = total - base < (unsigned)kP2PBlock ? total - base : (unsigned)kP2PBlock;
            const unsigned m = nself > base ? (nself - base < cnt ? nself - base : cnt) : 0u;
            const unsigned nch = (cnt - m + S - 1) / S;
            const unsigned cpad = m + nch * S;  // <= cnt + S - 1
            __syncthreads();                    // previous block fully consumed
            int b = 0;
            for (unsigned e = lane; e < cpad; e += 64) {
                // records past cnt repeat record cnt - 1 (of a neighbour box) with weight 0
                const unsigned p = base + (e < cnt ? e : cnt - 1);
                while (nb_pre[b + 1] <= p) ++b;
                const long j = (long)nb_first[b] + (p - nb_pre[b]);
                T *r = src + e * RW;
#pragma unroll
                for (int d = 0; d < D; ++d) r[d] = pos[d * n + j];
                r[D] = e < cnt ? pair_weight<K>(pw[j]) : T(0);
            }
            __syncthreads();
            if (small)
                p2p_block_mutual_tpl<4>(tpl, src, rev, m, nch, s, q, x, wt, acc, kern);
            else
                p2p_block_mutual_tpl<16>(tpl, src, rev, m, nch, s, q, x, wt, acc, kern);
            __syncthreads();
            // the neighbours' sums of this block (records m .. cnt): the first
            // pass stores, later passes add (one thread owns an entry in every pass)
            b = 0;
            for (unsigned e = lane; e < cnt; e += 64) {
                if (e < m) continue;
                const unsigned p = base + e;
                while (nb_pre[b + 1] <= p) ++b;
                const long j = (long)nb_first[b] + (p - nb_pre[b]);
                T *dst = slots + (long)(b - 1) * n + j;
                const T v = sym * rev[e];
                *dst = pass == 0 ? v : *dst + v;
            }
        }
        // the box's own sums: add the slices of each target (lanes s + S q')
#pragma unroll
        for (int k = 0; k < MT; ++k) {
            if (k < tpl) {
                T v = acc[k];
                for (unsigned o = S; o < 64; o <<= 1) v += __shfl_xor(v, (int)o);
                acc[k] = v;
            }
        }
        if (q == 0) {
#pragma unroll
            for (int k = 0; k < MT; ++k) {
                const unsigned i = p0 + s + S * k;
                if (k < tpl && i < p1) slots[(long)NU * n + i] = acc[k];
            }
        }
    }
}

template <typename T, int D, class K>
int launch_p2p(const fmma_p2p_args *a, const K &k, hipStream_t s)
{
    if (a->n_target_boxes <= 0) return FMMA_OK;
    if (a->slots) {
        if constexpr (kernel_symmetry<K>::value == 1 || kernel_symmetry<K>::value == -1) {
            if (a->t_pos != a->s_pos || a->t_start != a->s_start || a->n_t != a->n_s ||
                a->symmetry != kernel_symmetry<K>::value)
                return FMMA_ERR_INVALID;
            hipLaunchKernelGGL((p2p_mutual_kernel<T, D, K>), dim3((unsigned)a->n_target_boxes), dim3(64), 0, s, *a,
                               k);
        } else {
            return FMMA_ERR_INVALID;
        }
    } else {
        const long nwg = a->split_items ? a->n_split_items : a->n_target_boxes;
        if (nwg <= 0) return FMMA_OK;
        hipLaunchKernelGGL((p2p_kernel<T, D, K>), dim3((unsigned)nwg), dim3(64), 0, s, *a, k);
    }
    return hipGetLastError() == hipSuccess ? FMMA_OK : FMMA_ERR_HIP;
}

// ------------------------------------------------- balanced direct P2P
//
// With clustered points one target box can have hundreds of times the
// average near field, and its single workgroup then sets the kernel's time.
// The planner gives every box whose targets x neighbourhood records exceed
// kP2PSplitWork pair evaluations a set of work items: passes of at most
// `pass` targets, times chunks of the neighbourhood records sized so that
// an item does about kP2PSplitWork evaluations (at most kP2PSplitMaxChunks).
// Phase 0 marks the split boxes (parts = their chunk count) and counts the
// items and the largest chunk count; phase 1, once the library has sized
// the item list, writes the items.  Item order does not matter: every item
// writes its own sums, and the chunks of a target are added in chunk order.
constexpr unsigned long long kP2PSplitWork = 1ull << 21;
constexpr unsigned kP2PSplitMaxChunks = 64;

template <int D, class A>
__global__ __launch_bounds__(256) void p2p_split_plan_kernel(A a, unsigned pass, unsigned *parts, unsigned *counters,
                                                             unsigned *items, int phase)
{
    constexpr int NNB = D == 1 ? 3 : D == 2 ? 9 : 27;
    const long box = (long)blockIdx.x * 256 + threadIdx.x;
    if (box >= a.n_target_boxes) return;
    const unsigned ntb = a.t_start[box + 1] - a.t_start[box];
    unsigned total = 0;
    if (ntb > 0) {
        int bc[D];
        unsigned gbase;
        const int NL = box_coords_of<D>(a, (unsigned)box, bc, gbase);
        for (int k = 0; k < NNB; ++k) {
            int m[D], kk = k;
            bool ok = true;
#pragma unroll
            for (int d = D - 1; d >= 0; --d) {
                m[d] = bc[d] + (kk % 3) - 1;
                kk /= 3;
                ok = ok && m[d] >= 0 && m[d] < NL;
            }
            if (ok) {
                const unsigned mm = box_index_of<D>(a, m, gbase);
                total += a.s_start[mm + 1] - a.s_start[mm];
            }
        }
    }
    const unsigned long long work = (unsigned long long)ntb * total;
    if (work <= kP2PSplitWork) {
        if (phase == 0) parts[box] = 0;
        return;
    }
    const unsigned npass = (ntb + pass - 1) / pass, tp = ntb < pass ? ntb : pass;
    unsigned long long nch = ((unsigned long long)tp * total + kP2PSplitWork - 1) / kP2PSplitWork;
    nch = nch < 1 ? 1 : nch > kP2PSplitMaxChunks ? kP2PSplitMaxChunks : nch;
    const unsigned nitem = npass * (unsigned)nch;
    if (phase == 0) {
        parts[box] = (unsigned)nch;
        atomicAdd(&counters[0], nitem);
        atomicMax(&counters[1], (unsigned)nch);
    } else {
        const unsigned q = atomicAdd(&counters[2], nitem);
        for (unsigned p = 0; p < npass; ++p)
            for (unsigned c = 0; c < (unsigned)nch; ++c) {
                unsigned *it = items + 3 * (size_t)(q + p * (unsigned)nch + c);
                it[0] = (unsigned)box;
                it[1] = p;
                it[2] = c;
            }
    }
}

// The chunk sums of split boxes' targets, in chunk order (thread per sorted
// target; keys = its box): a solve writes base + sum (or adds the sum to ans
// or overwrites, as the P2P pass would), a gradient (dim > 0) writes the D
// components base[d] + sum.
template <typename T>
__global__ __launch_bounds__(256) void p2p_split_combine(const unsigned *keys, const unsigned *parts, const T *part,
                                                         long n, const T *base, const unsigned *perm, T *out,
                                                         int accumulate, int dim)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const unsigned nch = parts[keys[i]];
    if (nch <= 1) return;
    const long oi = perm[i];
    if (dim == 0) {
        T v = part[i];
        for (unsigned c = 1; c < nch; ++c) v += part[(long)c * n + i];
        if (base)
            out[oi] = base[i] + v;
        else if (accumulate)
            out[oi] += v;
        else
            out[oi] = v;
        return;
    }
    for (int d = 0; d < dim; ++d) {
        T v = part[(long)d * n + i];
        for (unsigned c = 1; c < nch; ++c) v += part[((long)c * dim + d) * n + i];
        out[oi * dim + d] = base ? base[(long)d * n + i] + v : v;
    }
}

// ------------------------------------------------------ P2P gradient
//
// The near field of the gradient with respect to the target coordinates,
// laid out as p2p_kernel (one wave per target box, slots x slices, sources
// of the 3^D neighbourhood staged through LDS) with D sums per target.
constexpr int kP2PGradTargets = 8;

template <int TPL, bool FAR, typename T, int D, int NT, class K>
__device__ __forceinline__ void p2p_grad_range(const T *src, unsigned j0, unsigned j1, unsigned q, unsigned nslice,
                                               const T (&x)[NT][D], T (&acc)[NT][D], const K &kern)
{
    constexpr int RW = p2p_rec<D>::W;
    const unsigned r0 = j0 % nslice;
    for (unsigned j = j0 + (q >= r0 ? q - r0 : q + nslice - r0); j < j1; j += nslice) {
        const T *r = src + j * RW;
        T y[D];
#pragma unroll
        for (int d = 0; d < D; ++d) y[d] = r[d];
        const T wj = r[D];
#pragma unroll
        for (int k = 0; k < TPL; ++k) {
            T g[D];
            pair_grad<FAR, T, D>(kern, x[k], y, g);
#pragma unroll
            for (int d = 0; d < D; ++d) acc[k][d] += wj * g[d];
        }
    }
}

template <int TPL, typename T, int D, int NT, class K>
__device__ __forceinline__ void p2p_grad_block(const T *src, unsigned cnt, unsigned nself, unsigned q,
                                               unsigned nslice, const T (&x)[NT][D], T (&acc)[NT][D], const K &kern)
{
    const unsigned m = nself < cnt ? nself : cnt;
    if (m > 0) p2p_grad_range<TPL, false>(src, 0, m, q, nslice, x, acc, kern);
    p2p_grad_range<TPL, true>(src, m, cnt, q, nslice, x, acc, kern);
}

template <typename T, int D, class K>
__global__ __launch_bounds__(64) void p2p_grad_kernel(fmma_p2p_grad_args a, K kern)
{
    constexpr int NNB = D == 1 ? 3 : D == 2 ? 9 : 27;
    constexpr int RW = p2p_rec<D>::W;
    constexpr int MT = kP2PGradTargets;
    __shared__ __align__(16) T src[kP2PBlock * RW];
    __shared__ unsigned nb_first[NNB], nb_pre[NNB + 1];

    unsigned box, t0, t1, chunk, nchunk;
    if (!p2p_work_item(a, box, t0, t1, chunk, nchunk)) return;
    const unsigned lane = threadIdx.x;
    const T *tp = static_cast<const T *>(a.t_pos);
    const T *sp = static_cast<const T *>(a.s_pos);
    const T *sw = static_cast<const T *>(a.s_w);
    const T *base = static_cast<const T *>(a.base);
    T *grad = static_cast<T *>(a.grad);
    const long nt = a.n_t, ns = a.n_s;

    // neighbour ranges, the box itself first (as p2p_kernel)
    unsigned cnt_nb = 0;
    if (lane < (unsigned)NNB) {
        int bc[D], m[D];
        unsigned gbase;
        const int NL = box_coords_of<D>(a, box, bc, gbase);
        bool ok = true;
        constexpr int centre = (NNB - 1) / 2;
        int k = lane == 0 ? centre : (int)lane <= centre ? (int)lane - 1 : (int)lane;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            m[d] = bc[d] + (k % 3) - 1;
            k /= 3;
            ok = ok && m[d] >= 0 && m[d] < NL;
        }
        unsigned first = 0;
        if (ok) {
            const unsigned mm = box_index_of<D>(a, m, gbase);
            first = a.s_start[mm];
            cnt_nb = a.s_start[mm + 1] - first;
        }
        nb_first[lane] = first;
    }
    unsigned incl = cnt_nb;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const unsigned y = __shfl_up(incl, o);
        if ((int)lane >= o) incl += y;
    }
    if (lane < (unsigned)NNB) nb_pre[lane + 1] = incl;
    if (lane == 0) nb_pre[0] = 0;
    __syncthreads();
    const unsigned total = nb_pre[NNB], nself = nb_pre[1];
    unsigned r_lo, r_hi;
    chunk_range(total, chunk, nchunk, r_lo, r_hi);

    const unsigned ntb = t1 - t0;
    unsigned nslot = 4;
    while (nslot < 64 && nslot * MT < ntb) nslot <<= 1;
    const unsigned nslice = 64 / nslot, s = lane % nslot, q = lane / nslot;

    for (unsigned pass0 = t0; pass0 < t1; pass0 += nslot * MT) {
        const unsigned npass = t1 - pass0 < nslot * MT ? t1 - pass0 : nslot * MT;
        const int tpl = (int)((npass + nslot - 1) / nslot);
        T x[MT][D], acc[MT][D];
#pragma unroll
        for (int k = 0; k < MT; ++k) {
            const unsigned i = pass0 + s + nslot * k;
            const unsigned ic = i < t1 ? i : t1 - 1;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                acc[k][d] = T(0);
                x[k][d] = k < tpl ? tp[d * nt + ic] : T(0);
            }
        }
        for (unsigned b0 = r_lo; b0 < r_hi; b0 += kP2PBlock) {
            const unsigned cnt = r_hi - b0 < (unsigned)kP2PBlock ? r_hi - b0 : (unsigned)kP2PBlock;
            __syncthreads();
            int b = 0;
            for (unsigned e = lane; e < cnt; e += 64) {
                const unsigned p = b0 + e;
                while (nb_pre[b + 1] <= p) ++b;
                const long j = (long)nb_first[b] + (p - nb_pre[b]);
                T *r = src + e * RW;
#pragma unroll
                for (int d = 0; d < D; ++d) r[d] = sp[d * ns + j];
                r[D] = sw[j];
            }
            __syncthreads();
            const unsigned ns_blk = nself > b0 ? nself - b0 : 0u;
            switch (tpl) {
            case 1: p2p_grad_block<1>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 2: p2p_grad_block<2>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 3: p2p_grad_block<3>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 4: p2p_grad_block<4>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 5: p2p_grad_block<5>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 6: p2p_grad_block<6>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            case 7: p2p_grad_block<7>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            default: p2p_grad_block<8>(src, cnt, ns_blk, q, nslice, x, acc, kern); break;
            }
        }
#pragma unroll
        for (int k = 0; k < MT; ++k) {
            if (k < tpl) {
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    T v = acc[k][d];
                    for (unsigned o = nslot; o < 64; o <<= 1) v += __shfl_xor(v, (int)o);
                    acc[k][d] = v;
                }
            }
        }
        if (q == 0) {
#pragma unroll
            for (int k = 0; k < MT; ++k) {
                const unsigned i = pass0 + s + nslot * k;
                if (k >= tpl || i >= t1) continue;
                if (nchunk > 1) {  // one chunk of a split box: its partial sums
#pragma unroll
                    for (int d = 0; d < D; ++d)
                        static_cast<T *>(a.split_part)[((long)chunk * D + d) * nt + i] = acc[k][d];
                    continue;
                }
                const long oi = a.t_perm[i];
#pragma unroll
                for (int d = 0; d < D; ++d) grad[oi * D + d] = base ? base[d * nt + i] + acc[k][d] : acc[k][d];
            }
        }
    }
}

template <typename T, int D, class K>
int launch_p2p_grad(const fmma_p2p_grad_args *a, const K &k, hipStream_t s)
{
    if constexpr (has_grad<K, T, D>::value) {
        const long nwg = a->split_items ? a->n_split_items : a->n_target_boxes;
        if (nwg <= 0) return FMMA_OK;
        hipLaunchKernelGGL((p2p_grad_kernel<T, D, K>), dim3((unsigned)nwg), dim3(64), 0, s, *a, k);
        return hipGetLastError() == hipSuccess ? FMMA_OK : FMMA_ERR_HIP;
    } else {
        return FMMA_ERR_UNSUPPORTED;
    }
}

// All-pairs gradient: one thread per target, sources in ascending order.
template <typename T, int D, class K>
__global__ __launch_bounds__(256) void exact_grad_kernel(fmma_exact_grad_args a, K kern)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n_t) return;
    const T *tg = static_cast<const T *>(a.targets);
    const T *sr = static_cast<const T *>(a.sources);
    const T *w = static_cast<const T *>(a.weights);
    T x[D], acc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        x[d] = tg[i * D + d];
        acc[d] = T(0);
    }
    for (long j = 0; j < a.n_s; ++j) {
        T y[D], g[D];
#pragma unroll
        for (int d = 0; d < D; ++d) y[d] = sr[j * D + d];
        kern.template grad<T, D>(x, y, g);
        const T wj = w ? w[j] : T(1);
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d] += wj * g[d];
    }
    T *out = static_cast<T *>(a.grad);
#pragma unroll
    for (int d = 0; d < D; ++d) out[i * D + d] = acc[d];
}

template <typename T, int D, class K>
int launch_exact_grad(const fmma_exact_grad_args *a, const K &k, hipStream_t s)
{
    if constexpr (has_grad<K, T, D>::value) {
        if (a->n_t <= 0) return FMMA_OK;
        hipLaunchKernelGGL((exact_grad_kernel<T, D, K>), dim3((unsigned)((a->n_t + 255) / 256)), dim3(256), 0, s, *a,
                           k);
        return hipGetLastError() == hipSuccess ? FMMA_OK : FMMA_ERR_HIP;
    } else {
        return FMMA_ERR_UNSUPPORTED;
    }
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
__device__ __forceinline__ T m2l_op_value(const fmma_m2l_op_args &a, T Len, int level, int oi, int kl, int km,
                                          const K &kern)
{
    const T len = Len / (T)(1u << level);
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
    return kern.template operator()<T, D>(X, Y);
}

template <typename T, int D, class K>
__global__ __launch_bounds__(256) void m2l_op_kernel(fmma_m2l_op_args a, K kern)
{
    const int oi = blockIdx.y;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const long PP = (long)a.P * a.P;
    if (e >= PP) return;
    const int kl = (int)(e / a.P), km = (int)(e % a.P);
    static_cast<T *>(a.K)[oi * PP + e] = m2l_op_value<T, D>(a, *static_cast<const T *>(a.Len), a.level, oi, kl, km, kern);
}

// Fragment index of element (row, col) of one offset's block.
__device__ __forceinline__ long frag_index(int format, int KP, int row, int col)
{
    if (format == FMMA_FRAG_BF16W) {
        const int lane = (row & 31) + 32 * ((col & 15) >> 3);
        return ((((long)(row >> 5) * (KP / 16) + (col >> 4)) * 64 + lane) << 3) + (col & 7);
    }
    if (format == FMMA_FRAG_F32W) {
        const int lane = (row & 31) + 32 * (col & 1), s = col >> 1;
        return ((((long)(row >> 5) * (KP / 8) + (s >> 2)) * 64 + lane) << 2) + (s & 3);
    }
    if (format == FMMA_FRAG_BF16) {
        const int lane = (row & 15) + 16 * ((col & 31) >> 3);
        return ((((long)(row >> 4) * (KP / 32) + (col >> 5)) * 64 + lane) << 3) + (col & 7);
    }
    const int lane = (row & 15) + 16 * (col & 3);
    return (((long)(row >> 4) * (KP / 4) + (col >> 2)) << 6) + lane;
}

// Operator sets of several (problem, level) pairs straight into fragment
// order, with the same values and roundings as m2l_op_kernel followed by the
// library's fragment conversion.  Each wave covers 64 consecutive fragment
// elements: 16 rows of K_o and 4 columns (16x16x4 formats: columns 4 s ..
// 4 s + 3; F32W: every second column from 8 g + lane / 32) or 8 (BF16).  Their node positions X (rows) and Y (columns) are computed once
// per wave into LDS with exactly m2l_op_value's expressions; each thread then
// evaluates one f(X, Y).  PAIRED: blockIdx.y is a pair (o, -o) and
// K_{-o}[col][row] = sym * K_o[row][col] is stored from the same value
// (translation invariance and f(y,x) = sym f(x,y)), halving the evaluations.
template <typename T, int D, class K, bool PAIRED>
__global__ __launch_bounds__(256) void m2l_op_frag_kernel(fmma_m2l_op_args a, K kern)
{
    __shared__ T sx[4][16][D], sy[4][8][D];
    const int b = blockIdx.z, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int oi = PAIRED ? a.pairs[2 * blockIdx.y] : (int)blockIdx.y;
    const int per = a.MP * a.KP;  // at most 512 x 512
    const bool bfw = a.format == FMMA_FRAG_BF16W;
    const bool bf = a.format == FMMA_FRAG_BF16 || bfw, wide = a.format == FMMA_FRAG_F32W;
    const int rw = (int)blockIdx.x * 256 + wave * 64;  // the wave's first element
    // The wave's elements are rows rbase + rs and columns cbase + cstep cs,
    // (rs, cs) from the lane alone (the fragment order of frag_rowcol).
    int rbase, cbase, cstep, rs, cs;
    if (bfw) {
        const int l0 = (rw >> 3) & 63, q = rw >> 9;
        const int st = q % (a.KP / 16), rb = q / (a.KP / 16);
        rbase = 32 * rb + (l0 & 31);
        cbase = 16 * st + 8 * (l0 >> 5);
        cstep = 1;
        rs = lane >> 3;
        cs = lane & 7;
    } else if (wide) {
        const int l0 = (rw >> 2) & 63, q = rw >> 8;
        const int g = q % (a.KP / 8), rb = q / (a.KP / 8);
        rbase = 32 * rb + (l0 & 31);
        cbase = 8 * g + (l0 >> 5);
        cstep = 2;
        rs = lane >> 2;
        cs = lane & 3;
    } else {
        const int nks = bf ? a.KP / 32 : a.KP / 4;
        const int q = bf ? (rw >> 9) : (rw >> 6);
        const int s = q % nks, rb = q / nks;
        rbase = 16 * rb;
        cbase = bf ? 32 * s + 8 * (((rw >> 3) & 63) >> 4) : 4 * s;
        cstep = 1;
        rs = bf ? ((rw >> 3) & 15) + (lane >> 3) : lane & 15;
        cs = bf ? lane & 7 : lane >> 4;
    }
    const int ncol = bf ? 8 : 4;
    const T Len = static_cast<const T *>(a.Len)[(long)a.batch_problem[b] * a.len_stride];
    const T len = Len / (T)(1u << a.batch_level[b]);
    if (lane < 16 + ncol) {
        const bool isrow = lane < 16;
        const int k = isrow ? rbase + lane : cbase + cstep * (lane - 16);
        if (k < a.P) {
            int dk = k;
#pragma unroll
            for (int d = D - 1; d >= 0; --d) {
                const T c = (T)a.nodes[dk % a.n_nodes];
                dk /= a.n_nodes;
                if (isrow)
                    sx[wave][lane][d] = len * (c / 2);
                else
                    sy[wave][lane - 16][d] = len * ((T)a.offsets[oi * D + d] + c / 2);
            }
        }
    }
    __syncthreads();
    const int r = rw + lane;
    if (r >= per) return;
    const int row = rbase + rs, col = cbase + cstep * cs;
    const bool in = row < a.P && col < a.P;
    T v = T(0);
    if (in) v = kern.template operator()<T, D>(sx[wave][rs], sy[wave][cs]);
    auto store = [&](int off, long e, T x) {
        const long g = ((long)b * a.n_off + off) * per + e;
        if (a.format == FMMA_FRAG_F64) {
            static_cast<double *>(a.Kf)[g] = (double)x;
        } else if (a.format == FMMA_FRAG_F32 || a.format == FMMA_FRAG_F32W) {
            static_cast<float *>(a.Kf)[g] = (float)x;
        } else {
            const float fv = (float)x;
            const __bf16 h = (__bf16)fv;
            static_cast<__bf16 *>(a.Kf)[g] = h;
            static_cast<__bf16 *>(a.Kf_lo)[g] = (__bf16)(fv - (float)h);
        }
    };
    store(oi, r, v);
    if constexpr (PAIRED) {
        constexpr T sym = (T)kernel_symmetry<K>::value;
        const int ni = a.pairs[2 * blockIdx.y + 1];
        if (in)
            store(ni, frag_index(a.format, a.KP, col, row), sym * v);
        else
            store(ni, r, T(0));  // padding of K_{-o} at the same place
    }
}

template <typename T, int D, class K>
int launch_m2l_op(const fmma_m2l_op_args *a, const K &k, hipStream_t s)
{
    if (a->format == FMMA_FRAG_RAW) {
        const long PP = (long)a->P * a->P;
        hipLaunchKernelGGL((m2l_op_kernel<T, D, K>), dim3((unsigned)((PP + 255) / 256), (unsigned)a->n_off),
                           dim3(256), 0, s, *a, k);
    } else {
        if (a->n_batch <= 0) return FMMA_OK;
        if (a->format == FMMA_FRAG_F64 && sizeof(T) != 8) return FMMA_ERR_INVALID;
        const long per = (long)a->MP * a->KP;
        const unsigned nb = (unsigned)((per + 255) / 256);
        constexpr bool sym = kernel_symmetry<K>::value == 1 || kernel_symmetry<K>::value == -1;
        if (sym && a->pairs && 2 * a->n_pairs == a->n_off && K::translation_invariant)
            hipLaunchKernelGGL((m2l_op_frag_kernel<T, D, K, sym>), dim3(nb, (unsigned)a->n_pairs, (unsigned)a->n_batch),
                               dim3(256), 0, s, *a, k);
        else
            hipLaunchKernelGGL((m2l_op_frag_kernel<T, D, K, false>),
                               dim3(nb, (unsigned)a->n_off, (unsigned)a->n_batch), dim3(256), 0, s, *a, k);
    }
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

// ----------------------------------------------------------------- M2P
//
// The far field of the tree and nrnmm solvers.  One wave per target box
// holds the box's targets in registers as the direct P2P does (slot/slice
// lanes, at most kP2PTargets per lane).  Far boxes are enumerated 64
// candidates at a time -- tree: for each level 2..leaf_level the 6^D children
// of the 3^D block around the parent of the target's box at that level;
// nrnmm: the whole grid, row-major -- and those at Chebyshev distance >= 2
// are compacted (ballot + popcount) into an LDS list of (corner, side,
// multipole offset).  The list's (node position, multipole) records are then
// staged through LDS in blocks of kP2PBlock and taken by every target, with
// f's far() form (a node of a far box never coincides with a target).
constexpr int kM2PList = 192;  // far boxes per list; flushed before another 64 could overflow it
constexpr int kM2PMaxNodes = 16;

template <typename T, int D, class K>
__device__ __forceinline__ void p2p_block_far(int tpl, const T *src, unsigned cnt, unsigned q, unsigned nslice,
                                              const T (&x)[kP2PTargets][D], T (&acc)[kP2PTargets], const K &kern)
{
    switch (tpl) {
    case 1: p2p_range<1, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    case 2: p2p_range<2, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    case 3: p2p_range<3, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    case 4: p2p_range<4, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    case 5: p2p_range<5, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    case 6: p2p_range<6, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    case 7: p2p_range<7, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    case 8: p2p_range<8, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    case 9: p2p_range<9, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    case 10: p2p_range<10, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    case 11: p2p_range<11, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    default: p2p_range<12, true>(src, 0, cnt, q, nslice, x, acc, kern); break;
    }
}

template <typename T, int D, class K>
__global__ __launch_bounds__(64) void m2p_kernel(fmma_m2p_args a, K kern)
{
    constexpr int RW = p2p_rec<D>::W;
    constexpr int C6 = D == 1 ? 6 : D == 2 ? 36 : 216;  // candidates per tree level
    __shared__ __align__(16) T src[kP2PBlock * RW];
    __shared__ T ent_rel[kM2PList * D];
    __shared__ T ent_len[kM2PList];
    __shared__ long ent_off[kM2PList];
    __shared__ double hc[kM2PMaxNodes];

    const unsigned box = blockIdx.x;
    const unsigned t0 = a.t_start[box], t1 = a.t_start[box + 1];
    if (t0 == t1) return;
    const unsigned lane = threadIdx.x;
    const unsigned nn = (unsigned)a.n_nodes, P = (unsigned)a.P;
    // (c_j + 1)/2 of the TYPE node, in double as the reference forms it
    if (lane < nn) hc[lane] = ((double)(T)a.nodes[lane] + 1.0) / 2.0;
    const T *tp = static_cast<const T *>(a.t_pos);
    const T *Wm = static_cast<const T *>(a.Wm);
    const T *org = static_cast<const T *>(a.origin);
    const T Len = *static_cast<const T *>(a.Len);
    const long nt = a.n_t;
    const int gn = a.grid_n, L = a.leaf_level;

    int bc[D];
    long ncand;
    if (gn > 0) {
        unsigned r = box;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            bc[d] = (int)(r % (unsigned)gn);
            r /= (unsigned)gn;
        }
        ncand = 1;
#pragma unroll
        for (int d = 0; d < D; ++d) ncand *= gn;
    } else {
        morton_decode<D>(box, bc);
        ncand = L >= 2 ? (long)(L - 1) * C6 : 0;
    }

    const unsigned ntb = t1 - t0;
    unsigned nslot = 4;
    while (nslot < 64 && nslot * kP2PTargets < ntb) nslot <<= 1;
    const unsigned nslice = 64 / nslot, s = lane % nslot, q = lane / nslot;

    for (unsigned pass0 = t0; pass0 < t1; pass0 += nslot * kP2PTargets) {
        const unsigned npass = t1 - pass0 < nslot * kP2PTargets ? t1 - pass0 : nslot * kP2PTargets;
        const int tpl = (int)((npass + nslot - 1) / nslot);
        T x[kP2PTargets][D], acc[kP2PTargets];
#pragma unroll
        for (int k = 0; k < kP2PTargets; ++k) {
            acc[k] = T(0);
            const unsigned i = pass0 + s + nslot * k;
            const unsigned ic = i < t1 ? i : t1 - 1;
#pragma unroll
            for (int d = 0; d < D; ++d) x[k][d] = k < tpl ? tp[d * nt + ic] : T(0);
        }
        int list_n = 0;
        for (long c0 = 0; c0 < ncand; c0 += 64) {
            const long c = c0 + lane;
            bool ok = false;
            int m[D], ref[D], side = 1;
            long lbase = 0;
            if (c < ncand) {
                if (gn > 0) {
                    long r = c;
#pragma unroll
                    for (int d = D - 1; d >= 0; --d) {
                        m[d] = (int)(r % gn);
                        r /= gn;
                        ref[d] = bc[d];
                    }
                    side = gn;
                } else {
                    const int l = 2 + (int)(c / C6);
                    int j = (int)(c - (long)(l - 2) * C6);
#pragma unroll
                    for (int d = D - 1; d >= 0; --d) {
                        ref[d] = bc[d] >> (L - l);
                        m[d] = 2 * ((ref[d] >> 1) - 1) + j % 6;
                        j /= 6;
                    }
                    side = 1 << l;
                    lbase = (long)P * (((1L << (D * l)) - 1) / ((1L << D) - 1));
                }
                int far = 0;
                ok = true;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    ok = ok && m[d] >= 0 && m[d] < side;
                    const int dd = m[d] > ref[d] ? m[d] - ref[d] : ref[d] - m[d];
                    far = dd > far ? dd : far;
                }
                ok = ok && far >= 2;
            }
            const unsigned long long mask = __ballot(ok);
            if (ok) {
#pragma clang fp contract(off)
                const int pos = list_n + (int)__popcll(mask & ((1ull << lane) - 1ull));
                const T len = Len / (T)side;
#pragma unroll
                for (int d = 0; d < D; ++d) ent_rel[pos * D + d] = len * (T)m[d] + org[d];
                ent_len[pos] = len;
                ent_off[pos] = gn > 0 ? c * (long)P : lbase + (long)morton_encode<D>(m) * P;
            }
            list_n += (int)__popcll(mask);
            if (list_n > kM2PList - 64 || c0 + 64 >= ncand) {
                const long nrec = (long)list_n * P;
                for (long rb = 0; rb < nrec; rb += kP2PBlock) {
                    const unsigned cnt = nrec - rb < kP2PBlock ? (unsigned)(nrec - rb) : (unsigned)kP2PBlock;
                    __syncthreads();  // list written, previous block consumed
                    for (unsigned e = lane; e < cnt; e += 64) {
#pragma clang fp contract(off)
                        const unsigned p = (unsigned)rb + e;
                        const unsigned ent = p / P, k = p - ent * P;
                        const double ln = (double)ent_len[ent];
                        T *r = src + e * RW;
                        unsigned kk = k;
#pragma unroll
                        for (int d = D - 1; d >= 0; --d) {
                            const unsigned dg = kk % nn;
                            kk /= nn;
                            r[d] = (T)(hc[dg] * ln + (double)ent_rel[ent * D + d]);
                        }
                        r[D] = pair_weight<K>(Wm[ent_off[ent] + k]);
                    }
                    __syncthreads();
                    p2p_block_far<T, D>(tpl, src, cnt, q, nslice, x, acc, kern);
                }
                __syncthreads();  // the list is rewritten next
                list_n = 0;
            }
        }
#pragma unroll
        for (int k = 0; k < kP2PTargets; ++k) {
            if (k < tpl) {
                T v = acc[k];
                for (unsigned o = nslot; o < 64; o <<= 1) v += __shfl_xor(v, (int)o);
                acc[k] = v;
            }
        }
        if (q == 0) {
            T *out = static_cast<T *>(a.out);
#pragma unroll
            for (int k = 0; k < kP2PTargets; ++k) {
                const unsigned i = pass0 + s + nslot * k;
                if (k < tpl && i < t1) out[i] = acc[k];
            }
        }
    }
}

template <typename T, int D, class K>
int launch_m2p(const fmma_m2p_args *a, const K &k, hipStream_t s)
{
    if (a->n_target_boxes <= 0 || a->n_t <= 0) return FMMA_OK;
    if (a->n_nodes < 1 || a->n_nodes > kM2PMaxNodes) return FMMA_ERR_UNSUPPORTED;
    hipLaunchKernelGGL((m2p_kernel<T, D, K>), dim3((unsigned)a->n_target_boxes), dim3(64), 0, s, *a, k);
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
    static int p2p_grad(const fmma_p2p_grad_args *a, void *ctx, void *s)
    {
        return launch_p2p_grad<T, D, K>(a, *static_cast<const K *>(ctx), (hipStream_t)s);
    }
    static int exact_grad(const fmma_exact_grad_args *a, void *ctx, void *s)
    {
        return launch_exact_grad<T, D, K>(a, *static_cast<const K *>(ctx), (hipStream_t)s);
    }
    static int m2p(const fmma_m2p_args *a, void *ctx, void *s)
    {
        return launch_m2p<T, D, K>(a, *static_cast<const K *>(ctx), (hipStream_t)s);
    }
    static fmma_kernel_ops make(const K *ctx)
    {
        fmma_kernel_ops o;
        o.dim = D;
        o.dtype = sizeof(T) == 8 ? FMMA_F64 : FMMA_F32;
        o.translation_invariant = K::translation_invariant ? 1 : 0;
        o.symmetry = kernel_symmetry<K>::value;
        o.ctx = const_cast<K *>(ctx);
        o.p2p = &p2p;
        o.exact = &exact;
        o.m2l_op = &m2l_op;
        o.m2l_general = &m2l_general;
        o.p2p_grad = has_grad<K, T, D>::value ? &p2p_grad : nullptr;
        o.exact_grad = has_grad<K, T, D>::value ? &exact_grad : nullptr;
        o.transposed = nullptr;
        o.m2p = &m2p;
        return o;
    }
};

}  // namespace detail
}  // namespace fmma
