// boxops.hip -- P2M, M2M, L2L, L2P on Morton-ordered boxes.
//
//   P2M  src/fmm_core.cpp:224-283   Wm[b][k] = sum_s w_s prod_d S_n(c_{k_d}, r_{s,d})
//   M2M  src/fmm_core.cpp:295-363   Wm[p][ko] += sum_sigma sum_ki prod_d S_n(c_ko, (c_ki + 2 sigma_d - 1)/2) Wm[c][ki]
//   L2L  src/fmm_core.cpp:507-568   Wl[c][ko] += sum_ki prod_d S_n((c_ko + 2 sigma_d - 1)/2, c_ki) Wl[p][ki]
//   L2P  src/fmm_core.cpp:580-615   u[t] = sum_k Wl[b][k] prod_d S_n(r_{t,d}, c_{k_d})
//
// One wave per box.  Particle data is structure-of-arrays in sorted order
// (coalesced); per-particle interpolation weights and the 1-D transfer tables
// sit in LDS; each lane owns KPL of the P coefficients of the box.  In Morton
// order the 2^D children of a parent are contiguous, so M2M reads and L2L
// writes one contiguous block per parent.
#include "boxops.hpp"

namespace fmma {
namespace lib {

template <int D, int KPL>
__device__ __forceinline__ void coeff_digits(int lane, int nn, int P, int (&dig)[KPL][D])
{
#pragma unroll
    for (int r = 0; r < KPL; ++r) {
        int k = lane + 64 * r;
        k = k < P ? k : 0;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            dig[r][d] = k % nn;
            k /= nn;
        }
    }
}

// ---------------------------------------------------------------------- P2M

template <typename T, int D, int KPL>
__global__ __launch_bounds__(64) void p2m_kernel(P2MArgs<T> a)
{
    __shared__ T sc[kMaxNodes * kMaxNodes];
    __shared__ T sv[64 * D * kMaxNodes];
    const unsigned box = blockIdx.x;
    const int lane = threadIdx.x, nn = a.nn, P = a.P;
    for (int i = lane; i < nn * nn; i += 64) sc[i] = a.coef[i];
    const unsigned s0 = a.start[box], s1 = a.start[box + 1];
    int dig[KPL][D];
    coeff_digits<D, KPL>(lane, nn, P, dig);
    T acc[KPL];
#pragma unroll
    for (int r = 0; r < KPL; ++r) acc[r] = T(0);
    const int NL = 1 << a.leaf_level;
    const T len = a.geom->len_leaf;
    __syncthreads();
    for (unsigned base = s0; base < s1; base += 64) {
        const unsigned i = base + lane;
        if (i < s1) {
            const T wt = a.w[i];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                T rel;
                leaf_coord_rel(a.pos[d * a.n + i], a.geom->origin[d], len, NL, rel);
                cheb_weights(sc, nn, rel, d == 0 ? wt : T(1), sv + (lane * D + d) * nn, 1);
            }
        }
        __syncthreads();
        const int cnt = (int)(s1 - base < 64u ? s1 - base : 64u);
        for (int p = 0; p < cnt; ++p) {
            const T *S = sv + p * D * nn;
#pragma unroll
            for (int r = 0; r < KPL; ++r) {
                T v = S[dig[r][0]];
#pragma unroll
                for (int d = 1; d < D; ++d) v *= S[d * nn + dig[r][d]];
                acc[r] += v;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < KPL; ++r) {
        const int k = lane + 64 * r;
        if (k < P) a.Wm[(size_t)box * P + k] = acc[r];
    }
}

// ---------------------------------------------------------------------- M2M

template <typename T, int D, int KPL>
__global__ __launch_bounds__(64) void m2m_kernel(const T *Wc, T *Wp, const T *tab, int nn, int P)
{
    __shared__ T st[2 * kMaxNodes * kMaxNodes];
    __shared__ T sw[(1 << D) * kMaxP];
    const unsigned parent = blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < 2 * nn * nn; i += 64) st[i] = tab[i];
    const T *src = Wc + (size_t)parent * (1u << D) * P;
    for (int i = lane; i < (1 << D) * P; i += 64) sw[i] = src[i];
    int dig[KPL][D];
    coeff_digits<D, KPL>(lane, nn, P, dig);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < KPL; ++r) {
        T acc = T(0);
        for (int sg = 0; sg < (1 << D); ++sg) {
            const T *w = sw + sg * P;
            const T *t0 = st + ((((sg >> (D - 1)) & 1) * nn + dig[r][0]) * nn);
            if constexpr (D == 1) {
                for (int a0 = 0; a0 < nn; ++a0) acc += t0[a0] * w[a0];
            } else if constexpr (D == 2) {
                const T *t1 = st + (((sg & 1) * nn + dig[r][1]) * nn);
                for (int a0 = 0; a0 < nn; ++a0) {
                    const T c0 = t0[a0];
                    for (int a1 = 0; a1 < nn; ++a1) acc += c0 * t1[a1] * w[a0 * nn + a1];
                }
            } else {
                const T *t1 = st + ((((sg >> 1) & 1) * nn + dig[r][1]) * nn);
                const T *t2 = st + (((sg & 1) * nn + dig[r][2]) * nn);
                for (int a0 = 0; a0 < nn; ++a0) {
                    const T c0 = t0[a0];
                    for (int a1 = 0; a1 < nn; ++a1) {
                        const T c01 = c0 * t1[a1];
                        const T *wr = w + (a0 * nn + a1) * nn;
                        for (int a2 = 0; a2 < nn; ++a2) acc += c01 * t2[a2] * wr[a2];
                    }
                }
            }
        }
        const int k = lane + 64 * r;
        if (k < P) Wp[(size_t)parent * P + k] = acc;
    }
}

// ---------------------------------------------------------------------- L2L
// L2L[s][ko][ki] = S_n((c_ko + 2s - 1)/2, c_ki) = tab[s][ki][ko] (S_n is symmetric).

template <typename T, int D, int KPL>
__global__ __launch_bounds__(64) void l2l_kernel(const T *Wp, T *Wc, const T *tab, int nn, int P)
{
    __shared__ T st[2 * kMaxNodes * kMaxNodes];
    __shared__ T sp[kMaxP];
    const unsigned parent = blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < 2 * nn * nn; i += 64) st[i] = tab[i];
    for (int i = lane; i < P; i += 64) sp[i] = Wp[(size_t)parent * P + i];
    int dig[KPL][D];
    coeff_digits<D, KPL>(lane, nn, P, dig);
    __syncthreads();
    for (int sg = 0; sg < (1 << D); ++sg) {
        T *dst = Wc + ((size_t)parent * (1u << D) + sg) * P;
#pragma unroll
        for (int r = 0; r < KPL; ++r) {
            T acc = T(0);
            const int s0 = (sg >> (D - 1)) & 1;
            if constexpr (D == 1) {
                for (int a0 = 0; a0 < nn; ++a0) acc += st[(s0 * nn + a0) * nn + dig[r][0]] * sp[a0];
            } else if constexpr (D == 2) {
                const int s1 = sg & 1;
                for (int a0 = 0; a0 < nn; ++a0) {
                    const T c0 = st[(s0 * nn + a0) * nn + dig[r][0]];
                    for (int a1 = 0; a1 < nn; ++a1)
                        acc += c0 * st[(s1 * nn + a1) * nn + dig[r][1]] * sp[a0 * nn + a1];
                }
            } else {
                const int s1 = (sg >> 1) & 1, s2 = sg & 1;
                for (int a0 = 0; a0 < nn; ++a0) {
                    const T c0 = st[(s0 * nn + a0) * nn + dig[r][0]];
                    for (int a1 = 0; a1 < nn; ++a1) {
                        const T c01 = c0 * st[(s1 * nn + a1) * nn + dig[r][1]];
                        const T *pr = sp + (a0 * nn + a1) * nn;
                        for (int a2 = 0; a2 < nn; ++a2) acc += c01 * st[(s2 * nn + a2) * nn + dig[r][2]] * pr[a2];
                    }
                }
            }
            const int k = lane + 64 * r;
            if (k < P) dst[k] += acc;
        }
    }
}

// ---------------------------------------------------------------------- L2P

template <typename T, int D>
__global__ __launch_bounds__(64) void l2p_kernel(L2PArgs<T> a)
{
    __shared__ T sc[kMaxNodes * kMaxNodes];
    __shared__ T sW[kMaxP];
    __shared__ T sv[64 * D * kMaxNodes];
    const unsigned box = blockIdx.x;
    const unsigned t0 = a.start[box], t1 = a.start[box + 1];
    if (t0 == t1) return;
    const int lane = threadIdx.x, nn = a.nn, P = a.P;
    for (int i = lane; i < nn * nn; i += 64) sc[i] = a.coef[i];
    for (int i = lane; i < P; i += 64) sW[i] = a.Wl[(size_t)box * P + i];
    __syncthreads();
    const int NL = 1 << a.leaf_level;
    const T len = a.geom->len_leaf;
    T *S = sv + lane * D * nn;
    for (unsigned base = t0; base < t1; base += 64) {
        const unsigned i = base + lane;
        if (i >= t1) break;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            T rel;
            leaf_coord_rel(a.pos[d * a.n + i], a.geom->origin[d], len, NL, rel);
            cheb_weights(sc, nn, rel, T(1), S + d * nn, 1);
        }
        T acc = T(0);
        if constexpr (D == 1) {
            for (int a0 = 0; a0 < nn; ++a0) acc += sW[a0] * S[a0];
        } else if constexpr (D == 2) {
            for (int a0 = 0; a0 < nn; ++a0)
                for (int a1 = 0; a1 < nn; ++a1) acc += sW[a0 * nn + a1] * (S[a0] * S[nn + a1]);
        } else {
            for (int a0 = 0; a0 < nn; ++a0)
                for (int a1 = 0; a1 < nn; ++a1) {
                    const T s01 = S[a0] * S[nn + a1];
                    const T *wr = sW + (a0 * nn + a1) * nn;
                    for (int a2 = 0; a2 < nn; ++a2) acc += wr[a2] * (s01 * S[2 * nn + a2]);
                }
        }
        a.out[i] = acc;
    }
}

// ------------------------------------------------------------------ launchers

static int kpl_of(int P) { return P <= 64 ? 1 : P <= 128 ? 2 : P <= 256 ? 4 : 8; }

#define FMMA_KPL_SWITCH(P, CALL)      \
    switch (kpl_of(P)) {              \
    case 1: { constexpr int KPL = 1; CALL; } break; \
    case 2: { constexpr int KPL = 2; CALL; } break; \
    case 4: { constexpr int KPL = 4; CALL; } break; \
    default: { constexpr int KPL = 8; CALL; } break; \
    }

template <typename T, int D>
int p2m(const P2MArgs<T> &a, long nbox, hipStream_t st)
{
    if (nbox <= 0) return FMMA_OK;
    FMMA_KPL_SWITCH(a.P, hipLaunchKernelGGL((p2m_kernel<T, D, KPL>), dim3((unsigned)nbox), dim3(64), 0, st, a));
    FMMA_LAUNCHED("p2m");
    return FMMA_OK;
}

template <typename T, int D>
int m2m(const T *Wc, T *Wp, long nparent, const T *tab, int nn, int P, hipStream_t st)
{
    if (nparent <= 0) return FMMA_OK;
    FMMA_KPL_SWITCH(P, hipLaunchKernelGGL((m2m_kernel<T, D, KPL>), dim3((unsigned)nparent), dim3(64), 0, st, Wc,
                                          Wp, tab, nn, P));
    FMMA_LAUNCHED("m2m");
    return FMMA_OK;
}

template <typename T, int D>
int l2l(const T *Wp, T *Wc, long nparent, const T *tab, int nn, int P, hipStream_t st)
{
    if (nparent <= 0) return FMMA_OK;
    FMMA_KPL_SWITCH(P, hipLaunchKernelGGL((l2l_kernel<T, D, KPL>), dim3((unsigned)nparent), dim3(64), 0, st, Wp,
                                          Wc, tab, nn, P));
    FMMA_LAUNCHED("l2l");
    return FMMA_OK;
}

template <typename T, int D>
int l2p(const L2PArgs<T> &a, long nbox, hipStream_t st)
{
    if (nbox <= 0) return FMMA_OK;
    hipLaunchKernelGGL((l2p_kernel<T, D>), dim3((unsigned)nbox), dim3(64), 0, st, a);
    FMMA_LAUNCHED("l2p");
    return FMMA_OK;
}

#define FMMA_BOX_INST(T, D)                                                                 \
    template int p2m<T, D>(const P2MArgs<T> &, long, hipStream_t);                          \
    template int m2m<T, D>(const T *, T *, long, const T *, int, int, hipStream_t);         \
    template int l2l<T, D>(const T *, T *, long, const T *, int, int, hipStream_t);         \
    template int l2p<T, D>(const L2PArgs<T> &, long, hipStream_t);
FMMA_BOX_INST(double, 1)
FMMA_BOX_INST(double, 2)
FMMA_BOX_INST(double, 3)
FMMA_BOX_INST(float, 1)
FMMA_BOX_INST(float, 2)
FMMA_BOX_INST(float, 3)

}  // namespace lib
}  // namespace fmma
