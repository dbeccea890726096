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

// The geometry and leaf level of a leaf box: the plan's, or in a batch of
// independent problems those of the box's problem.
template <typename T, class A>
__device__ __forceinline__ int box_geom(const A &a, unsigned box, const Geom<T> *&g)
{
    if (a.box_prob) {
        const unsigned p = a.box_prob[box];
        g = a.geoms + p;
        return a.levels[p];
    }
    g = a.geom;
    return a.leaf_level;
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
    const Geom<T> *gm;
    const int lv = box_geom<T>(a, box, gm), NL = a.grid_n > 0 ? a.grid_n : 1 << lv;
    const T len = gm->len_leaf;
    __syncthreads();
    for (unsigned base = s0; base < s1; base += 64) {
        const unsigned i = base + lane;
        if (i < s1) {
            const T wt = a.w[i];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                T rel;
                leaf_coord_rel(a.pos[d * a.n + i], gm->origin[d], len, NL, rel);
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
__global__ __launch_bounds__(64) void m2m_kernel(const T *Wc, T *Wp, const T *tab, int nn, int P, const long *items)
{
    __shared__ T st[2 * kMaxNodes * kMaxNodes];
    __shared__ T sw[(1 << D) * kMaxP];
    const unsigned parent = blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < 2 * nn * nn; i += 64) st[i] = tab[i];
    const T *src = Wc + (items ? (size_t)items[2 * parent] : (size_t)parent * (1u << D) * P);
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
        if (k < P) Wp[(items ? (size_t)items[2 * parent + 1] : (size_t)parent * P) + k] = acc;
    }
}

// ---------------------------------------------------------------------- L2L
// L2L[s][ko][ki] = S_n((c_ko + 2s - 1)/2, c_ki) = tab[s][ki][ko] (S_n is symmetric).

template <typename T, int D, int KPL>
__global__ __launch_bounds__(64) void l2l_kernel(const T *Wp, T *Wc, const T *tab, int nn, int P, const long *items)
{
    __shared__ T st[2 * kMaxNodes * kMaxNodes];
    __shared__ T sp[kMaxP];
    const unsigned parent = blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < 2 * nn * nn; i += 64) st[i] = tab[i];
    for (int i = lane; i < P; i += 64) sp[i] = Wp[(items ? (size_t)items[2 * parent + 1] : (size_t)parent * P) + i];
    int dig[KPL][D];
    coeff_digits<D, KPL>(lane, nn, P, dig);
    __syncthreads();
    for (int sg = 0; sg < (1 << D); ++sg) {
        T *dst = Wc + (items ? (size_t)items[2 * parent] : (size_t)parent * (1u << D) * P) + (size_t)sg * P;
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
    const Geom<T> *gm;
    const int NL = 1 << box_geom<T>(a, box, gm);
    const T len = gm->len_leaf;
    T *S = sv + lane * D * nn;
    for (unsigned base = t0; base < t1; base += 64) {
        const unsigned i = base + lane;
        if (i >= t1) break;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            T rel;
            leaf_coord_rel(a.pos[d * a.n + i], gm->origin[d], len, NL, rel);
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

// ======================================================================
// Compile-time node count NN (2..8): the per-particle interpolation weights
// live in registers or right-sized LDS, and M2M / L2L apply the 1-D
// transfer matrices one dimension at a time (D*NN multiply-adds per
// coefficient and child instead of NN^D).

template <typename T, int NN>
__device__ __forceinline__ void cheb_weights_ct(const T *coef, T r, T scale, T (&out)[NN])
{
    T tj[NN];
    tj[0] = T(1);
    if constexpr (NN > 1) tj[1] = r;
#pragma unroll
    for (int j = 2; j < NN; ++j) tj[j] = T(2) * r * tj[j - 1] - tj[j - 2];
#pragma unroll
    for (int k = 0; k < NN; ++k) {
        T v = T(0);
#pragma unroll
        for (int j = 0; j < NN; ++j) v += coef[k * NN + j] * tj[j];
        out[k] = v * scale;
    }
}

constexpr int ipow_c(int b, int e) { return e == 0 ? 1 : b * ipow_c(b, e - 1); }

// P2M: one wave per source box.  The box's points are staged 64 at a time,
// each lane computing one point's Chebyshev weights S_d (D x NN, the first
// scaled by the point's weight) into LDS.  Then the multipole
// Wm[a0..a_{D-1}] = sum_p S_0[p][a0] ... S_{D-1}[p][a_{D-1}] is accumulated as
// a sum of outer products: lane (g, c) holds the trailing index tuple c of
// the last D - 1 dimensions (NN^(D-1) of them) and all NN coefficients along
// dimension 0 in registers, and takes every G-th point (g = lane group, G =
// 64 / NN^(D-1) groups); the groups are added at the end.  Per point and lane
// that is one product of D - 1 weights and NN multiply-adds, for NN
// coefficients.
template <typename T, int D, int NN>
__global__ __launch_bounds__(64) void p2m_ct(P2MArgs<T> a)
{
    constexpr int P = ipow_c(NN, D);
    constexpr int NT = ipow_c(NN, D - 1);  // trailing index tuples
    constexpr int G = NT <= 64 ? 64 / NT : 1;
    static_assert(NT <= 64, "at most 64 trailing index tuples (8 nodes in 3-D)");
    // rows of an odd number of values: the 64 lanes writing their rows hit
    // different banks (an even stride of 16 doubles put every lane on the
    // same two banks)
    constexpr int SV = (D * NN) | 1;
    __shared__ T sc[NN * NN];
    __shared__ T sv[64][SV];
    __shared__ T red[G > 1 ? G - 1 : 1][NT * NN];
    const unsigned box = blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < NN * NN; i += 64) sc[i] = a.coef[i];
    const unsigned s0 = a.start[box], s1 = a.start[box + 1];
    const int g = lane / NT, c = lane - g * NT;
    const bool act = g < G;
    int dig[D > 1 ? D - 1 : 1];
    {
        int k = act ? c : 0;
#pragma unroll
        for (int d = D - 1; d >= 1; --d) {
            dig[d - 1] = k % NN;
            k /= NN;
        }
    }
    T acc[NN];
#pragma unroll
    for (int k = 0; k < NN; ++k) acc[k] = T(0);
    const Geom<T> *gm;
    const int lv = box_geom<T>(a, box, gm), NL = a.grid_n > 0 ? a.grid_n : 1 << lv;
    const T len = gm->len_leaf;
    __syncthreads();
    for (unsigned base = s0; base < s1; base += 64) {
        const unsigned i = base + lane;
        if (i < s1) {
            const T wt = a.w[i];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                T rel;
                leaf_coord_rel(a.pos[d * a.n + i], gm->origin[d], len, NL, rel);
                T out[NN];
                cheb_weights_ct<T, NN>(sc, rel, d == 0 ? wt : T(1), out);
#pragma unroll
                for (int k = 0; k < NN; ++k) sv[lane][d * NN + k] = out[k];
            }
        }
        __syncthreads();
        const int cnt = (int)(s1 - base < 64u ? s1 - base : 64u);
        if (act) {
            for (int p = g; p < cnt; p += G) {
                T t = T(1);
#pragma unroll
                for (int d = 1; d < D; ++d) t *= sv[p][d * NN + dig[d - 1]];
#pragma unroll
                for (int k = 0; k < NN; ++k) acc[k] += sv[p][k] * t;
            }
        }
        __syncthreads();
    }
    // groups 1..G-1 hand their sums to group 0
    if constexpr (G > 1) {
        if (act && g > 0)
#pragma unroll
            for (int k = 0; k < NN; ++k) red[g - 1][k * NT + c] = acc[k];
        __syncthreads();
        if (g == 0)
#pragma unroll
            for (int k = 0; k < NN; ++k)
                for (int h = 1; h < G; ++h) acc[k] += red[h - 1][k * NT + c];
    }
    if (g == 0)
#pragma unroll
        for (int k = 0; k < NN; ++k) a.Wm[(size_t)box * P + k * NT + c] = acc[k];
}

// P2M in 3-D for NN <= 6 with fewer LDS reads per multiply-add: lane (g, a2)
// keeps all NN^2 sums (a0, a1) of its last index a2 in registers and takes
// every G-th point of a staged chunk (G = 64 / NN groups): per point it reads
// the point's NN weights of dimensions 0 and 1 (the same addresses for the
// lanes of a group) and one weight of dimension 2, forms S1[a1] S2[a2] and
// adds S0[a0] times that to the NN^2 sums.  The groups are added at the end
// in a fixed order through LDS.
template <typename T, int NN>
__global__ __launch_bounds__(64) void p2m3_ct(P2MArgs<T> a)
{
    constexpr int D = 3, P = NN * NN * NN, G = 64 / NN, SV = (D * NN) | 1;
    constexpr int NS = 64 * SV > G * P ? 64 * SV : G * P;
    __shared__ T sc[NN * NN];
    __shared__ T smem[NS];  // the staged weights, then the groups' sums
    T(*sv)[SV] = reinterpret_cast<T(*)[SV]>(smem);
    T(*red)[P] = reinterpret_cast<T(*)[P]>(smem);
    const unsigned box = blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < NN * NN; i += 64) sc[i] = a.coef[i];
    const unsigned s0 = a.start[box], s1 = a.start[box + 1];
    const int g = lane / NN, a2 = lane - g * NN;
    const bool act = g < G;
    T acc[NN][NN];
#pragma unroll
    for (int i = 0; i < NN; ++i)
#pragma unroll
        for (int j = 0; j < NN; ++j) acc[i][j] = T(0);
    const Geom<T> *gm;
    const int lv = box_geom<T>(a, box, gm), NL = a.grid_n > 0 ? a.grid_n : 1 << lv;
    const T len = gm->len_leaf;
    __syncthreads();
    for (unsigned base = s0; base < s1; base += 64) {
        const unsigned i = base + lane;
        if (i < s1) {
            const T wt = a.w[i];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                T rel;
                leaf_coord_rel(a.pos[d * a.n + i], gm->origin[d], len, NL, rel);
                T out[NN];
                cheb_weights_ct<T, NN>(sc, rel, d == 0 ? wt : T(1), out);
#pragma unroll
                for (int k = 0; k < NN; ++k) sv[lane][d * NN + k] = out[k];
            }
        }
        __syncthreads();
        const int cnt = (int)(s1 - base < 64u ? s1 - base : 64u);
        if (act) {
            for (int p = g; p < cnt; p += G) {
                const T w2 = sv[p][2 * NN + a2];
                T t1[NN];
#pragma unroll
                for (int j = 0; j < NN; ++j) t1[j] = sv[p][NN + j] * w2;
#pragma unroll
                for (int k = 0; k < NN; ++k) {
                    const T w0 = sv[p][k];
#pragma unroll
                    for (int j = 0; j < NN; ++j) acc[k][j] += w0 * t1[j];
                }
            }
        }
        __syncthreads();
    }
    if (act)
#pragma unroll
        for (int k = 0; k < NN; ++k)
#pragma unroll
            for (int j = 0; j < NN; ++j) red[g][(k * NN + j) * NN + a2] = acc[k][j];
    __syncthreads();
    for (int e = lane; e < P; e += 64) {
        T v = red[0][e];
        for (int h = 1; h < G; ++h) v += red[h][e];
        a.Wm[(size_t)box * P + e] = v;
    }
}

// out[..k_d..] = sum_c M[k_d][c] in[..c..] along dimension d of an NN^D
// tensor (dimension 0 most significant); M is NN x NN (row = output index).
template <typename T, int D, int NN>
__device__ __forceinline__ void apply_dim(const T *M, bool transpose, int d, const T *in, T *out, int lane)
{
    constexpr int P = ipow_c(NN, D);
    const int stride = d == D - 1 ? 1 : d == D - 2 ? NN : NN * NN;
    for (int e = lane; e < P; e += 64) {
        const int kd = (e / stride) % NN;
        const int base = e - kd * stride;
        T v = T(0);
#pragma unroll
        for (int c = 0; c < NN; ++c) v += (transpose ? M[c * NN + kd] : M[kd * NN + c]) * in[base + c * stride];
        out[e] = v;
    }
}

// M2M: one wave per parent.  For each child sigma the transfer
// prod_d S_n(c_ko, (c_ki + 2 sigma_d - 1)/2) is applied dimension by
// dimension through LDS and summed over the children in registers.
template <typename T, int D, int NN>
__global__ __launch_bounds__(64) void m2m_ct(const T *Wc, T *Wp, const T *tab, const long *items)
{
    constexpr int P = ipow_c(NN, D), KPL = (P + 63) / 64, NC = 1 << D;
    __shared__ T st[2 * NN * NN];
    __shared__ T sw[NC * P];
    __shared__ T tA[P], tB[P];
    const unsigned parent = blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < 2 * NN * NN; i += 64) st[i] = tab[i];
    const T *src = Wc + (items ? (size_t)items[2 * parent] : (size_t)parent * NC * P);
    for (int i = lane; i < NC * P; i += 64) sw[i] = src[i];
    T acc[KPL];
#pragma unroll
    for (int r = 0; r < KPL; ++r) acc[r] = T(0);
    __syncthreads();
    for (int sg = 0; sg < NC; ++sg) {
        const T *cur = sw + sg * P;
        // dimensions D-1 .. 1 through LDS
        T *bufs[2] = {tA, tB};
        int nb = 0;
        for (int d = D - 1; d >= 1; --d) {
            const int bit = (sg >> (D - 1 - d)) & 1;
            apply_dim<T, D, NN>(st + bit * NN * NN, false, d, cur, bufs[nb], lane);
            __syncthreads();
            cur = bufs[nb];
            nb ^= 1;
        }
        // dimension 0 into the accumulators
        const T *M = st + ((sg >> (D - 1)) & 1) * NN * NN;
        const int stride = ipow_c(NN, D - 1);
#pragma unroll
        for (int r = 0; r < KPL; ++r) {
            const int e = lane + 64 * r;
            if (e < P) {
                const int k0 = e / stride, base = e - k0 * stride;
                T v = T(0);
#pragma unroll
                for (int c = 0; c < NN; ++c) v += M[k0 * NN + c] * cur[base + c * stride];
                acc[r] += v;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < KPL; ++r) {
        const int k = lane + 64 * r;
        if (k < P) Wp[(items ? (size_t)items[2 * parent + 1] : (size_t)parent * P) + k] = acc[r];
    }
}

// L2L: one wave per parent; each child gets the transposed separable transfer.
template <typename T, int D, int NN>
__global__ __launch_bounds__(64) void l2l_ct(const T *Wp, T *Wc, const T *tab, const long *items)
{
    constexpr int P = ipow_c(NN, D), KPL = (P + 63) / 64, NC = 1 << D;
    __shared__ T st[2 * NN * NN];
    __shared__ T sp[P], tA[P], tB[P];
    const unsigned parent = blockIdx.x;
    const int lane = threadIdx.x;
    for (int i = lane; i < 2 * NN * NN; i += 64) st[i] = tab[i];
    for (int i = lane; i < P; i += 64) sp[i] = Wp[(items ? (size_t)items[2 * parent + 1] : (size_t)parent * P) + i];
    __syncthreads();
    for (int sg = 0; sg < NC; ++sg) {
        const T *cur = sp;
        T *bufs[2] = {tA, tB};
        int nb = 0;
        for (int d = D - 1; d >= 1; --d) {
            const int bit = (sg >> (D - 1 - d)) & 1;
            apply_dim<T, D, NN>(st + bit * NN * NN, true, d, cur, bufs[nb], lane);
            __syncthreads();
            cur = bufs[nb];
            nb ^= 1;
        }
        const T *M = st + ((sg >> (D - 1)) & 1) * NN * NN;
        const int stride = ipow_c(NN, D - 1);
        T *dst = Wc + (items ? (size_t)items[2 * parent] : (size_t)parent * NC * P) + (size_t)sg * P;
#pragma unroll
        for (int r = 0; r < KPL; ++r) {
            const int e = lane + 64 * r;
            if (e < P) {
                const int k0 = e / stride, base = e - k0 * stride;
                T v = T(0);
#pragma unroll
                for (int c = 0; c < NN; ++c) v += M[c * NN + k0] * cur[base + c * stride];
                dst[e] += v;
            }
        }
        __syncthreads();
    }
}

// M2M and L2L with one wave per child (2^D waves per parent): the same
// transfers as m2m_ct / l2l_ct, in the same order of operations, so the
// results are bit-identical, but the children's passes run side by side
// instead of one after another (a short serial chain per parent, which is
// what the few-parent upper levels are bound by).  M2M adds the children's
// contributions in child order through LDS.
template <int D, int NN>
constexpr bool kWideTransfer = ipow_c(NN, D) * (3 << D) * 8 <= 48 * 1024;

template <typename T, int D, int NN>
__global__ __launch_bounds__(64 << D) void m2m_w(const T *Wc, T *Wp, const T *tab, const long *items)
{
    constexpr int P = ipow_c(NN, D), KPL = (P + 63) / 64, NC = 1 << D;
    __shared__ T st[2 * NN * NN];
    __shared__ T sw[NC * P];  // the children's multipoles, then their contributions
    __shared__ T tA[NC][P], tB[NC][P];
    const unsigned parent = blockIdx.x;
    const int tid = threadIdx.x, sg = tid >> 6, lane = tid & 63;
    for (int i = tid; i < 2 * NN * NN; i += 64 * NC) st[i] = tab[i];
    const T *src = Wc + (items ? (size_t)items[2 * parent] : (size_t)parent * NC * P);
    for (int i = tid; i < NC * P; i += 64 * NC) sw[i] = src[i];
    __syncthreads();
    const T *cur = sw + sg * P;
    T *bufs[2] = {tA[sg], tB[sg]};
    int nb = 0;
    for (int d = D - 1; d >= 1; --d) {
        const int bit = (sg >> (D - 1 - d)) & 1;
        apply_dim<T, D, NN>(st + bit * NN * NN, false, d, cur, bufs[nb], lane);
        __syncthreads();
        cur = bufs[nb];
        nb ^= 1;
    }
    const T *M = st + ((sg >> (D - 1)) & 1) * NN * NN;
    constexpr int stride = ipow_c(NN, D - 1);
    T v[KPL];
#pragma unroll
    for (int r = 0; r < KPL; ++r) {
        const int e = lane + 64 * r;
        v[r] = T(0);
        if (e < P) {
            const int k0 = e / stride, base = e - k0 * stride;
#pragma unroll
            for (int c = 0; c < NN; ++c) v[r] += M[k0 * NN + c] * cur[base + c * stride];
        }
    }
    __syncthreads();  // every wave is done with sw
#pragma unroll
    for (int r = 0; r < KPL; ++r) {
        const int e = lane + 64 * r;
        if (e < P) sw[sg * P + e] = v[r];
    }
    __syncthreads();
    T *dst = Wp + (items ? (size_t)items[2 * parent + 1] : (size_t)parent * P);
    for (int e = tid; e < P; e += 64 * NC) {
        T acc = T(0);
#pragma unroll
        for (int c = 0; c < NC; ++c) acc += sw[c * P + e];
        dst[e] = acc;
    }
}

template <typename T, int D, int NN>
__global__ __launch_bounds__(64 << D) void l2l_w(const T *Wp, T *Wc, const T *tab, const long *items)
{
    constexpr int P = ipow_c(NN, D), KPL = (P + 63) / 64, NC = 1 << D;
    __shared__ T st[2 * NN * NN];
    __shared__ T sp[P];
    __shared__ T tA[NC][P], tB[NC][P];
    const unsigned parent = blockIdx.x;
    const int tid = threadIdx.x, sg = tid >> 6, lane = tid & 63;
    for (int i = tid; i < 2 * NN * NN; i += 64 * NC) st[i] = tab[i];
    for (int i = tid; i < P; i += 64 * NC)
        sp[i] = Wp[(items ? (size_t)items[2 * parent + 1] : (size_t)parent * P) + i];
    __syncthreads();
    const T *cur = sp;
    T *bufs[2] = {tA[sg], tB[sg]};
    int nb = 0;
    for (int d = D - 1; d >= 1; --d) {
        const int bit = (sg >> (D - 1 - d)) & 1;
        apply_dim<T, D, NN>(st + bit * NN * NN, true, d, cur, bufs[nb], lane);
        __syncthreads();
        cur = bufs[nb];
        nb ^= 1;
    }
    const T *M = st + ((sg >> (D - 1)) & 1) * NN * NN;
    constexpr int stride = ipow_c(NN, D - 1);
    T *dst = Wc + (items ? (size_t)items[2 * parent] : (size_t)parent * NC * P) + (size_t)sg * P;
#pragma unroll
    for (int r = 0; r < KPL; ++r) {
        const int e = lane + 64 * r;
        if (e < P) {
            const int k0 = e / stride, base = e - k0 * stride;
            T v = T(0);
#pragma unroll
            for (int c = 0; c < NN; ++c) v += M[c * NN + k0] * cur[base + c * stride];
            dst[e] += v;
        }
    }
}

// L2P over sorted targets: workgroup of 256 threads = 256 consecutive
// targets.  Their boxes form one contiguous Morton range (targets are
// sorted by box); when that range is short its locals are staged in LDS,
// otherwise they are read from global memory.  The two sources are separate
// code paths: through one pointer that may be either, every load would be a
// flat load waited for on its own.
constexpr int kL2PElems = 2048;  // LDS coefficients staged per workgroup

// The local expansion W (NN^D coefficients) at a point with Chebyshev
// weights S, contracted one dimension at a time, innermost (last) first:
// independent short sums instead of one chain over all coefficients.
template <typename T, int D, int NN, typename PW>
__device__ __forceinline__ T l2p_contract(PW W, const T (&S)[D][NN])
{
    T acc = T(0);
    if constexpr (D == 1) {
#pragma unroll
        for (int a0 = 0; a0 < NN; ++a0) acc += W[a0] * S[0][a0];
    } else if constexpr (D == 2) {
#pragma unroll
        for (int a0 = 0; a0 < NN; ++a0) {
            T t = T(0);
#pragma unroll
            for (int a1 = 0; a1 < NN; ++a1) t += W[a0 * NN + a1] * S[1][a1];
            acc += S[0][a0] * t;
        }
    } else {
#pragma unroll
        for (int a0 = 0; a0 < NN; ++a0) {
            T u = T(0);
#pragma unroll
            for (int a1 = 0; a1 < NN; ++a1) {
                T t = T(0);
#pragma unroll
                for (int a2 = 0; a2 < NN; ++a2) t += W[(a0 * NN + a1) * NN + a2] * S[2][a2];
                u += S[1][a1] * t;
            }
            acc += S[0][a0] * u;
        }
    }
    return acc;
}

// The L2P value of sorted target i (i < a.n) of a workgroup whose locals
// were staged (kb0 = the first box of the range) or not.
template <typename T, int D, int NN>
__device__ __forceinline__ T l2p_point(const L2PArgs<T> &a, long i, const T *sc, const T *sW, unsigned kb0,
                                       bool staged)
{
    constexpr int P = ipow_c(NN, D);
    const unsigned box = a.keys[i];
    const Geom<T> *gm;
    const int NL = 1 << box_geom<T>(a, box, gm);
    const T len = gm->len_leaf;
    T S[D][NN];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        T rel;
        leaf_coord_rel(a.pos[d * a.n + i], gm->origin[d], len, NL, rel);
        cheb_weights_ct<T, NN>(sc, rel, T(1), S[d]);
    }
    // explicit address spaces: the two contractions cannot be merged into
    // one over a generic pointer
    typedef const __attribute__((address_space(3))) T *LdsPtr;
    typedef const __attribute__((address_space(1))) T *GlobalPtr;
    if (staged) return l2p_contract<T, D, NN>((LdsPtr)(sW + (size_t)(box - kb0) * P), S);
    return l2p_contract<T, D, NN>((GlobalPtr)(a.Wl + (size_t)box * P), S);
}

// Stages the Chebyshev coefficients and, when the workgroup's boxes are few
// enough, their locals; returns whether it staged them.
template <typename T, int D, int NN, int NBX>
__device__ __forceinline__ bool l2p_stage(const L2PArgs<T> &a, long i0, T *sc, T *sW, unsigned &kb0)
{
    constexpr int P = ipow_c(NN, D);
    const long ilast = (i0 + 255 < a.n ? i0 + 255 : a.n - 1);
    kb0 = a.keys[i0];
    const unsigned kb1 = a.keys[ilast];
    const bool staged = kb1 - kb0 < (unsigned)NBX;
    for (int e = threadIdx.x; e < NN * NN; e += 256) sc[e] = a.coef[e];
    if (staged) {
        const T *src = a.Wl + (size_t)kb0 * P;
        const int cnt = (int)(kb1 - kb0 + 1) * P;
        for (int e = threadIdx.x; e < cnt; e += 256) sW[e] = src[e];
    }
    __syncthreads();
    return staged;
}

template <typename T, int D, int NN>
__global__ __launch_bounds__(256) void l2p_ct(L2PArgs<T> a)
{
    constexpr int P = ipow_c(NN, D);
    constexpr int NBX = kL2PElems / P > 0 ? kL2PElems / P : 1;
    __shared__ T sc[NN * NN];
    __shared__ T sW[NBX * P];
    const long i0 = (long)blockIdx.x * 256;
    const long i = i0 + threadIdx.x;
    unsigned kb0;
    const bool staged = l2p_stage<T, D, NN, NBX>(a, i0, sc, sW, kb0);
    if (i >= a.n) return;
    a.out[i] = l2p_point<T, D, NN>(a, i, sc, sW, kb0, staged);
}

// L2P fused with the mutual P2P form's combine pass: one thread per sorted
// point computes its L2P value as l2p_ct does and adds the point's P2P
// slots in p2p_combine_kernel's order, so ans is written once and the L2P
// values make no round trip through memory.  Bit-identical to l2p_ct
// followed by p2p_combine_kernel with base = the L2P values.
template <typename T, int D, int NN>
__global__ __launch_bounds__(256) void l2p_combine_ct(L2PArgs<T> a, P2PCombineArgs<T> c)
{
    constexpr int P = ipow_c(NN, D);
    constexpr int NBX = kL2PElems / P > 0 ? kL2PElems / P : 1;
    constexpr int NU = ((D == 1 ? 3 : D == 2 ? 9 : 27) - 1) / 2;
    __shared__ T sc[NN * NN];
    __shared__ T sW[NBX * P];
    const long i0 = (long)blockIdx.x * 256;
    const long i = i0 + threadIdx.x;
    unsigned kb0;
    const bool staged = l2p_stage<T, D, NN, NBX>(a, i0, sc, sW, kb0);
    if (i >= a.n) return;
    T v = c.slots[(long)NU * c.n + i];
#pragma unroll
    for (int u = 0; u < NU; ++u) v += c.slots[(long)u * c.n + i];
    const T base = l2p_point<T, D, NN>(a, i, sc, sW, kb0, staged);
    c.ans[c.perm[i]] = base + v;
}

#define FMMA_NN_SWITCH(NNV, CALL)                          \
    switch (NNV) {                                         \
    case 2: { constexpr int NN = 2; CALL; } return true;  \
    case 3: { constexpr int NN = 3; CALL; } return true;  \
    case 4: { constexpr int NN = 4; CALL; } return true;  \
    case 5: { constexpr int NN = 5; CALL; } return true;  \
    case 6: { constexpr int NN = 6; CALL; } return true;  \
    case 7: { constexpr int NN = 7; CALL; } return true;  \
    case 8: { constexpr int NN = 8; CALL; } return true;  \
    default: return false;                                 \
    }

template <typename T, int D>
static bool p2m_fast(const P2MArgs<T> &a, long nbox, hipStream_t st)
{
    if constexpr (D == 3) {
        switch (a.nn) {
        case 4: hipLaunchKernelGGL((p2m3_ct<T, 4>), dim3((unsigned)nbox), dim3(64), 0, st, a); return true;
        case 5: hipLaunchKernelGGL((p2m3_ct<T, 5>), dim3((unsigned)nbox), dim3(64), 0, st, a); return true;
        case 6: hipLaunchKernelGGL((p2m3_ct<T, 6>), dim3((unsigned)nbox), dim3(64), 0, st, a); return true;
        default: break;
        }
    }
    FMMA_NN_SWITCH(a.nn, hipLaunchKernelGGL((p2m_ct<T, D, NN>), dim3((unsigned)nbox), dim3(64), 0, st, a));
}

template <typename T, int D, int NN>
static void launch_m2m(const T *Wc, T *Wp, long nparent, const T *tab, hipStream_t st, const long *items)
{
    if constexpr (kWideTransfer<D, NN>)
        hipLaunchKernelGGL((m2m_w<T, D, NN>), dim3((unsigned)nparent), dim3(64 << D), 0, st, Wc, Wp, tab, items);
    else
        hipLaunchKernelGGL((m2m_ct<T, D, NN>), dim3((unsigned)nparent), dim3(64), 0, st, Wc, Wp, tab, items);
}

template <typename T, int D, int NN>
static void launch_l2l(const T *Wp, T *Wc, long nparent, const T *tab, hipStream_t st, const long *items)
{
    if constexpr (kWideTransfer<D, NN>)
        hipLaunchKernelGGL((l2l_w<T, D, NN>), dim3((unsigned)nparent), dim3(64 << D), 0, st, Wp, Wc, tab, items);
    else
        hipLaunchKernelGGL((l2l_ct<T, D, NN>), dim3((unsigned)nparent), dim3(64), 0, st, Wp, Wc, tab, items);
}

template <typename T, int D>
static bool m2m_fast(const T *Wc, T *Wp, long nparent, const T *tab, int nn, hipStream_t st, const long *items)
{
    FMMA_NN_SWITCH(nn, (launch_m2m<T, D, NN>(Wc, Wp, nparent, tab, st, items)));
}

template <typename T, int D>
static bool l2l_fast(const T *Wp, T *Wc, long nparent, const T *tab, int nn, hipStream_t st, const long *items)
{
    FMMA_NN_SWITCH(nn, (launch_l2l<T, D, NN>(Wp, Wc, nparent, tab, st, items)));
}

template <typename T, int D>
static bool l2p_fast(const L2PArgs<T> &a, hipStream_t st)
{
    const unsigned nb = (unsigned)((a.n + 255) / 256);
    FMMA_NN_SWITCH(a.nn, hipLaunchKernelGGL((l2p_ct<T, D, NN>), dim3(nb), dim3(256), 0, st, a));
}

template <typename T, int D>
static bool l2p_combine_fast(const L2PArgs<T> &a, const P2PCombineArgs<T> &c, hipStream_t st)
{
    const unsigned nb = (unsigned)((a.n + 255) / 256);
    FMMA_NN_SWITCH(a.nn, hipLaunchKernelGGL((l2p_combine_ct<T, D, NN>), dim3(nb), dim3(256), 0, st, a, c));
}

// ----------------------------------------------------------- L2P gradient
//
// d/dx_d of the L2P interpolant u(x) = sum_k Wl[b][k] prod_e S_n(r_e, c_{k_e})
// (src/fmm_core.cpp:602-612), r_e = 2((x_e - o_e)/len - b_e) - 1: the
// Chebyshev basis differentiated, S_n'(r, c_k) = sum_j coef[k][j] T_j'(r) with
// T_j' = j U_{j-1} (U from its three-term recurrence), times dr/dx = 2/len.
// One thread per sorted target; out is [D][n] in sorted order.

template <typename T, int D, int NMAX>
__global__ __launch_bounds__(256) void l2p_grad_kernel(L2PArgs<T> a)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int n = a.nn;
    const unsigned box = a.keys[i];
    const T *W = a.Wl + (size_t)box * a.P;
    const Geom<T> *gm;
    const int NL = 1 << box_geom<T>(a, box, gm);
    const T len = gm->len_leaf;
    T S[D][NMAX], dS[D][NMAX];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        T r;
        leaf_coord_rel(a.pos[d * a.n + i], gm->origin[d], len, NL, r);
        for (int k = 0; k < n; ++k) {
            S[d][k] = a.coef[k * n];
            dS[d][k] = T(0);
        }
        T tp = T(1), t = r;        // T_{j-1}, T_j
        T up = T(0), u = T(1);     // U_{j-2}, U_{j-1}
        for (int j = 1; j < n; ++j) {
            const T dt = T(j) * u;  // T_j'
            for (int k = 0; k < n; ++k) {
                S[d][k] += a.coef[k * n + j] * t;
                dS[d][k] += a.coef[k * n + j] * dt;
            }
            const T tn = T(2) * r * t - tp;
            tp = t;
            t = tn;
            const T un = T(2) * r * u - up;
            up = u;
            u = un;
        }
    }
    T g[D];
#pragma unroll
    for (int d = 0; d < D; ++d) g[d] = T(0);
    if constexpr (D == 1) {
        for (int a0 = 0; a0 < n; ++a0) g[0] += W[a0] * dS[0][a0];
    } else if constexpr (D == 2) {
        for (int a0 = 0; a0 < n; ++a0)
            for (int a1 = 0; a1 < n; ++a1) {
                const T w = W[a0 * n + a1];
                g[0] += w * (dS[0][a0] * S[1][a1]);
                g[1] += w * (S[0][a0] * dS[1][a1]);
            }
    } else {
        for (int a0 = 0; a0 < n; ++a0)
            for (int a1 = 0; a1 < n; ++a1) {
                const T s01 = S[0][a0] * S[1][a1], d01 = dS[0][a0] * S[1][a1], s0d1 = S[0][a0] * dS[1][a1];
                for (int a2 = 0; a2 < n; ++a2) {
                    const T w = W[(a0 * n + a1) * n + a2];
                    g[0] += w * (d01 * S[2][a2]);
                    g[1] += w * (s0d1 * S[2][a2]);
                    g[2] += w * (s01 * dS[2][a2]);
                }
            }
    }
    const T scale = T(2) / len;
#pragma unroll
    for (int d = 0; d < D; ++d) a.out[d * a.n + i] = g[d] * scale;
}

// ------------------------------------------------------------ P2P combine
//
// Final sums of the mutual P2P form: point i's own sum, then slot u = 0 ..
// n_upper - 1 (what box B - o_u added, or zero), in that fixed order.  One
// thread per sorted point; every read is coalesced.

template <typename T, int D>
__global__ __launch_bounds__(256) void p2p_combine_kernel(P2PCombineArgs<T> a)
{
    constexpr int NU = ((D == 1 ? 3 : D == 2 ? 9 : 27) - 1) / 2;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    T v = a.slots[(long)NU * a.n + i];
#pragma unroll
    for (int u = 0; u < NU; ++u) v += a.slots[(long)u * a.n + i];
    const unsigned oi = a.perm[i];
    if (a.base)
        a.ans[oi] = a.base[i] + v;
    else if (a.accumulate)
        a.ans[oi] += v;
    else
        a.ans[oi] = v;
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
    if (!p2m_fast<T, D>(a, nbox, st))
        FMMA_KPL_SWITCH(a.P, hipLaunchKernelGGL((p2m_kernel<T, D, KPL>), dim3((unsigned)nbox), dim3(64), 0, st, a));
    FMMA_LAUNCHED("p2m");
    return FMMA_OK;
}

template <typename T, int D>
int m2m(const T *Wc, T *Wp, long nparent, const T *tab, int nn, int P, hipStream_t st, const long *items)
{
    if (nparent <= 0) return FMMA_OK;
    if (!m2m_fast<T, D>(Wc, Wp, nparent, tab, nn, st, items))
    FMMA_KPL_SWITCH(P, hipLaunchKernelGGL((m2m_kernel<T, D, KPL>), dim3((unsigned)nparent), dim3(64), 0, st, Wc,
                                          Wp, tab, nn, P, items));
    FMMA_LAUNCHED("m2m");
    return FMMA_OK;
}

template <typename T, int D>
int l2l(const T *Wp, T *Wc, long nparent, const T *tab, int nn, int P, hipStream_t st, const long *items)
{
    if (nparent <= 0) return FMMA_OK;
    if (!l2l_fast<T, D>(Wp, Wc, nparent, tab, nn, st, items))
    FMMA_KPL_SWITCH(P, hipLaunchKernelGGL((l2l_kernel<T, D, KPL>), dim3((unsigned)nparent), dim3(64), 0, st, Wp,
                                          Wc, tab, nn, P, items));
    FMMA_LAUNCHED("l2l");
    return FMMA_OK;
}

template <typename T, int D>
int l2p(const L2PArgs<T> &a, long nbox, hipStream_t st)
{
    if (nbox <= 0 || a.n <= 0) return FMMA_OK;
    if (!l2p_fast<T, D>(a, st)) hipLaunchKernelGGL((l2p_kernel<T, D>), dim3((unsigned)nbox), dim3(64), 0, st, a);
    FMMA_LAUNCHED("l2p");
    return FMMA_OK;
}

template <typename T, int D>
int l2p_grad(const L2PArgs<T> &a, hipStream_t st)
{
    if (a.n <= 0) return FMMA_OK;
    const unsigned nb = (unsigned)((a.n + 255) / 256);
    if (a.nn <= 8)
        hipLaunchKernelGGL((l2p_grad_kernel<T, D, 8>), dim3(nb), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((l2p_grad_kernel<T, D, kMaxNodes>), dim3(nb), dim3(256), 0, st, a);
    FMMA_LAUNCHED("l2p_grad");
    return FMMA_OK;
}

// out[i] = s * out[i] (the sign of an antisymmetric kernel's transpose), or
// out[i] = 0 with s == 0.
template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(T *out, long n, T s)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = s == T(0) ? T(0) : s * out[i];
}

template <typename T>
int scale(T *out, long n, T s, hipStream_t st)
{
    if (n <= 0) return FMMA_OK;
    hipLaunchKernelGGL((scale_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, n, s);
    FMMA_LAUNCHED("scale");
    return FMMA_OK;
}
template int scale<double>(double *, long, double, hipStream_t);
template int scale<float>(float *, long, float, hipStream_t);

template <typename T, int D>
int p2p_combine(const P2PCombineArgs<T> &a, hipStream_t st)
{
    if (a.n <= 0) return FMMA_OK;
    hipLaunchKernelGGL((p2p_combine_kernel<T, D>), dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a);
    FMMA_LAUNCHED("p2p_combine");
    return FMMA_OK;
}

template <typename T, int D>
int l2p_combine(const L2PArgs<T> &a, const P2PCombineArgs<T> &c, hipStream_t st)
{
    if (a.n <= 0) return FMMA_OK;
    if (a.n != c.n || !l2p_combine_fast<T, D>(a, c, st)) return FMMA_ERR_UNSUPPORTED;
    FMMA_LAUNCHED("l2p_combine");
    return FMMA_OK;
}

#define FMMA_BOX_INST(T, D)                                                                 \
    template int l2p_grad<T, D>(const L2PArgs<T> &, hipStream_t);                           \
    template int l2p_combine<T, D>(const L2PArgs<T> &, const P2PCombineArgs<T> &, hipStream_t); \
    template int p2p_combine<T, D>(const P2PCombineArgs<T> &, hipStream_t);                 \
    template int p2m<T, D>(const P2MArgs<T> &, long, hipStream_t);                          \
    template int m2m<T, D>(const T *, T *, long, const T *, int, int, hipStream_t, const long *); \
    template int l2l<T, D>(const T *, T *, long, const T *, int, int, hipStream_t, const long *); \
    template int l2p<T, D>(const L2PArgs<T> &, long, hipStream_t);
FMMA_BOX_INST(double, 1)
FMMA_BOX_INST(double, 2)
FMMA_BOX_INST(double, 3)
FMMA_BOX_INST(float, 1)
FMMA_BOX_INST(float, 2)
FMMA_BOX_INST(float, 3)

}  // namespace lib
}  // namespace fmma
