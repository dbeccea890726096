// boxops.hpp -- P2M / M2M / L2L / L2P and the P2P combine pass (boxops.hip).
#pragma once
#include "common.hpp"

namespace fmma {
namespace lib {

template <typename T>
struct P2MArgs {
    const unsigned *start;  // [nbox+1] source ranges, Morton order
    const T *pos;           // [D][n] sorted coordinates
    const T *w;             // [n] sorted weights
    long n;
    const Geom<T> *geom;
    int leaf_level;
    const T *coef;          // [nn][nn] Chebyshev weight coefficients
    int nn, P;
    T *Wm;                  // [nbox][P] leaf multipoles (written for every box)
    // a batch of problems (else NULL): problem of each leaf box, per-problem geometry and leaf level
    const unsigned *box_prob;
    const Geom<T> *geoms;
    const int *levels;
    int grid_n;  // > 0: the boxes are a row-major grid_n^D grid (nrnmm), leaf_level unused
};

template <typename T>
struct L2PArgs {
    const unsigned *start;  // [nbox+1] target ranges
    const unsigned *keys;   // [n] leaf box (Morton) of each sorted target
    const T *pos;           // [D][n] sorted target coordinates
    long n;
    const Geom<T> *geom;
    int leaf_level;
    const T *coef;
    int nn, P;
    const T *Wl;            // [nbox][P] leaf locals
    T *out;                 // [n] sorted order
    const unsigned *box_prob;  // batch (else NULL), as in P2MArgs
    const Geom<T> *geoms;
    const int *levels;
};

// Final sums of the mutual P2P form (fmma_p2p_args::slots).
template <typename T>
struct P2PCombineArgs {
    const T *slots;         // [(3^D-1)/2 + 1][n], every entry written by the P2P pass
    long n;
    const T *base;          // [n] sorted order (L2P result), or NULL
    const unsigned *perm;   // [n] original index of each sorted point
    T *ans;                 // [n] original order
    int accumulate;         // base == NULL: 1 adds onto ans, 0 overwrites
};

template <typename T, int D>
int p2m(const P2MArgs<T> &a, long nbox, hipStream_t st);

// Wp[p] = M2M of children Wc[8p .. 8p+7] for p < nparent (overwrites Wp).
// With items (a batch): parent q reads the 2^D child rows at Wc + items[2q]
// and writes Wp + items[2q+1] (element offsets).
template <typename T, int D>
int m2m(const T *Wc, T *Wp, long nparent, const T *tab, int nn, int P, hipStream_t st,
        const long *items = nullptr);

// Wc[child] += L2L of Wp[parent] for every child of every parent; items as in m2m
// (items[2q] the children, items[2q+1] the parent).
template <typename T, int D>
int l2l(const T *Wp, T *Wc, long nparent, const T *tab, int nn, int P, hipStream_t st,
        const long *items = nullptr);

template <typename T, int D>
int l2p(const L2PArgs<T> &a, long nbox, hipStream_t st);

template <typename T, int D>
int p2p_combine(const P2PCombineArgs<T> &a, hipStream_t st);

// l2p and p2p_combine (with base = the L2P values) in one pass over the
// sorted points, bit-identical to the two; FMMA_ERR_UNSUPPORTED (nothing
// launched) for node counts without the compile-time kernels.
template <typename T, int D>
int l2p_combine(const L2PArgs<T> &a, const P2PCombineArgs<T> &c, hipStream_t st);

// Gradient of the L2P interpolant with respect to the target coordinates:
// a.out is [D][n], sorted order.
template <typename T, int D>
int l2p_grad(const L2PArgs<T> &a, hipStream_t st);

// out[i] *= s for i < n, or out[i] = 0 when s == 0.
template <typename T>
int scale(T *out, long n, T s, hipStream_t st);

}  // namespace lib
}  // namespace fmma
