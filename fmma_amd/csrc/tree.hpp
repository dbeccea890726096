// tree.hpp -- tree-build entry points (tree.hip).
#pragma once
#include "common.hpp"

namespace fmma {
namespace lib {

constexpr int kBBoxParts = 1024;  // partial min/max blocks of the bounding-cube reduction
template <int D>
constexpr int kBBoxStride = 2 * D + 1;  // per block: min[D], max[D], non-finite flag

// The point sort: keys[n] (in: leaf keys; out: sorted), vals[n] (out: the
// original index of each sorted point), keys_alt/vals_alt (ping-pong),
// hist[radix_hist_words(n)], the sorted coordinates (SoA, D x n) and
// weights, and SoA buffers of the same sizes for the passes in between (ws,
// ws_alt unused without weights).
template <typename T>
struct PointSortBuffers {
    unsigned *keys, *vals, *keys_alt, *vals_alt, *hist;
    T *pos, *ws, *pos_alt, *ws_alt;
};

// Arguments of one pass of the point sort (tree.hip).
template <typename T>
struct PointPass {
    const unsigned *kin, *vin;  // vin unused in the first pass (the index is the value)
    unsigned *kout, *vout;
    const T *pin, *win;  // first pass: AoS coordinates, weights (NULL: 1); then SoA
    T *pout, *wout;
    long n;
    int shift;
    const unsigned *offs;  // scattered digit starts, [256][ntiles]
    unsigned ntiles;
};

// The bounding cube of targets and sources; len_leaf = Len / n_side.
template <typename T, int D>
int bbox(const T *t, long nt, const T *s, long ns, T *part, Geom<T> *g, int n_side, hipStream_t st);

// Leaf box of each point: Morton key on the 2^leaf_level grid, or with
// grid_n > 0 the row-major index on a grid_n^D grid.  With hist0 also the
// point sort's first-pass histogram (sort_points with hist0_ready).
template <typename T, int D>
int leaf_keys(const T *pts, long n, const Geom<T> *g, int leaf_level, int grid_n, unsigned *keys, hipStream_t st,
              unsigned *hist0 = nullptr);

size_t radix_hist_words(long n);

// Stable sort of the points (pts: n x D AoS; w: weights, NULL for 1) by
// b.keys, the low `bits` bits significant: b.keys sorted, b.vals the
// original index of each sorted point, b.pos its coordinates (SoA), b.ws its
// weight when with_w.  The coordinates and weights move with the keys in
// every pass, so no gather by permutation follows.
template <typename T, int D>
int sort_points(PointSortBuffers<T> &b, long n, int bits, const T *pts, const T *w, bool with_w, hipStream_t st,
                bool hist0_ready = false);

int box_ranges(const unsigned *keys, long n, unsigned nbox, unsigned *start, hipStream_t st);

template <typename T, int D>
int gather(const T *pts, const T *w, const unsigned *perm, long n, T *pos, T *ws, hipStream_t st);

}  // namespace lib
}  // namespace fmma
