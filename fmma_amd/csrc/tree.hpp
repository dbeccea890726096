// tree.hpp -- tree-build entry points (tree.hip).
#pragma once
#include "common.hpp"

namespace fmma {
namespace lib {

constexpr int kBBoxParts = 1024;  // partial min/max blocks of the bounding-cube reduction
template <int D>
constexpr int kBBoxStride = 2 * D + 1;  // per block: min[D], max[D], non-finite flag

struct SortBuffers {
    unsigned *keys;      // [n] in: leaf keys; out: sorted keys
    unsigned *vals;      // [n] out: original index of each sorted point
    unsigned *keys_alt;  // [n] ping-pong
    unsigned *vals_alt;  // [n]
    unsigned *hist;      // [radix_hist_words(n)]
};

// The bounding cube of targets and sources; len_leaf = Len / n_side.
template <typename T, int D>
int bbox(const T *t, long nt, const T *s, long ns, T *part, Geom<T> *g, int n_side, hipStream_t st);

// Values per packed point record: D coordinates and the weight, padded to
// 16 bytes (D = 1) or 32 bytes (D = 2, 3) for double.
template <int D>
constexpr int kRecStride = D == 1 ? 2 : 4;

// Leaf box of each point: Morton key on the 2^leaf_level grid, or with
// grid_n > 0 the row-major index on a grid_n^D grid.  With rec != NULL also
// each point's packed record (coordinates, then w[i], or 1 without w).
template <typename T, int D>
int leaf_keys(const T *pts, long n, const Geom<T> *g, int leaf_level, int grid_n, unsigned *keys, hipStream_t st,
              const T *w = nullptr, T *rec = nullptr);

// Stable sort of b.keys (low `bits` bits significant); b.vals gets the
// permutation.  bits == 0 only fills b.vals with 0..n-1.
int radix_sort(SortBuffers &b, long n, int bits, hipStream_t st);
size_t radix_hist_words(long n);

int box_ranges(const unsigned *keys, long n, unsigned nbox, unsigned *start, hipStream_t st);

template <typename T, int D>
int gather(const T *pts, const T *w, const unsigned *perm, long n, T *pos, T *ws, hipStream_t st);

// pos[d*n + i] = rec[perm[i]] coordinates, ws[i] (if not NULL) its weight.
template <typename T, int D>
int gather_records(const T *rec, const unsigned *perm, long n, T *pos, T *ws, hipStream_t st);

}  // namespace lib
}  // namespace fmma
