// batch.hpp -- tree-build kernels for a batch of independent problems (batch.hip).
#pragma once
#include "common.hpp"

namespace fmma {
namespace lib {

// Geometry of every problem p: the bounding cube of its targets
// t[t_off[p] .. t_off[p+1]) and sources s[s_off[p] .. s_off[p+1]) (s may be
// NULL when the targets are the sources), with the same arithmetic as the
// single-problem bbox, and len_leaf for its leaf level levels[p].
template <typename T, int D>
int batch_bbox(const T *t, const long *t_off, const T *s, const long *s_off, const int *levels, int nprob,
               Geom<T> *geoms, hipStream_t st);

// Leaf key of every point i of the concatenated set: box_base[p] + the Morton
// code of its leaf box in problem p = the problem with off[p] <= i < off[p+1].
template <typename T, int D>
int batch_keys(const T *pts, long n, const long *off, int nprob, const Geom<T> *geoms, const int *levels,
               const unsigned *box_base, unsigned *keys, hipStream_t st);

}  // namespace lib
}  // namespace fmma
