// m2l.hpp -- M2L as batched MFMA contractions (m2l.hip).
#pragma once
#include <vector>

#include "common.hpp"

namespace fmma {
namespace lib {

constexpr int kMaxLevels = 16;
constexpr int kTileParents = 64;  // target boxes per workgroup (one child class of 64 parents)

// Separation vectors.  "far" = every o in [-3,3]^D with max|o_d| >= 2, in
// row-major order (the reference's 7^D operator table minus the near block,
// src/fmm_core.cpp:391-402).  A target box's child class sigma (its
// coordinate parities) fixes which far offsets are interactions: o_d in
// [-2,3] when b_d is even, [-3,2] when odd -- the children of the parent's
// neighbours (src/fmm_core.cpp:21-37) -- 6^D - 3^D of them for every class.
struct OffsetTables {
    int D = 0, n_far = 0, n_cls = 0;
    std::vector<int> far;    // [n_far][D]
    std::vector<int> cls_f;  // [2^D][n_cls] index into far
    std::vector<int> cls_o;  // [2^D][n_cls][D]
    void build(int dim);
};

// A workgroup's tile: the sigma = cls children of parents p0 .. p0+np-1 of a
// level, with the element offsets of that level's multipoles / locals
// (w_off, rows of P), operand copy (c_off, rows of KP) and operator
// fragments (k_off) -- per problem in a batch.
struct M2LTask {
    int level, cls;
    unsigned p0;
    int np;
    long w_off, c_off, k_off;
    // A split level (few workgroups, each a long chain of offset steps): the
    // task takes the class's offsets [i0, i1) and writes its partial sums at
    // element o_off of M2LKArgs::Wpart.  i1 == 0: every offset, into Wl.
    int i0 = 0, i1 = 0;
    long o_off = 0;
};

// A split level's partial sums: part p of the level's count = boxes x P
// elements sits at part_off + p * count of Wpart; m2l_sum_parts adds the
// parts in order into Wl at w_off.
struct M2LSplit {
    int level, parts;
    long w_off, count, part_off;
};

// Workgroups a level's launch should reach before its offsets are split:
// eight per CU.
constexpr long kM2LSplitTarget = 2048;

// A run of tasks with one tile width (one launch).
struct M2LGroup {
    int tb, first, count;
    bool pipelined;
};

// Tasks of levels leaf_level .. 2 of one problem (or of only_level >= 0);
// the offsets are filled from per-level tables (element offsets) given by
// the caller.  The tile width of a level is tb_of_level[l] when given and
// positive, else the widest of 256/128/64 that keeps >= 512 workgroups.
// With splits, a pipelined level whose launch would have at most half of
// split_target workgroups is cut into 2..8 parts of at least 16 offsets each
// (as many as keep the workgroups <= split_target); *splits lists them and
// the elements of Wpart they take (return value of m2l_split_elems).
std::vector<M2LTask> make_m2l_tasks(int D, int leaf_level, int P, int mode, std::vector<M2LGroup> &groups,
                                    const long *w_off, const long *c_off, const long *k_off,
                                    const int *tb_of_level = nullptr, int only_level = -1,
                                    std::vector<M2LSplit> *splits = nullptr, long split_target = 0);
long m2l_split_elems(const std::vector<M2LSplit> &splits);
bool m2l_pipe_supported(int D, int KP, int mode, int tb);

// Kernel arguments of the MFMA contraction.
struct M2LKArgs {
    const M2LTask *tasks;
    const int *cls_f;
    const int *cls_o;
    int n_cls;
    int P, MP, KP;              // MP = 16*ceil(P/16) rows, KP = padded contraction length
    const void *Wm;
    void *Wl;
    const void *Kf;             // operator fragments (f64 / f32), or bf16 hi part
    const void *Kf_lo;          // bf16 lo part
    const void *Wc;             // multipoles as MFMA operands, rows of KP: TYPE / float / bf16 hi
    const void *Wc_lo;          // bf16 lo part (mode 2)
    void *Wpart;                // partial sums of split levels (TYPE)
};

// Shape of the operator fragments for a precision: rows MP, contraction KP.
void m2l_shape(int P, int precision, int &MP, int &KP);
// The FP32 form at a padded length of 128 runs on the 32x32x2 MFMA
// (fragments FMMA_FRAG_F32W, MP = 128); the fragment format of a form.
bool m2l_wide(int mode, int KP);
int m2l_frag_format(int P, int mode);
size_t m2l_frag_elems(int n_far, int P, int precision);  // elements per level (per part for bf16)

// Kraw [n_far][P][P] (TYPE) -> fragments of level l.
template <typename T>
int m2l_fragments(const T *Kraw, int n_far, int P, int precision, void *Kf, void *Kf_lo, hipStream_t st);

// nbox multipole rows of P converted to padded operand rows of KP.
template <typename T>
int m2l_convert(const T *W, long nbox, int P, int KP, int mode, void *Wc, void *Wc_lo, hipStream_t st);

// One launch per task group: workgroup per task, blockDim 256, grid.y = row chunks.
template <typename T, int D>
int m2l_apply(const M2LKArgs &a, const std::vector<M2LGroup> &groups, int mode, hipStream_t st);

// Wl of each split level = the sum of its parts, in part order.
template <typename T>
int m2l_sum_parts(T *Wl, const T *Wpart, const std::vector<M2LSplit> &splits, hipStream_t st);

}  // namespace lib
}  // namespace fmma
