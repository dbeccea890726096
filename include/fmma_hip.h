/*
 * fmma_hip.h -- C-ABI of the MI355X black-box FMM library (libfmma_hip.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  Each
 * entry point names the reference (fockl/FMMA) interface it replaces.  The
 * reference has no C-ABI of its own; its surface is the C++ class
 * fmma::FMMA<TYPE,DIM> (include/fmma/fmma.hpp:52-170) and the pybind11 module
 * pyfmma (cpython/pyfmma.cpp:140-168).  include/fmma/fmma.hpp in this tree
 * rebuilds that class on top of these functions.
 *
 * Threading: a plan is not re-entrant (as the reference object is not:
 * src/fmm.cpp:23-28 mutates Depth); use one plan per thread.  All work is
 * stream-ordered on the plan's stream; the solve entry points return after
 * that stream has finished the solve.
 */
#ifndef FMMA_HIP_H
#define FMMA_HIP_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- status codes */
enum {
    FMMA_OK = 0,
    FMMA_ERR_INVALID = 1,     /* bad argument (dimension, order, null pointer, ...) */
    FMMA_ERR_SIZE = 2,        /* source_weight size != source size (src/fmm.cpp:32-35) */
    FMMA_ERR_SOLVER = 3,      /* unknown solver type (src/fmma.cpp:183-186) */
    FMMA_ERR_HIP = 4,         /* a HIP runtime call failed; fmma_last_error() has the text */
    FMMA_ERR_UNSUPPORTED = 5  /* valid for the reference, not provided on the GPU */
};

/* ------------------------------------------------------------- data types */
enum { FMMA_F64 = 0, FMMA_F32 = 1 };

/* Built-in kernels f(x,y).  The reference takes any std::function
 * (include/fmma/fmma.hpp:55-87); a C caller names one of these instead, and a
 * C++ caller compiled with hipcc can pass its own functor (fmma_kernel_ops). */
enum {
    FMMA_KERNEL_INV_R = 0,     /* 1/|x-y|, inf at r=0: the reference default (fmma.hpp:55-63) */
    FMMA_KERNEL_INV_R_REG = 1, /* 1/|x-y|, 0 at r=0 */
    FMMA_KERNEL_INV_DIFF = 2,  /* 1/(x_0-y_0) */
    FMMA_KERNEL_LOG_R = 3,     /* log|x-y|, 0 at r=0 */
    FMMA_KERNEL_GAUSS = 4,     /* exp(-|x-y|^2/(2h^2)), params[0]=h */
    FMMA_KERNEL_YUKAWA = 5,    /* exp(-kappa|x-y|)/|x-y|, 0 at r=0, params[0]=kappa */
    FMMA_KERNEL_SUM_R2 = 6,    /* sum_d (x_d-y_d)/|x-y|^2 (test/src/fmma/fmm.cpp:82-91) */
    FMMA_KERNEL_SUM2_R2 = 7,   /* sum_d (x_d+y_d)^2/|x-y|^2, not translation invariant
                                  (test/src/fmma/fmm.cpp:151-161) */
    FMMA_KERNEL_CUSTOM = 100   /* set by fmma_plan_set_kernel_ops */
};

/* Arithmetic of the M2L contraction.  The reference computes it in TYPE
 * (src/fmm_core.cpp:451 -> src/math.cpp:263-294); the other forms are new. */
enum {
    FMMA_M2L_NATIVE = 0, /* TYPE: FP64 MFMA for double, FP32 MFMA for float */
    FMMA_M2L_FP32 = 1,   /* FP32 MFMA per offset, summed over offsets in TYPE */
    FMMA_M2L_BF16X3 = 2  /* split-BF16 (hi*hi + hi*lo + lo*hi) MFMA, FP32 per offset */
};

typedef struct fmma_plan_s *fmma_plan;

/* ---------------------------------------------- kernel-dependent launchers
 * The stages whose arithmetic depends on f: P2P, the exact solver and the
 * M2L operator tables.  The library fills this table for the built-in
 * kernels; include/fmma/fmma_device.hpp fills it for a user functor compiled
 * in the caller's hipcc translation unit.  All pointers are device pointers;
 * coordinates are structure-of-arrays (dimension-major). */
typedef struct {
    int dim, dtype;
    long n_target_boxes;       /* leaf boxes, Morton order */
    int leaf_level;            /* boxes per side = 1 << leaf_level */
    const unsigned *t_start;   /* [n_boxes+1] target ranges in sorted order */
    const unsigned *s_start;   /* [n_boxes+1] source ranges in sorted order */
    const void *t_pos;         /* [dim][n_t] sorted target coordinates */
    const void *s_pos;         /* [dim][n_s] sorted source coordinates */
    const void *s_w;           /* [n_s] sorted weights */
    long n_t, n_s;
    const void *base;          /* [n_t] value to add to, sorted order (L2P result), or NULL */
    const unsigned *t_perm;    /* [n_t] original index of each sorted target */
    void *ans;                 /* [n_t] caller's output, original order */
    int accumulate;            /* base==NULL: 1 adds onto ans (leaf level < 2), 0 overwrites */
    /* Mutual form, used when slots != NULL: targets and sources are the same
     * sorted points (t_pos == s_pos, t_start == s_start) and f(y,x) =
     * symmetry * f(x,y).  Box A evaluates itself and each neighbour B that
     * follows it in the row-major 3^dim block once, adding to A's sums and,
     * times symmetry, to B's.  The launcher writes every slots[u][i] (u <
     * n_upper = (3^dim-1)/2: what point i of box B received from box B - o_u,
     * o_u the u-th such offset, or 0 when that box is outside the grid or
     * empty) and slots[n_upper][i] (point i's own sum), sorted order, and
     * leaves ans alone; the library's combine pass adds them in a fixed order. */
    void *slots;               /* [n_upper+1][n_t] TYPE, or NULL for the direct form */
    int symmetry;              /* +1 or -1 in the mutual form */
    /* A batch of independent problems (else NULL): leaf box b belongs to
     * problem box_problem[b], whose boxes problem_box_base[p] ..
     * problem_box_base[p+1]-1 are its Morton-ordered grid of side
     * 1 << problem_leaf_level[p]; leaf_level is then unused. */
    const unsigned *box_problem;
    const unsigned *problem_box_base;
    const int *problem_leaf_level;
    /* > 0: the leaf boxes are a row-major grid of grid_n per side (the nrnmm
     * solver) instead of a Morton-ordered 2^leaf_level tree. */
    int grid_n;
    /* Balanced direct form (slots == NULL), for boxes whose near field is
     * far above the average (clustered points).  split_parts[box] > 0 marks a
     * box the library split into work items; a launch without split_items
     * skips those boxes.  A launch with split_items runs one workgroup per
     * item split_items[3q .. 3q+2] = (box, pass, chunk): the box's targets
     * t_start[box] + pass * split_pass .. (at most split_pass of them)
     * against chunk `chunk` of the concatenated neighbourhood records (the
     * box's own sources first, then the other neighbour boxes in row-major
     * order) cut into split_parts[box] chunks of equal size.  With more than
     * one chunk the item writes its sums to split_part[chunk * n_t + i]
     * (sorted order) and leaves ans alone; the library adds the chunks of a
     * target in chunk order. */
    const unsigned *split_parts;  /* [n_boxes], or NULL */
    const unsigned *split_items;  /* [n_split_items][3], or NULL */
    long n_split_items;
    unsigned split_pass;
    void *split_part;             /* [max parts][n_t] TYPE */
} fmma_p2p_args;

typedef struct {
    int dim, dtype;
    long n_t, n_s;
    const void *targets;       /* [n_t][dim] caller layout (array of structures) */
    const void *weights;       /* [n_s] or NULL for all-ones */
    const void *sources;       /* [n_s][dim] */
    void *ans;                 /* [n_t] */
    void *partial;             /* [split][n_t] scratch of the library, or NULL when split == 1 */
    int split;                 /* number of source slices summed in fixed order */
} fmma_exact_args;

/* Near field of the gradient with respect to the target coordinates:
 * grad[i][d] = base[d][i] + sum_j w_j d/dx_d f(x_i, y_j) over the sources of
 * the 3^dim neighbour boxes, for the fmm solver's gradient.  Same tree
 * arrays as fmma_p2p_args. */
typedef struct {
    int dim, dtype;
    long n_target_boxes;
    int leaf_level;
    const unsigned *t_start;   /* [n_boxes+1] */
    const unsigned *s_start;   /* [n_boxes+1] */
    const void *t_pos;         /* [dim][n_t] sorted */
    const void *s_pos;         /* [dim][n_s] sorted */
    const void *s_w;           /* [n_s] sorted */
    long n_t, n_s;
    const void *base;          /* [dim][n_t] far-field gradient, sorted order, or NULL */
    const unsigned *t_perm;    /* [n_t] */
    void *grad;                /* [n_t][dim] caller's output, original order, overwritten */
    const unsigned *box_problem;  /* a batch, as in fmma_p2p_args (else NULL) */
    const unsigned *problem_box_base;
    const int *problem_leaf_level;
    int grid_n;                   /* as in fmma_p2p_args */
    /* The balanced form as in fmma_p2p_args; chunk sums at
     * split_part[(chunk * dim + d) * n_t + i]. */
    const unsigned *split_parts;
    const unsigned *split_items;
    long n_split_items;
    unsigned split_pass;
    void *split_part;             /* [max parts][dim][n_t] TYPE */
} fmma_p2p_grad_args;

/* Far field of the tree and nrnmm solvers (M2P, src/fmm_core.cpp:625-661 and
 * src/nrnmm.cpp:79-91): out[i] = sum over the far boxes s of target i and
 * their nodes k of f(x_i, y_{s,k}) Wm[s][k], node position
 * y_{s,k,d} = (c_{k_d} + 1)/2 len + (len s_d + origin_d), len = Len / (boxes
 * per side of s's level), k = sum_d k_d n^(dim-1-d).
 *  grid_n == 0 (tree): the target boxes are the Morton-ordered 2^leaf_level
 *  grid; target i's far boxes are the interaction lists of its box at every
 *  level l = 2..leaf_level (the children of the parent's neighbours that are
 *  not neighbours, src/fmm_core.cpp:15-63), whose multipoles sit in Morton
 *  order at Wm + P (2^(dim l) - 1)/(2^dim - 1).
 *  grid_n > 0 (nrnmm): the target boxes are the row-major grid_n^dim grid and
 *  the far boxes every box of it at Chebyshev distance >= 2, multipoles
 *  [grid_n^dim][P] row-major. */
typedef struct {
    int dim, dtype;
    int n_nodes, P;
    const double *nodes;       /* [n_nodes] Chebyshev nodes, device */
    long n_target_boxes;
    const unsigned *t_start;   /* [n_target_boxes+1] target ranges in sorted order */
    const void *t_pos;         /* [dim][n_t] sorted target coordinates */
    long n_t;
    const void *origin;        /* [dim] TYPE, device: the bounding cube's corner */
    const void *Len;           /* TYPE, device: its side */
    int leaf_level, grid_n;
    const void *Wm;
    void *out;                 /* [n_t] sorted order, overwritten */
} fmma_m2p_args;

/* All-pairs gradient: grad[i][d] = sum_j w_j d/dx_d f(x_i, y_j). */
typedef struct {
    int dim, dtype;
    long n_t, n_s;
    const void *targets;       /* [n_t][dim] */
    const void *weights;       /* [n_s] or NULL for all-ones */
    const void *sources;       /* [n_s][dim] */
    void *grad;                /* [n_t][dim] */
} fmma_exact_grad_args;

/* Element order of M2L operator fragments (include/fmma/detail/m2l_layout.hpp):
 * 16x16x4 MFMA operands (F64, F32), 16x16x32 (BF16), 32x32x2 FP32 operands
 * in groups of four steps (F32W) or 32x32x16 BF16 operands (BF16W), the
 * last two for MP = KP = 128. */
enum {
    FMMA_FRAG_RAW = 0,
    FMMA_FRAG_F64 = 1,
    FMMA_FRAG_F32 = 2,
    FMMA_FRAG_BF16 = 3,
    FMMA_FRAG_F32W = 4,
    FMMA_FRAG_BF16W = 5
};

/* The bounding cube lives on the device (it is reduced there); the box side
 * at a level is len = Len / (TYPE)(1 << level), as in src/fmm_core.cpp:376.
 * format FMMA_FRAG_RAW: one level into K.  Otherwise n_batch operator sets,
 * set b at level batch_level[b] with cube side Len[batch_problem[b] *
 * len_stride], written straight into fragment order (MP rows, KP columns per
 * offset) at Kf (+ Kf_lo for the BF16 lo part), set b at b * n_off * MP * KP. */
typedef struct {
    int dim, dtype;
    int n_nodes;               /* nodes per dimension n = poly_ord + 1 */
    int P;                     /* n^dim */
    const double *nodes;       /* [n] Chebyshev nodes (device) */
    int level;                 /* boxes per side 1 << level (FMMA_FRAG_RAW) */
    const void *Len;           /* TYPE, device: side of the bounding cube */
    const int *offsets;        /* [n_off][dim] separation vectors o (device) */
    int n_off;
    void *K;                   /* [n_off][P][P] K[o][kl][km] = f(len*c_kl/2, len*(o + c_km/2)), TYPE */
    int format, MP, KP, n_batch, len_stride;
    const int *batch_level;    /* [n_batch] device */
    const int *batch_problem;  /* [n_batch] device */
    void *Kf, *Kf_lo;
    /* Optional, fragment formats of a kernel with symmetry +-1: n_pairs
     * (offset, -offset) index pairs; only the first of each pair is
     * evaluated, K_{-o}[km][kl] = symmetry * K_o[kl][km] is written from it. */
    const int *pairs;          /* [n_pairs][2] device, or NULL */
    int n_pairs;
} fmma_m2l_op_args;

typedef struct {
    int dim, dtype;
    int n_nodes, P;
    const double *nodes;
    int level;                 /* boxes per side 1 << level */
    const void *Len;           /* TYPE, device */
    const void *origin;        /* [dim] TYPE, device */
    const void *Wm;            /* [boxes][P] multipoles, Morton order, TYPE */
    void *Wl;                  /* [boxes][P] locals, overwritten */
} fmma_m2l_general_args;

/* Each launcher gets the table's `ctx` back (the functor's state, or the
 * built-in kernel's parameters) and the plan's hipStream_t; it returns an
 * FMMA_* status and must not synchronise. */
typedef struct fmma_kernel_ops_s {
    int dim, dtype;
    int translation_invariant; /* 1: f(x,y) = g(x-y); M2L may use per-offset operators */
    int symmetry;              /* +1: f(y,x) = f(x,y), -1: f(y,x) = -f(x,y), 0: neither or unknown;
                                  nonzero lets P2P use the mutual form when targets are the sources */
    void *ctx;
    int (*p2p)(const fmma_p2p_args *, void *ctx, void *stream);
    int (*exact)(const fmma_exact_args *, void *ctx, void *stream);
    int (*m2l_op)(const fmma_m2l_op_args *, void *ctx, void *stream);
    int (*m2l_general)(const fmma_m2l_general_args *, void *ctx, void *stream);
    /* Optional (NULL when f has no gradient): grad_x f for fmma_plan_gradient. */
    int (*p2p_grad)(const fmma_p2p_grad_args *, void *ctx, void *stream);
    int (*exact_grad)(const fmma_exact_grad_args *, void *ctx, void *stream);
    /* The table of f^T(x,y) = f(y,x), for fmma_plan_solve_transpose when
     * symmetry == 0 (NULL otherwise: f^T is then symmetry * f). */
    const struct fmma_kernel_ops_s *transposed;
    /* Optional (NULL: the tree and nrnmm solvers are unavailable): M2P. */
    int (*m2p)(const fmma_m2p_args *, void *ctx, void *stream);
} fmma_kernel_ops;

/* ------------------------------------------------------------ plan life */

/* Replaces constructing fmma::FMMA<TYPE,DIM> (include/fmma/fmma.hpp:130,
 * src/fmma.cpp:14-17).  dim in {1,2,3}; dtype FMMA_F64 or FMMA_F32.  The new
 * plan has the reference's defaults: solver "exact", poly_ord 1, Depth -1
 * (auto), trans_sym_flag 0, kernel FMMA_KERNEL_INV_R. */
int fmma_plan_create(fmma_plan *plan, int dim, int dtype);

/* Replaces ~FMMA (src/fmma.cpp:26-29).  Frees device memory. */
int fmma_plan_destroy(fmma_plan plan);

/* FMMA::set_solver_type (include/fmma/fmma.hpp:92-95).  "exact",
 * "exact_matvec", "fmm", "tree" and "nrnmm" run on the GPU (the last two need
 * a launcher table with m2p, which every built-in and set_fn table has);
 * anything else makes solve return FMMA_ERR_SOLVER, the reference's exit path
 * (src/fmma.cpp:183-186). */
int fmma_plan_set_solver_type(fmma_plan plan, const char *type);

/* FMMA::set_poly_ord (include/fmma/fmma.hpp:108-111): n = poly_ord+1 nodes per
 * dimension.  0 <= poly_ord <= 15. */
int fmma_plan_set_poly_ord(fmma_plan plan, int poly_ord);

/* FMMA::set_Depth (include/fmma/fmma.hpp:116-119).  <= 0 selects the auto
 * rule (int)(ln N_targets / DIM) + 1 at the next fmm or tree solve, which
 * then sticks (src/fmm.cpp:23-28, src/tree.cpp:20-25). */
int fmma_plan_set_depth(fmma_plan plan, int depth);
int fmma_plan_get_depth(fmma_plan plan, int *depth);

/* FMMA::set_trans_sym_flag (include/fmma/fmma.hpp:124-127). */
int fmma_plan_set_trans_sym_flag(fmma_plan plan, int flag);

/* FMMA::set_nrn_N (include/fmma/fmma.hpp:100-103): the nrnmm solver's boxes
 * per side.  <= 0 selects the auto rule (int)(sqrt(3) M^(1/(2 DIM))) + 1, M
 * the source count, at the next nrnmm solve, which then sticks
 * (src/nrnmm.cpp:20-22).  The getter has no counterpart in the reference. */
int fmma_plan_set_nrn_N(fmma_plan plan, int nrn_N);
int fmma_plan_get_nrn_N(fmma_plan plan, int *nrn_N);

/* Stands in for FMMA::set_fn / set_fn1 / set_fn2 (include/fmma/fmma.hpp:65-87)
 * with a built-in kernel.  params may be NULL when nparams == 0. */
int fmma_plan_set_kernel(fmma_plan plan, int kernel_id, const double *params, int nparams);

/* Stands in for FMMA::set_fn with a user functor: the table is copied. */
int fmma_plan_set_kernel_ops(fmma_plan plan, const fmma_kernel_ops *ops);

/* New: M2L arithmetic, FMMA_M2L_*. */
int fmma_plan_set_m2l_precision(fmma_plan plan, int precision);

/* New: the HIP stream the plan runs on (hipStream_t; NULL = null stream). */
int fmma_plan_set_stream(fmma_plan plan, void *stream);

/* New: record per-stage HIP-event times into the time log (default on). */
int fmma_plan_set_timing(fmma_plan plan, int enable);

/* New: allow the mutual P2P form (default 1).  It is used only when targets
 * and sources are the same array and the kernel declares a symmetry; it
 * evaluates each pair of distinct neighbour boxes once instead of twice and
 * needs (3^dim+1)/2 * n TYPE values of scratch.  The reference evaluates every
 * target box against its whole neighbourhood (src/fmm.cpp:122-141). */
int fmma_plan_set_p2p_mutual(fmma_plan plan, int enable);

/* New: balance the direct P2P and its gradient over work items when some
 * target boxes have far more near-field pairs than the average (default 1;
 * clustered points): such a box's targets are cut into passes and its
 * neighbourhood sources into chunks whose sums are added in a fixed order.
 * Costs one small device-to-host read per direct P2P.  The reference runs
 * one loop per target box (src/fmm.cpp:122-141). */
int fmma_plan_set_p2p_balance(fmma_plan plan, int enable);

/* ------------------------------------------------------------------ solve */

/* FMMA::solve(target, source_weight, source, ans) (src/fmma.cpp:169-207),
 * and with weights == NULL FMMA::solve(target, source, ans)
 * (src/fmma.cpp:143-160).  DEVICE pointers: targets [n_t][dim], weights [n_s],
 * sources [n_s][dim], ans [n_t], all of the plan's dtype.  targets may equal
 * sources (same pointer and count): the tree is then built once.  As in the
 * reference, ans is overwritten when the fmm leaf level is >= 2 and added to
 * otherwise, and left untouched when n_t or n_s is 0. */
int fmma_plan_solve(fmma_plan plan, long n_t, const void *targets, const void *weights,
                    long n_s, const void *sources, void *ans);

/* No counterpart in the reference: fmma_plan_solve for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  The same solve, arguments and results (to rounding), but every
 * device buffer it needs (tree, multipoles, operators, near-field slots) is
 * carved from one static 64 MiB scratch array per GPU that the library holds
 * until the process ends; it calls no hipMalloc or hipFree and creates no
 * events, so it records no stage times, and it uses the rule's M2L tile
 * widths (or fmma_plan_set_m2l_tile's) instead of the autotuner.  A solve
 * that needs more than the scratch (3-D FP64 at 4 nodes per dimension:
 * beyond about 150,000 points) fails with FMMA_ERR_SIZE before ans is
 * written.  Calls that overlap on one GPU share its scratch and race; use one
 * stream (or serialise the calls).  The auto Depth / nrn_N it picks is kept
 * by the plan, as with fmma_plan_solve. */
int fmma_plan_solve_noalloc(fmma_plan plan, long n_t, const void *targets, const void *weights,
                            long n_s, const void *sources, void *ans);

/* The same with HOST pointers: the library stages them through device
 * memory (the copies are outside the per-stage time log). */
int fmma_plan_solve_host(fmma_plan plan, long n_t, const void *targets, const void *weights,
                         long n_s, const void *sources, void *ans);

/* New (differentiation): the gradient of fmma_plan_solve's result with
 * respect to the target coordinates, grad[i][d] = d ans_i / d x_{i,d}, for
 * the solvers "fmm" (the derivative of the fmm approximation: the L2P
 * interpolant differentiated through the Chebyshev basis of
 * src/fmm_core.cpp:602-612, plus sum_j w_j grad_x f over the neighbour boxes)
 * and "exact" / "exact_matvec" (all pairs).  DEVICE pointers; grad is
 * [n_t][dim], overwritten.  reuse = 1 takes the local expansions of this
 * plan's previous fmm solve when it had the same pointers, counts and
 * settings; otherwise, and with reuse = 0, the fmm passes run again.  Needs a
 * kernel with a gradient (every built-in, or fmma_kernel_ops.p2p_grad). */
int fmma_plan_gradient(fmma_plan plan, long n_t, const void *targets, const void *weights,
                       long n_s, const void *sources, void *grad, int reuse);
int fmma_plan_gradient_host(fmma_plan plan, long n_t, const void *targets, const void *weights,
                            long n_s, const void *sources, void *grad);

/* New (differentiation): the transpose in the weights, out[j] = sum_i cot[i]
 * f(x_i, y_j) for j < n_s, i.e. d(sum_i cot_i ans_i)/d w_j.  For "fmm" it is
 * the fmm solver run with targets and sources swapped and f transposed, on
 * the tree and Depth that fmma_plan_solve(targets, ., sources) uses (Depth
 * from n_t when still automatic), so it is the exact adjoint of that solve's
 * operator up to rounding.  DEVICE pointers: targets [n_t][dim], cot [n_t],
 * sources [n_s][dim], out [n_s], overwritten. */
int fmma_plan_solve_transpose(fmma_plan plan, long n_t, const void *targets, const void *cot,
                              long n_s, const void *sources, void *out);
int fmma_plan_solve_transpose_host(fmma_plan plan, long n_t, const void *targets, const void *cot,
                                   long n_s, const void *sources, void *out);

/* New (batching): nprob independent problems with this plan's settings in
 * one pass of batched kernels.  Problem p has t_counts[p] targets and
 * s_counts[p] sources, stored one after another in the concatenated device
 * arrays targets [sum t_counts][dim], weights [sum s_counts] (NULL for
 * all-ones), sources [sum s_counts][dim]; ans [sum t_counts] is overwritten.
 * Each problem is solved as fmma_plan_solve would solve it alone: its own
 * bounding cube, and its own Depth (the plan's when set, else the auto rule
 * on t_counts[p]; each must be >= 3).  targets == sources with equal counts
 * means every problem's targets are its sources.  Solver "fmm" with
 * per-offset M2L operators (trans_sym_flag 1 or a translation-invariant
 * kernel).  The first call for a layout also runs the M2L tile autotuner
 * (fmma_plan_set_m2l_autotune); later calls replay one captured hipGraph
 * (fmma_plan_set_batch_graph) while the layout, the settings and every
 * address stay the same.  No reference counterpart: the reference solves one
 * problem per call (src/fmm.cpp:15-144). */
int fmma_plan_solve_batch(fmma_plan plan, int n_problems, const long *t_counts, const void *targets,
                          const void *weights, const long *s_counts, const void *sources, void *ans);
int fmma_plan_set_batch_graph(fmma_plan plan, int enable);   /* default 1 */
int fmma_plan_set_m2l_autotune(fmma_plan plan, int enable);  /* default 1 */
/* 0 (default): the autotuner's (or, with it off, the rule's) M2L tile widths;
 * 16, 32, 64, 128 or 256: that many target boxes per workgroup at every level
 * where a pipelined kernel has the width (the rule's elsewhere), for tests and
 * measurements.  No reference counterpart. */
int fmma_plan_set_m2l_tile(fmma_plan plan, int tile);
/* M2L tile widths (target boxes per workgroup) per number of parent boxes of
 * a level: the autotuner's choices so far for batches (batch = 1), or the
 * widths the last single fmm solve's tasks use (0).  Returns the number of
 * entries.  The single-solve tuner runs on the first fmm solve of a (Depth,
 * M2L form) when autotuning is on. */
int fmma_plan_m2l_tiles(fmma_plan plan, int batch, long *parents, int *tile, int max);

/* Per-stage milliseconds, accumulated over solves like the reference's
 * time_log (include/fmma/fmma.hpp:168, src/fmm.cpp:57-141); the total is
 * overwritten per solve (src/fmma.cpp:189).  Returns the number of entries;
 * names[i] stay valid until the next solve.  Pass max = 0 to count. */
int fmma_plan_time_log(fmma_plan plan, const char **names, double *ms, int max);
int fmma_plan_clear_time_log(fmma_plan plan);

/* Device pointers to a level's coefficients after an fmm solve, Morton box
 * order ([boxes][P], the plan's dtype), for stage checks.  which: 0 = multipole
 * Wm, 1 = local Wl.  Also returns P and the box count of that level. */
int fmma_plan_level_coeffs(fmma_plan plan, int which, int level, const void **ptr, long *n_boxes, int *P);

/* Leaf box of every point of the last fmm solve, as the reference's row-major
 * box id (src/fmm_core.cpp:258-267), in the caller's point order, written to
 * a HOST array of n longs.  which: 0 = sources, 1 = targets. */
int fmma_plan_point_boxes(fmma_plan plan, int which, long n, long *boxes);

/* Morton index <-> row-major box id (src/fmm_core.cpp:205-212) at a level. */
long fmma_morton_to_rowmajor(int dim, int level, long morton);

/* Thread-local text of the last error (never NULL). */
const char *fmma_last_error(void);

/* Library version string. */
const char *fmma_version(void);

#ifdef __cplusplus
}
#endif
#endif
