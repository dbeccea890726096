// plan.hip -- the solve driver and the C-ABI of libfmma_hip.
//
// FMMA<TYPE,DIM>::solve / fmm / exact (src/fmma.cpp:143-207, src/fmm.cpp:15-144)
// become a plan object that owns grow-only device buffers and runs every
// stage as stream-ordered kernels:
//   bounding cube -> leaf keys -> stable radix sort -> box ranges -> gather
//   -> P2M -> M2M (levels L-1..2) -> M2L (levels 2..L, one MFMA launch)
//   -> L2L (levels 2..L-1) -> L2P -> P2P (writes ans in caller order).
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/fmma/detail/pair_kernels.hpp"
#include "batch.hpp"
#include "boxops.hpp"
#include "builtins.hpp"
#include "m2l.hpp"
#include "tree.hpp"

namespace fmma {
namespace lib {

static thread_local std::string g_err = "";

void set_error(const std::string &msg) { g_err = msg; }

int hip_fail(hipError_t e, const char *what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return FMMA_ERR_HIP;
}

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

// ------------------------------------------------------------ no-alloc scratch
//
// fmma_plan_solve_noalloc takes every device buffer of its solve from
// `scratch`, one static array per GPU that lives until the process ends, so
// that it calls no hipMalloc / hipFree.  Its device address comes from
// hipGetSymbolAddress, once per device with that device current (host code
// never passes the array's name to a kernel).  Calls that overlap on one GPU
// share the array and race; each call initialises everything it reads.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// The region a no-alloc solve carves its buffers from (bump allocation),
// active on the calling thread for the duration of that solve.
struct Arena {
    char *base;
    size_t cap, used;
};
static thread_local Arena *t_arena = nullptr;

struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool borrowed = false;  // carved from the no-alloc scratch: never freed
    int ensure(size_t bytes)
    {
        if (t_arena) return carve(bytes);
        if (bytes <= cap) return FMMA_OK;
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        borrowed = false;
        const hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return hip_fail(e, "hipMalloc");
        }
        cap = bytes;
        return FMMA_OK;
    }
    int carve(size_t bytes)
    {
        if (borrowed && bytes <= cap) return FMMA_OK;
        if (p && !borrowed) {
            set_error("internal: a buffer of the no-alloc solve was allocated before it");
            return FMMA_ERR_INVALID;
        }
        const size_t off = (t_arena->used + 255) & ~size_t(255);
        if (off > t_arena->cap || bytes > t_arena->cap - off) {
            set_error("fmma_plan_solve_noalloc: this solve needs more than the " +
                      std::to_string(kNoAllocScratchBytes >> 20) + " MiB of no-alloc scratch (" +
                      std::to_string((off + bytes + (1 << 20) - 1) >> 20) +
                      " MiB or more); use fmma_plan_solve or a smaller problem");
            return FMMA_ERR_SIZE;
        }
        p = t_arena->base + off;
        cap = bytes;
        borrowed = true;
        t_arena->used = off + bytes;
        return FMMA_OK;
    }
    template <class X>
    X *as() const { return static_cast<X *>(p); }
    ~DBuf()
    {
        if (p && !borrowed) (void)hipFree(p);
    }
};

// ------------------------------------------------ host Chebyshev (FP64, libm)
// T_j(x) = cos(j acos x), S_n(x,y) = 1/n + sum_j (2/n) T_j(x) T_j(y)
// (src/math.cpp:14-50); nodes c_k = cos((2k+1) pi / (2n)) (src/fmm_core.cpp:240-243).
static double host_T(int j, double x) { return std::cos(j * std::acos(x)); }

static double host_S(int n, double x, double y)
{
    double s = 1.0 / n;
    for (int j = 1; j < n; ++j) s += 2.0 / n * host_T(j, x) * host_T(j, y);
    return s;
}

// The layout of a batch of independent problems (fmma_plan_solve_batch),
// kept while the problem sizes and the plan's settings stay the same.
struct BatchPlan {
    std::vector<long> tc, sc;  // the key: point counts, shared, settings generation, M2L mode
    bool shared = false;
    unsigned long gen = 0;
    int mode = -1;

    int nprob = 0, Lmax = 0, bits = 0;
    long nt = 0, ns = 0, NB = 0, rows = 0;
    std::vector<int> L;                  // leaf level per problem
    std::vector<unsigned> bbase;         // [nprob+1] first global leaf box
    std::vector<long> row_off;           // [nprob][kMaxLevels] first row (of P) of a level
    std::vector<long> step_first, step_count;  // M2M steps, then L2L steps: ranges of items
    int n_m2m_steps = 0;
    std::vector<int> ent_level, ent_prob;      // M2L operator sets (problem, level)
    std::vector<M2LGroup> groups;
    bool tuned = false;                  // tile widths chosen (autotuner or rule)
    DBuf d_toff, d_soff, d_levels, d_bbase, d_box_prob, d_geoms, d_items, d_tasks, d_ent_level, d_ent_prob;

    hipGraph_t graph = nullptr;          // the captured solve and every address it was captured with
    hipGraphExec_t exec = nullptr;
    std::vector<const void *> g_bufs;

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~BatchPlan() { drop_graph(); }
};

// "P2P sum" is the combine pass of the mutual P2P form (zero for the direct form)
static const char *kStageNames[] = {"tree", "P2M", "M2M", "M2L", "L2L", "L2P", "P2P", "P2P sum"};
static const char *kTreeStageNames[] = {"tree", "P2M", "M2M", "M2P", "P2P"};
static const char *kNrnStageNames[] = {"tree", "P2M", "M2P", "P2P"};
constexpr int kStages = 8;

struct Plan {
    int dim = 3, dtype = FMMA_F64;
    std::string solver = "exact";
    int poly_ord = 1, depth = -1, trans_sym = 0, nrn_N = -1;
    int kernel_id = FMMA_KERNEL_INV_R;
    BuiltinParams kparams;
    fmma_kernel_ops ops{};
    int precision = FMMA_M2L_NATIVE;
    hipStream_t stream = nullptr;
    bool timing = true;
    bool p2p_mutual = true;
    std::map<std::string, double> time_log;
    std::vector<std::string> names_out;
    hipEvent_t ev[kStages + 1] = {};
    bool events_ready = false;
    void *h_geom = nullptr;         // pinned host copy of the bounding cube
    hipEvent_t geom_ev = nullptr;   // recorded after that copy

    DBuf geom, bbox_part;
    DBuf s_keys, s_vals, s_keys_alt, s_vals_alt, s_hist, s_pos, s_w, s_start;
    DBuf t_keys, t_vals, t_keys_alt, t_vals_alt, t_hist, t_pos, t_start;
    DBuf Wm, Wl, Kraw, Kf, Kf_lo, Wc, Wc_lo, l2p_out, exact_part, p2p_slots;
    DBuf d_nodes, d_coef, d_tab, d_far, d_cls_f, d_cls_o, d_tasks, d_op_levels, d_op_prob, d_pairs;
    int n_pairs = 0;
    DBuf h_t, h_s, h_w, h_ans;
    DBuf pos_alt;  // SoA coordinates and weights between the passes of the point sort

    // batches: the layout, graph replay, the M2L tile autotuner and its choices
    std::unique_ptr<BatchPlan> batch;
    bool batch_graph = true, m2l_autotune = true;
    bool batch_eager = false;             // the last batched solve ran eagerly (stage marks recorded)
    int forced_tb = 0;                    // > 0: this M2L tile width wherever a pipelined kernel has it
    std::map<long, int> tuned_tb;        // batches: (dim, P, mode, parents per level) -> tile width
    std::map<long, int> tuned_tb_single; // single problems: (dim, P, mode, parents per level) -> tile width
    std::map<long, int> used_tb_single;  // the last single solve's tasks: parents per level -> tile width
    int tuned_set_L = -1, tuned_set_mode = -1;  // the single-problem tasks carry tuned widths
    hipStream_t cap_stream = nullptr;     // private stream for graph capture

    // cached set-up keyed by (poly_ord, leaf level, mode)
    int set_n = -1, set_L = -1, set_mode = -1;
    OffsetTables offs;
    int n_tasks = 0;
    std::vector<M2LGroup> groups;
    bool p2p_balance = true;       // direct P2P: split boxes with a near field far above the average
    DBuf split_parts, split_cnt, split_items, split_part;
    std::vector<M2LSplit> splits;  // M2L levels whose offsets are split over several tasks
    DBuf Wpart;                    // their partial sums
    long lvlc_off[kMaxLevels + 1] = {};
    long lvl_off[kMaxLevels + 1] = {};
    long kf_off[kMaxLevels] = {};
    int last_L = -1, last_shared = 0;
    bool last_m2l_general = false;  // the last fmm passes evaluated f per box pair in M2L
    long last_nt = 0, last_ns = 0;

    // The tree and local expansions of the last fmm passes, for a gradient
    // that reuses them: valid while the inputs and settings are unchanged.
    struct View {
        bool valid = false;
        const void *tg = nullptr, *w = nullptr, *sr = nullptr;
        long nt = 0, ns = 0;
        unsigned long gen = 0;
        int L = -1;
        bool shared = false;
        const void *t_pos = nullptr;
        const unsigned *t_start = nullptr, *t_perm = nullptr, *t_keys = nullptr;
    } view;
    unsigned long gen = 1;  // bumped by every setter that changes results
    DBuf grad_far;

    ~Plan()
    {
        batch.reset();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        if (events_ready)
            for (auto &e : ev) (void)hipEventDestroy(e);
        if (geom_ev) (void)hipEventDestroy(geom_ev);
        if (h_geom) (void)hipHostFree(h_geom);
    }

    size_t tsize() const { return dtype == FMMA_F64 ? 8 : 4; }
    int P() const
    {
        int p = 1;
        for (int d = 0; d < dim; ++d) p *= poly_ord + 1;
        return p;
    }
    int mode() const
    {
        if (precision == FMMA_M2L_BF16X3) return 2;
        if (precision == FMMA_M2L_FP32 || dtype == FMMA_F32) return 1;
        return 0;
    }
    void mark(int i)
    {
        if (timing && events_ready) (void)hipEventRecord(ev[i], stream);
    }
};

template <typename T>
static void to_T(const std::vector<double> &src, std::vector<unsigned char> &dst)
{
    dst.resize(src.size() * sizeof(T));
    T *o = reinterpret_cast<T *>(dst.data());
    for (size_t i = 0; i < src.size(); ++i) o[i] = (T)src[i];
}

static int upload(DBuf &b, const void *src, size_t bytes, hipStream_t st)
{
    int rc = b.ensure(bytes);
    if (rc) return rc;
    if (bytes) FMMA_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
    return FMMA_OK;
}

static long tune_key(int D, int P, int mode, long npar)
{
    return (((long)D * 1024 + P) * 4 + mode) * (1L << 32) + npar;
}

// The tile width of each level's tasks (by parents per level), as reported
// by fmma_plan_m2l_tiles for single solves.
static void record_tiles(Plan &pl, const std::vector<M2LTask> &tasks, int D)
{
    pl.used_tb_single.clear();
    for (const M2LGroup &g : pl.groups)
        for (int i = g.first; i < g.first + g.count; ++i) pl.used_tb_single[1L << (D * (tasks[i].level - 1))] = g.tb;
}

// Constant tables that depend on (poly_ord, dim, dtype): nodes, Chebyshev
// weight coefficients, M2M/L2L transfer tables, M2L offset tables.
static int setup_basis(Plan &pl)
{
    const int n = pl.poly_ord + 1, D = pl.dim;
    hipStream_t st = pl.stream;
    int rc;
    if (pl.set_n != n || pl.offs.D != D) {
        std::vector<double> nodes(n), coef(n * n), tab(2 * n * n);
        for (int k = 0; k < n; ++k) nodes[k] = std::cos((2.0 * k + 1.0) / (2 * pl.poly_ord + 2) * M_PI);
        // nodes as TYPE, as the reference stores them (chebyshev_node_all)
        std::vector<double> tn(n);
        for (int k = 0; k < n; ++k) tn[k] = pl.dtype == FMMA_F64 ? nodes[k] : (double)(float)nodes[k];
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j) coef[k * n + j] = j == 0 ? 1.0 / n : 2.0 / n * host_T(j, tn[k]);
        for (int s = 0; s < 2; ++s)
            for (int ko = 0; ko < n; ++ko)
                for (int ki = 0; ki < n; ++ki)
                    tab[(s * n + ko) * n + ki] = host_S(n, tn[ko], (tn[ki] + (2 * s - 1)) / 2);
        std::vector<unsigned char> cb, tb;
        if (pl.dtype == FMMA_F64) {
            to_T<double>(coef, cb);
            to_T<double>(tab, tb);
        } else {
            to_T<float>(coef, cb);
            to_T<float>(tab, tb);
        }
        if ((rc = upload(pl.d_nodes, tn.data(), n * sizeof(double), st))) return rc;
        if ((rc = upload(pl.d_coef, cb.data(), cb.size(), st))) return rc;
        if ((rc = upload(pl.d_tab, tb.data(), tb.size(), st))) return rc;
        pl.offs.build(D);
        if ((rc = upload(pl.d_far, pl.offs.far.data(), pl.offs.far.size() * sizeof(int), st))) return rc;
        if ((rc = upload(pl.d_cls_f, pl.offs.cls_f.data(), pl.offs.cls_f.size() * sizeof(int), st))) return rc;
        if ((rc = upload(pl.d_cls_o, pl.offs.cls_o.data(), pl.offs.cls_o.size() * sizeof(int), st))) return rc;
        // (o, -o) pairs of far offsets, for the paired operator build
        std::vector<int> pairs;
        for (int f = 0; f < pl.offs.n_far; ++f)
            for (int g = f + 1; g < pl.offs.n_far; ++g) {
                bool neg = true;
                for (int d = 0; d < D; ++d) neg = neg && pl.offs.far[g * D + d] == -pl.offs.far[f * D + d];
                if (neg) {
                    pairs.push_back(f);
                    pairs.push_back(g);
                }
            }
        pl.n_pairs = (int)pairs.size() / 2;
        if ((rc = upload(pl.d_pairs, pairs.data(), pairs.size() * sizeof(int), st))) return rc;
        // keep host vectors alive until the copies are done
        FMMA_HIP(hipStreamSynchronize(st));
        pl.set_n = n;
        pl.set_L = -1;
    }
    return FMMA_OK;
}

// ... and those that depend on the leaf level and the M2L mode.
// Workgroups below which an M2L level's offsets are split over several
// tasks (none in a no-alloc solve: its scratch holds no partial sums).
static long split_target() { return t_arena ? 0 : kM2LSplitTarget; }

static int setup(Plan &pl, int L)
{
    const int n = pl.poly_ord + 1, D = pl.dim, P = pl.P(), mode = pl.mode();
    (void)n;
    hipStream_t st = pl.stream;
    int rc;
    if ((rc = setup_basis(pl))) return rc;
    if (pl.set_L != L || pl.set_mode != mode) {
        long off = 0;
        for (int l = 0; l <= L; ++l) {
            pl.lvl_off[l] = off;
            off += (long)P << (D * l);
        }
        pl.lvl_off[L + 1] = off;
        if ((rc = pl.Wm.ensure(off * pl.tsize()))) return rc;
        if ((rc = pl.Wl.ensure(off * pl.tsize()))) return rc;
        if (L >= 2) {
            const size_t fe = m2l_frag_elems(pl.offs.n_far, P, mode);
            const size_t esz = mode == 0 ? 8 : mode == 1 ? 4 : 2;
            for (int l = 0; l < kMaxLevels; ++l) pl.kf_off[l] = l >= 2 ? (long)(l - 2) * (long)fe : 0;
            if ((rc = pl.Kf.ensure(fe * esz * (L - 1)))) return rc;
            if (mode == 2 && (rc = pl.Kf_lo.ensure(fe * esz * (L - 1)))) return rc;
            // padded operand rows of KP for levels 2..L
            int MP, KP;
            m2l_shape(P, mode, MP, KP);
            long offc = 0;
            for (int l = 0; l <= L; ++l) {
                pl.lvlc_off[l] = offc;
                if (l >= 2) offc += (long)KP << (D * l);
            }
            const size_t opsz = mode == 0 ? pl.tsize() : mode == 1 ? 4 : 2;
            if ((rc = pl.Wc.ensure(offc * opsz))) return rc;
            if (mode == 2 && (rc = pl.Wc_lo.ensure(offc * 2))) return rc;
            std::vector<M2LTask> tasks = make_m2l_tasks(D, L, P, mode, pl.groups, pl.lvl_off, pl.lvlc_off, pl.kf_off,
                                                        nullptr, -1, &pl.splits, split_target());
            if ((rc = pl.Wpart.ensure((size_t)m2l_split_elems(pl.splits) * pl.tsize()))) return rc;
            record_tiles(pl, tasks, D);
            std::vector<int> op_levels, op_prob(L - 1, 0);
            for (int l = 2; l <= L; ++l) op_levels.push_back(l);
            if ((rc = upload(pl.d_op_levels, op_levels.data(), op_levels.size() * sizeof(int), st))) return rc;
            if ((rc = upload(pl.d_op_prob, op_prob.data(), op_prob.size() * sizeof(int), st))) return rc;
            pl.n_tasks = (int)tasks.size();
            if ((rc = upload(pl.d_tasks, tasks.data(), tasks.size() * sizeof(M2LTask), st))) return rc;
            FMMA_HIP(hipStreamSynchronize(st));
        } else {
            pl.n_tasks = 0;
        }
        pl.set_L = L;
        pl.set_mode = mode;
        pl.tuned_set_L = -1;
    }
    return FMMA_OK;
}

// Reads back the bounding cube just computed and rejects inputs the
// reference cannot bin (Geom::status), before any kernel uses the cube.
static int geom_status_error(int status)
{
    if (status == kGeomNonFinite)
        return fail(FMMA_ERR_INVALID, "a target or source coordinate is NaN or infinite");
    if (status == kGeomZeroLen)
        return fail(FMMA_ERR_INVALID, "every target and source is at the same position: the bounding cube has "
                                      "size 0 and its boxes are undefined");
    return FMMA_OK;
}

template <typename T>
static int check_geom(Plan &pl, Geom<T> *out = nullptr)
{
    Geom<T> g;
    FMMA_HIP(hipMemcpyAsync(&g, pl.geom.p, sizeof(g), hipMemcpyDeviceToHost, pl.stream));
    FMMA_HIP(hipStreamSynchronize(pl.stream));
    if (out) *out = g;
    return geom_status_error(g.status);
}

// The same in two steps: geom_fetch queues the copy of the bounding cube to
// pinned host memory and returns at once, geom_wait waits for that copy
// (long done by then) and checks it.  The kernels queued in between (tree,
// P2M, M2M) write only the plan's own buffers and stay in bounds on any cube
// (leaf coordinates are clamped), so the solve still fails before anything
// reaches ans.  Without an allocation the check is made at once
// (check_geom): the no-allocation entry point allocates no pinned memory.
template <typename T>
static int geom_fetch(Plan &pl, Geom<T> *now, bool &pending)
{
    pending = false;
    if (t_arena) return check_geom<T>(pl, now);
    if (!pl.h_geom) FMMA_HIP(hipHostMalloc(&pl.h_geom, sizeof(Geom<double>), hipHostMallocDefault));
    if (!pl.geom_ev) FMMA_HIP(hipEventCreateWithFlags(&pl.geom_ev, hipEventDisableTiming));
    FMMA_HIP(hipMemcpyAsync(pl.h_geom, pl.geom.p, sizeof(Geom<T>), hipMemcpyDeviceToHost, pl.stream));
    FMMA_HIP(hipEventRecord(pl.geom_ev, pl.stream));
    pending = true;
    return FMMA_OK;
}

template <typename T>
static int geom_wait(Plan &pl, Geom<T> *out, bool pending)
{
    if (!pending) return FMMA_OK;
    FMMA_HIP(hipEventSynchronize(pl.geom_ev));
    std::memcpy(out, pl.h_geom, sizeof(Geom<T>));
    return geom_status_error(out->status);
}

// With trans_sym_flag 0 the reference evaluates f at absolute node positions
// len*(b + (1+c)/2) + origin (src/fmm_core.cpp:466-487); per-offset
// operators use the relative ones of flag 1 (src/fmm_core.cpp:407-408).  For
// a translation-invariant f the two differ only by the rounding of the
// absolute positions, about eps * |position| / len_leaf relative: measured
// 2.4e-17 times that ratio in FP64 (4.0e-12 for [1e5, 1e5 + 1)^3 at leaf
// level 3).  Past these ratios (FP64: about 1e-13; FP32: a tenth of its
// 3e-5 bar) a flag-0 solve evaluates f per box pair as the reference does.
template <typename T>
static bool absolute_positions_matter(const Geom<T> &g, int D, int L)
{
    const double limit = sizeof(T) == 8 ? 4096.0 : 256.0;
    double R = 0.0;
    for (int d = 0; d < D; ++d) {
        const double o = (double)g.origin[d];
        R = std::max(R, std::max(std::fabs(o), std::fabs(o + (double)g.Len)));
    }
    return R * std::ldexp(1.0, L) > limit * (double)g.Len;
}

// Sort one point set into its leaf boxes: keys -> radix sort -> ranges ->
// gather.  Boxes: the Morton-ordered 2^L grid, or with grid_n > 0 the
// row-major grid_n^D grid.
template <typename T, int D>
static int build_tree(Plan &pl, const T *pts, const T *w, long n, int L, bool with_w, DBuf &keys, DBuf &vals,
                      DBuf &keys_alt, DBuf &vals_alt, DBuf &hist, DBuf &pos, DBuf *wbuf, DBuf &start,
                      int grid_n = 0)
{
    unsigned nbox = 1u << (D * L);
    int bits = D * L;
    if (grid_n > 0) {
        nbox = 1;
        for (int d = 0; d < D; ++d) nbox *= (unsigned)grid_n;
        bits = 0;
        while (bits < 32 && (1ull << bits) < nbox) ++bits;
    }
    int rc;
    if ((rc = keys.ensure(n * 4)) || (rc = vals.ensure(n * 4)) || (rc = keys_alt.ensure(n * 4)) ||
        (rc = vals_alt.ensure(n * 4)) || (rc = hist.ensure(radix_hist_words(n) * 4)) ||
        (rc = pos.ensure((size_t)n * D * sizeof(T))) || (rc = start.ensure(((size_t)nbox + 1) * 4)))
        return rc;
    if (with_w && (rc = wbuf->ensure((size_t)n * sizeof(T)))) return rc;
    if ((rc = pl.pos_alt.ensure((size_t)(D + 1) * n * sizeof(T)))) return rc;
    const Geom<T> *g = pl.geom.as<Geom<T>>();
    if ((rc = leaf_keys<T, D>(pts, n, g, L, grid_n, keys.as<unsigned>(), pl.stream,
                              bits > 0 ? hist.as<unsigned>() : nullptr)))
        return rc;
    PointSortBuffers<T> sb{keys.as<unsigned>(), vals.as<unsigned>(), keys_alt.as<unsigned>(),
                           vals_alt.as<unsigned>(), hist.as<unsigned>(), pos.as<T>(),
                           with_w ? wbuf->as<T>() : nullptr, pl.pos_alt.as<T>(), pl.pos_alt.as<T>() + (size_t)D * n};
    if ((rc = sort_points<T, D>(sb, n, bits, pts, w, with_w, pl.stream, bits > 0))) return rc;
    return box_ranges(keys.as<unsigned>(), n, nbox, start.as<unsigned>(), pl.stream);
}

// The M2L tile autotuner for one problem: per level, time every pipelined
// tile width (16..256 target boxes per workgroup) on the level's own tasks
// (a warm-up and 3 runs, HIP events), keep the fastest, and rebuild the
// plan's tasks with the choices.  Runs on the first fmm solve of a (Depth,
// form); the choices are cached by (dimension, P, form, parents per level).
template <typename T, int D>
static int tune_single(Plan &pl, const M2LKArgs &ka, int L)
{
    const int P = pl.P(), mode = pl.mode();
    hipStream_t st = pl.stream;
    int rc;
    int tb_of[kMaxLevels] = {};
    for (int l = 2; l <= L; ++l) {
        const long npar = 1L << (D * (l - 1));
        if (pl.forced_tb > 0) {
            tb_of[l] = m2l_pipe_supported(D, ka.KP, mode, pl.forced_tb) ? pl.forced_tb : 0;
            continue;
        }
        const long key = tune_key(D, P, mode, npar);
        auto it = pl.tuned_tb_single.find(key);
        if (it != pl.tuned_tb_single.end()) {
            tb_of[l] = it->second;
            continue;
        }
        float best = 1e30f;
        int best_tb = 0;
        for (int tb : {16, 32, 64, 128, 256}) {
            if (!m2l_pipe_supported(D, ka.KP, mode, tb) || tb > 4 * npar) continue;
            int one[kMaxLevels] = {};
            one[l] = tb;
            std::vector<M2LGroup> g;
            std::vector<M2LSplit> sp;
            std::vector<M2LTask> t = make_m2l_tasks(D, L, P, mode, g, pl.lvl_off, pl.lvlc_off, pl.kf_off, one, l, &sp,
                                                    split_target());
            if ((rc = pl.Wpart.ensure((size_t)m2l_split_elems(sp) * pl.tsize()))) return rc;
            DBuf dt;
            if ((rc = upload(dt, t.data(), t.size() * sizeof(M2LTask), st))) return rc;
            M2LKArgs a2 = ka;
            a2.tasks = dt.as<M2LTask>();
            a2.Wpart = pl.Wpart.p;
            if ((rc = m2l_apply<T, D>(a2, g, mode, st))) return rc;  // warm-up
            hipEvent_t e0, e1;
            FMMA_HIP(hipEventCreate(&e0));
            FMMA_HIP(hipEventCreate(&e1));
            FMMA_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < 3 && !rc; ++r) rc = m2l_apply<T, D>(a2, g, mode, st);
            FMMA_HIP(hipEventRecord(e1, st));
            FMMA_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FMMA_HIP(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (rc) return rc;
            if (ms < best) {
                best = ms;
                best_tb = tb;
            }
        }
        pl.tuned_tb_single[key] = best_tb;
        tb_of[l] = best_tb;
    }
    std::vector<M2LTask> tasks = make_m2l_tasks(D, L, P, mode, pl.groups, pl.lvl_off, pl.lvlc_off, pl.kf_off, tb_of,
                                                -1, &pl.splits, split_target());
    if ((rc = pl.Wpart.ensure((size_t)m2l_split_elems(pl.splits) * pl.tsize()))) return rc;
    record_tiles(pl, tasks, D);
    pl.n_tasks = (int)tasks.size();
    if ((rc = upload(pl.d_tasks, tasks.data(), tasks.size() * sizeof(M2LTask), st))) return rc;
    FMMA_HIP(hipStreamSynchronize(st));
    pl.tuned_set_L = L;
    return FMMA_OK;
}

// Tree, P2M, M2M, M2L and L2L (stages 0-4); records pl.view.
template <typename T, int D>
static int fmm_passes(Plan &pl, long nt, const T *tg, const T *w, long ns, const T *sr)
{
    pl.view.valid = false;
    const int L = pl.depth - 1;
    if (D * L > 30 || L >= kMaxLevels)
        return fail(FMMA_ERR_UNSUPPORTED, "Depth " + std::to_string(pl.depth) + " too deep for dimension " +
                                              std::to_string(D) + " (at most 2^30 leaf boxes)");
    const int n = pl.poly_ord + 1, P = pl.P(), mode = pl.mode();
    if (P > kMaxP) return fail(FMMA_ERR_UNSUPPORTED, "(poly_ord+1)^DIM > 512 is not supported");
    int rc;
    if ((rc = setup(pl, L))) return rc;
    hipStream_t st = pl.stream;
    const bool shared = (const void *)tg == (const void *)sr && nt == ns;
    const unsigned nbox = 1u << (D * L);

    pl.mark(0);
    if ((rc = pl.geom.ensure(sizeof(Geom<T>))) || (rc = pl.bbox_part.ensure(kBBoxParts * kBBoxStride<D> * sizeof(T))))
        return rc;
    if ((rc = bbox<T, D>(tg, nt, sr, ns, pl.bbox_part.as<T>(), pl.geom.as<Geom<T>>(), 1 << L, st))) return rc;
    Geom<T> hgeom;
    bool geom_pending;
    if ((rc = geom_fetch<T>(pl, &hgeom, geom_pending))) return rc;
    if ((rc = build_tree<T, D>(pl, sr, w, ns, L, true, pl.s_keys, pl.s_vals, pl.s_keys_alt, pl.s_vals_alt,
                               pl.s_hist, pl.s_pos, &pl.s_w, pl.s_start)))
        return rc;
    const T *t_pos = pl.s_pos.as<T>();
    const unsigned *t_start = pl.s_start.as<unsigned>(), *t_perm = pl.s_vals.as<unsigned>();
    const unsigned *t_keys = pl.s_keys.as<unsigned>();
    if (!shared) {
        if ((rc = build_tree<T, D>(pl, tg, nullptr, nt, L, false, pl.t_keys, pl.t_vals, pl.t_keys_alt,
                                   pl.t_vals_alt, pl.t_hist, pl.t_pos, nullptr, pl.t_start)))
            return rc;
        t_pos = pl.t_pos.as<T>();
        t_start = pl.t_start.as<unsigned>();
        t_perm = pl.t_vals.as<unsigned>();
        t_keys = pl.t_keys.as<unsigned>();
    }
    pl.mark(1);

    T *Wm = pl.Wm.as<T>(), *Wl = pl.Wl.as<T>();
    P2MArgs<T> pa{pl.s_start.as<unsigned>(), pl.s_pos.as<T>(), pl.s_w.as<T>(), ns, pl.geom.as<Geom<T>>(), L,
                  pl.d_coef.as<T>(), n, P, Wm + pl.lvl_off[L]};
    if ((rc = p2m<T, D>(pa, nbox, st))) return rc;
    pl.mark(2);

    for (int l = L - 1; l >= 2; --l)
        if ((rc = m2m<T, D>(Wm + pl.lvl_off[l + 1], Wm + pl.lvl_off[l], 1L << (D * l), pl.d_tab.as<T>(), n, P, st)))
            return rc;
    pl.mark(3);
    if ((rc = geom_wait<T>(pl, &hgeom, geom_pending))) return rc;

    if (L >= 2) {
        const bool sym =
            pl.trans_sym == 1 || (pl.ops.translation_invariant && !absolute_positions_matter(hgeom, D, L));
        pl.last_m2l_general = !sym;
        if (sym) {
            // the operators of levels 2..L straight into fragment order
            fmma_m2l_op_args oa{};
            oa.dim = D;
            oa.dtype = pl.dtype;
            oa.n_nodes = n;
            oa.P = P;
            oa.nodes = pl.d_nodes.as<double>();
            oa.Len = &pl.geom.as<Geom<T>>()->Len;
            oa.offsets = pl.d_far.as<int>();
            oa.n_off = pl.offs.n_far;
            oa.format = m2l_frag_format(P, mode);
            m2l_shape(P, mode, oa.MP, oa.KP);
            oa.n_batch = L - 1;
            oa.len_stride = 0;
            oa.batch_level = pl.d_op_levels.as<int>();
            oa.batch_problem = pl.d_op_prob.as<int>();
            oa.Kf = pl.Kf.p;
            oa.Kf_lo = pl.Kf_lo.p;
            oa.pairs = pl.d_pairs.as<int>();
            oa.n_pairs = pl.n_pairs;
            rc = pl.ops.m2l_op(&oa, pl.ops.ctx, st);
            if (rc == FMMA_ERR_UNSUPPORTED) {
                // a launcher table without the fragment form: one level at a time through Kraw
                if ((rc = pl.Kraw.ensure((size_t)pl.offs.n_far * P * P * pl.tsize()))) return rc;
                for (int l = 2; l <= L; ++l) {
                    fmma_m2l_op_args ra{};
                    ra.dim = D;
                    ra.dtype = pl.dtype;
                    ra.n_nodes = n;
                    ra.P = P;
                    ra.nodes = pl.d_nodes.as<double>();
                    ra.level = l;
                    ra.Len = &pl.geom.as<Geom<T>>()->Len;
                    ra.offsets = pl.d_far.as<int>();
                    ra.n_off = pl.offs.n_far;
                    ra.K = pl.Kraw.p;
                    ra.format = FMMA_FRAG_RAW;
                    if ((rc = pl.ops.m2l_op(&ra, pl.ops.ctx, st))) return rc;
                    char *kf = pl.Kf.as<char>(), *kl = pl.Kf_lo.as<char>();
                    const size_t esz = mode == 0 ? 8 : mode == 1 ? 4 : 2;
                    if ((rc = m2l_fragments<T>(pl.Kraw.as<T>(), pl.offs.n_far, P, mode, kf + pl.kf_off[l] * esz,
                                               mode == 2 ? kl + pl.kf_off[l] * esz : nullptr, st)))
                        return rc;
                }
            } else if (rc) {
                return rc;
            }
            M2LKArgs ka{};
            ka.tasks = pl.d_tasks.as<M2LTask>();
            ka.cls_f = pl.d_cls_f.as<int>();
            ka.cls_o = pl.d_cls_o.as<int>();
            ka.n_cls = pl.offs.n_cls;
            ka.P = P;
            m2l_shape(P, mode, ka.MP, ka.KP);
            ka.Wm = Wm;
            ka.Wl = Wl;
            ka.Kf = pl.Kf.p;
            ka.Kf_lo = pl.Kf_lo.p;
            ka.Wc = pl.Wc.p;
            ka.Wc_lo = pl.Wc_lo.p;
            {
                const size_t opsz = mode == 0 ? sizeof(T) : mode == 1 ? 4 : 2;
                char *wc = pl.Wc.as<char>() + pl.lvlc_off[2] * opsz;
                char *wl = mode == 2 ? pl.Wc_lo.as<char>() + pl.lvlc_off[2] * 2 : nullptr;
                const long nbox = (pl.lvl_off[L + 1] - pl.lvl_off[2]) / P;
                if ((rc = m2l_convert<T>(Wm + pl.lvl_off[2], nbox, P, ka.KP, mode, wc, wl, st))) return rc;
            }
            if ((pl.m2l_autotune || pl.forced_tb > 0) && pl.tuned_set_L != L && (rc = tune_single<T, D>(pl, ka, L)))
                return rc;
            ka.tasks = pl.d_tasks.as<M2LTask>();
            ka.Wpart = pl.Wpart.p;
            if ((rc = m2l_apply<T, D>(ka, pl.groups, mode, st))) return rc;
            if ((rc = m2l_sum_parts<T>(Wl, pl.Wpart.as<T>(), pl.splits, st))) return rc;
        } else {
            for (int l = 2; l <= L; ++l) {
                fmma_m2l_general_args ga{D, pl.dtype, n, P, pl.d_nodes.as<double>(), l,
                                         &pl.geom.as<Geom<T>>()->Len, pl.geom.as<Geom<T>>()->origin,
                                         Wm + pl.lvl_off[l], Wl + pl.lvl_off[l]};
                if ((rc = pl.ops.m2l_general(&ga, pl.ops.ctx, st))) return rc;
            }
        }
    }
    pl.mark(4);

    for (int l = 2; l + 1 <= L; ++l)
        if ((rc = l2l<T, D>(Wl + pl.lvl_off[l], Wl + pl.lvl_off[l + 1], 1L << (D * l), pl.d_tab.as<T>(), n, P, st)))
            return rc;
    pl.mark(5);
    Plan::View &v = pl.view;
    v.tg = tg;
    v.w = w;
    v.sr = sr;
    v.nt = nt;
    v.ns = ns;
    v.gen = pl.gen;
    v.L = L;
    v.shared = shared;
    v.t_pos = t_pos;
    v.t_start = t_start;
    v.t_perm = t_perm;
    v.t_keys = t_keys;
    v.valid = true;
    pl.last_L = L;
    pl.last_shared = shared ? 1 : 0;
    pl.last_nt = nt;
    pl.last_ns = ns;
    return FMMA_OK;
}

// The direct P2P (dim_out 0) or its gradient (dim_out = D) balanced over
// work items when some target boxes have a near field far above the
// average (clustered points): the planner marks those boxes on the device,
// one read-back gives the item count, then the unsplit boxes, the items and
// the chunk sums run.  Without split boxes (or in a no-alloc solve) this is
// the plain launch.  t_keys: the box of each sorted target.
template <typename T, int D, class A, class Launch>
static int balanced_p2p(Plan &pl, A qa, const unsigned *t_keys, unsigned pass, int dim_out, const T *base, T *out,
                        int accumulate, Launch launch)
{
    const long nbox = qa.n_target_boxes;
    if (t_arena || !pl.p2p_balance || !t_keys || nbox <= 1) return launch(qa);
    hipStream_t st = pl.stream;
    int rc;
    if ((rc = pl.split_parts.ensure((size_t)nbox * 4)) || (rc = pl.split_cnt.ensure(16))) return rc;
    unsigned *parts = pl.split_parts.as<unsigned>(), *cnt = pl.split_cnt.as<unsigned>();
    FMMA_HIP(hipMemsetAsync(cnt, 0, 16, st));
    const unsigned nb = (unsigned)((nbox + 255) / 256);
    hipLaunchKernelGGL((detail::p2p_split_plan_kernel<D, A>), dim3(nb), dim3(256), 0, st, qa, pass, parts, cnt,
                       (unsigned *)nullptr, 0);
    FMMA_LAUNCHED("p2p_split_plan");
    unsigned h[4];
    FMMA_HIP(hipMemcpyAsync(h, cnt, 16, hipMemcpyDeviceToHost, st));
    FMMA_HIP(hipStreamSynchronize(st));
    if (h[0] == 0) return launch(qa);
    const size_t per = (size_t)qa.n_t * (dim_out > 0 ? dim_out : 1) * sizeof(T);
    if ((rc = pl.split_items.ensure((size_t)h[0] * 3 * 4)) || (h[1] > 1 && (rc = pl.split_part.ensure(per * h[1]))))
        return rc;
    hipLaunchKernelGGL((detail::p2p_split_plan_kernel<D, A>), dim3(nb), dim3(256), 0, st, qa, pass, parts, cnt,
                       pl.split_items.as<unsigned>(), 1);
    FMMA_LAUNCHED("p2p_split_plan");
    qa.split_parts = parts;
    if ((rc = launch(qa))) return rc;  // the boxes that were not split
    A qi = qa;
    qi.split_items = pl.split_items.as<unsigned>();
    qi.n_split_items = (long)h[0];
    qi.split_pass = pass;
    qi.split_part = pl.split_part.p;
    if ((rc = launch(qi))) return rc;
    if (h[1] > 1) {
        hipLaunchKernelGGL((detail::p2p_split_combine<T>), dim3((unsigned)((qa.n_t + 255) / 256)), dim3(256), 0, st,
                           t_keys, parts, pl.split_part.as<T>(), qa.n_t, base, qa.t_perm, out, accumulate, dim_out);
        FMMA_LAUNCHED("p2p_split_combine");
    }
    return FMMA_OK;
}

// P2P over the leaf neighbours (the tree's Morton boxes of leaf level L, or
// with grid_n > 0 the row-major grid_n^D boxes): ans[perm[i]] = base[i] +
// near sum (base NULL: the near sum, added onto ans when accumulate).  The
// mutual form runs when targets are the sources and f declares its symmetry;
// mark_mid >= 0 records that stage mark between its two launches.
template <typename T, int D>
static int p2p_stage(Plan &pl, int L, int grid_n, unsigned nbox, const unsigned *t_start, const T *t_pos,
                     const unsigned *t_perm, long nt, long ns, bool shared, const T *base, T *ans, int accumulate,
                     int mark_mid)
{
    hipStream_t st = pl.stream;
    int rc;
    const bool mutual = shared && pl.p2p_mutual && nbox > 1 && (pl.ops.symmetry == 1 || pl.ops.symmetry == -1);
    const int n_upper = ((D == 1 ? 3 : D == 2 ? 9 : 27) - 1) / 2;
    if (mutual && (rc = pl.p2p_slots.ensure((size_t)(n_upper + 1) * nt * sizeof(T)))) return rc;
    fmma_p2p_args qa{D, pl.dtype, (long)nbox, L, t_start, pl.s_start.as<unsigned>(), t_pos, pl.s_pos.as<T>(),
                     pl.s_w.as<T>(), nt, ns, base, t_perm, ans, accumulate,
                     mutual ? pl.p2p_slots.p : nullptr, mutual ? pl.ops.symmetry : 0,
                     nullptr, nullptr, nullptr, grid_n};
    if (mutual) {
        if ((rc = pl.ops.p2p(&qa, pl.ops.ctx, st))) return rc;
    } else {
        const unsigned *t_keys = shared ? pl.s_keys.as<unsigned>() : pl.t_keys.as<unsigned>();
        auto launch = [&](const fmma_p2p_args &x) { return pl.ops.p2p(&x, pl.ops.ctx, st); };
        if ((rc = balanced_p2p<T, D>(pl, qa, t_keys, 64u * detail::kP2PTargets, 0, base, ans, accumulate, launch)))
            return rc;
    }
    if (mark_mid >= 0) pl.mark(mark_mid);
    if (mutual) {
        P2PCombineArgs<T> ca{pl.p2p_slots.as<T>(), nt, base, t_perm, ans, accumulate};
        if ((rc = p2p_combine<T, D>(ca, st))) return rc;
    }
    return FMMA_OK;
}

// L2P and P2P (stages 5-7) on the passes recorded in pl.view.
template <typename T, int D>
static int fmm_eval(Plan &pl, T *ans)
{
    const Plan::View &v = pl.view;
    const int L = v.L, n = pl.poly_ord + 1, P = pl.P();
    const long nt = v.nt, ns = v.ns;
    const bool shared = v.shared;
    const unsigned nbox = 1u << (D * L);
    const T *t_pos = static_cast<const T *>(v.t_pos);
    const unsigned *t_start = v.t_start, *t_perm = v.t_perm, *t_keys = v.t_keys;
    T *Wl = pl.Wl.as<T>();
    hipStream_t st = pl.stream;
    int rc;
    const T *base = nullptr;
    L2PArgs<T> la{t_start, t_keys, t_pos, nt, pl.geom.as<Geom<T>>(), L, pl.d_coef.as<T>(), n, P, Wl + pl.lvl_off[L],
                  nullptr};
    const bool mutual = shared && pl.p2p_mutual && nbox > 1 && (pl.ops.symmetry == 1 || pl.ops.symmetry == -1);
    if (L >= 2 && mutual && n >= 2 && n <= 8) {
        // the mutual near field, then L2P inside its combine pass (stage "P2P sum")
        pl.mark(6);
        const int n_upper = ((D == 1 ? 3 : D == 2 ? 9 : 27) - 1) / 2;
        if ((rc = pl.p2p_slots.ensure((size_t)(n_upper + 1) * nt * sizeof(T)))) return rc;
        fmma_p2p_args qa{D, pl.dtype, (long)nbox, L, t_start, pl.s_start.as<unsigned>(), t_pos, pl.s_pos.as<T>(),
                         pl.s_w.as<T>(), nt, ns, nullptr, t_perm, ans, 0, pl.p2p_slots.p, pl.ops.symmetry,
                         nullptr, nullptr, nullptr, 0};
        if ((rc = pl.ops.p2p(&qa, pl.ops.ctx, st))) return rc;
        pl.mark(7);
        P2PCombineArgs<T> ca{pl.p2p_slots.as<T>(), nt, nullptr, t_perm, ans, 0};
        if ((rc = l2p_combine<T, D>(la, ca, st))) return rc;
        pl.mark(8);
        return FMMA_OK;
    }
    if (L >= 2) {
        if ((rc = pl.l2p_out.ensure((size_t)nt * sizeof(T)))) return rc;
        la.out = pl.l2p_out.as<T>();
        if ((rc = l2p<T, D>(la, nbox, st))) return rc;
        base = pl.l2p_out.as<T>();
    }
    pl.mark(6);
    if ((rc = p2p_stage<T, D>(pl, L, 0, nbox, t_start, t_pos, t_perm, nt, ns, shared, base, ans, L < 2 ? 1 : 0, 7)))
        return rc;
    pl.mark(8);
    return FMMA_OK;
}

template <typename T, int D>
static int run_fmm(Plan &pl, long nt, const T *tg, const T *w, long ns, const T *sr, T *ans)
{
    int rc;
    if ((rc = fmm_passes<T, D>(pl, nt, tg, w, ns, sr))) return rc;
    return fmm_eval<T, D>(pl, ans);
}

// Gradient of the fmm result with respect to the target coordinates: the
// L2P interpolant differentiated, plus the near field of grad_x f.
template <typename T, int D>
static int fmm_eval_grad(Plan &pl, T *grad)
{
    const Plan::View &v = pl.view;
    const int L = v.L, n = pl.poly_ord + 1, P = pl.P();
    const long nt = v.nt, ns = v.ns;
    const unsigned nbox = 1u << (D * L);
    hipStream_t st = pl.stream;
    int rc;
    const T *base = nullptr;
    if (L >= 2) {
        if ((rc = pl.grad_far.ensure((size_t)D * nt * sizeof(T)))) return rc;
        L2PArgs<T> la{v.t_start, v.t_keys, static_cast<const T *>(v.t_pos), nt, pl.geom.as<Geom<T>>(), L,
                      pl.d_coef.as<T>(), n, P, pl.Wl.as<T>() + pl.lvl_off[L], pl.grad_far.as<T>()};
        if ((rc = l2p_grad<T, D>(la, st))) return rc;
        base = pl.grad_far.as<T>();
    }
    fmma_p2p_grad_args ga{D, pl.dtype, (long)nbox, L, v.t_start, pl.s_start.as<unsigned>(), v.t_pos,
                          pl.s_pos.as<T>(), pl.s_w.as<T>(), nt, ns, base, v.t_perm, grad};
    auto launch = [&](const fmma_p2p_grad_args &x) { return pl.ops.p2p_grad(&x, pl.ops.ctx, st); };
    return balanced_p2p<T, D>(pl, ga, v.t_keys, 64u * detail::kP2PGradTargets, D, base, grad, 0, launch);
}

template <typename T, int D>
static int run_exact(Plan &pl, long nt, const T *tg, const T *w, long ns, const T *sr, T *ans)
{
    long waves = (nt + 63) / 64;
    int split = 1;
    if (waves < 8192) {
        long s = 8192 / (waves > 0 ? waves : 1);
        long smax = ns / 1024;
        s = s < smax ? s : smax;
        s = s < 256 ? s : 256;
        split = s > 1 ? (int)s : 1;
    }
    int rc;
    if (split > 1 && (rc = pl.exact_part.ensure((size_t)split * nt * sizeof(T)))) return rc;
    fmma_exact_args ea{D, pl.dtype, nt, ns, tg, w, sr, ans, split > 1 ? pl.exact_part.p : nullptr, split};
    return pl.ops.exact(&ea, pl.ops.ctx, pl.stream);
}

// Both point sets into their leaf boxes (the sources with their weights);
// returns the target side's sorted arrays.
template <typename T, int D>
static int sort_points(Plan &pl, long nt, const T *tg, const T *w, long ns, const T *sr, int L, int grid_n,
                       const T *&t_pos, const unsigned *&t_start, const unsigned *&t_perm)
{
    int rc;
    if ((rc = build_tree<T, D>(pl, sr, w, ns, L, true, pl.s_keys, pl.s_vals, pl.s_keys_alt, pl.s_vals_alt,
                               pl.s_hist, pl.s_pos, &pl.s_w, pl.s_start, grid_n)))
        return rc;
    t_pos = pl.s_pos.as<T>();
    t_start = pl.s_start.as<unsigned>();
    t_perm = pl.s_vals.as<unsigned>();
    if ((const void *)tg != (const void *)sr || nt != ns) {
        if ((rc = build_tree<T, D>(pl, tg, nullptr, nt, L, false, pl.t_keys, pl.t_vals, pl.t_keys_alt,
                                   pl.t_vals_alt, pl.t_hist, pl.t_pos, nullptr, pl.t_start, grid_n)))
            return rc;
        t_pos = pl.t_pos.as<T>();
        t_start = pl.t_start.as<unsigned>();
        t_perm = pl.t_vals.as<unsigned>();
    }
    return FMMA_OK;
}

// The tree solver (src/tree.cpp:12-90): P2M and M2M as in the fmm passes,
// then M2P from the interaction lists of levels 2..L straight to the
// targets, then P2P; ans is overwritten (src/tree.cpp:64-66).  Stage marks:
// 0 tree, 1 P2M, 2 M2M, 3 M2P, 4 P2P, 5 end.
template <typename T, int D>
static int run_tree(Plan &pl, long nt, const T *tg, const T *w, long ns, const T *sr, T *ans)
{
    pl.view.valid = false;
    pl.last_L = -1;
    const int L = pl.depth - 1;
    if (D * L > 30 || L >= kMaxLevels)
        return fail(FMMA_ERR_UNSUPPORTED, "Depth " + std::to_string(pl.depth) + " too deep for dimension " +
                                              std::to_string(D) + " (at most 2^30 leaf boxes)");
    const int n = pl.poly_ord + 1, P = pl.P();
    if (P > kMaxP) return fail(FMMA_ERR_UNSUPPORTED, "(poly_ord+1)^DIM > 512 is not supported");
    int rc;
    if ((rc = setup_basis(pl))) return rc;
    long off[kMaxLevels + 1], o = 0;
    for (int l = 0; l <= L; ++l) {
        off[l] = o;
        o += (long)P << (D * l);
    }
    if ((rc = pl.Wm.ensure(o * pl.tsize()))) return rc;
    hipStream_t st = pl.stream;
    const bool shared = (const void *)tg == (const void *)sr && nt == ns;
    const unsigned nbox = 1u << (D * L);

    pl.mark(0);
    if ((rc = pl.geom.ensure(sizeof(Geom<T>))) || (rc = pl.bbox_part.ensure(kBBoxParts * kBBoxStride<D> * sizeof(T))))
        return rc;
    if ((rc = bbox<T, D>(tg, nt, sr, ns, pl.bbox_part.as<T>(), pl.geom.as<Geom<T>>(), 1 << L, st))) return rc;
    if ((rc = check_geom<T>(pl))) return rc;
    const T *t_pos;
    const unsigned *t_start, *t_perm;
    if ((rc = sort_points<T, D>(pl, nt, tg, w, ns, sr, L, 0, t_pos, t_start, t_perm))) return rc;
    pl.mark(1);

    T *Wm = pl.Wm.as<T>();
    P2MArgs<T> pa{pl.s_start.as<unsigned>(), pl.s_pos.as<T>(), pl.s_w.as<T>(), ns, pl.geom.as<Geom<T>>(), L,
                  pl.d_coef.as<T>(), n, P, Wm + off[L]};
    if ((rc = p2m<T, D>(pa, nbox, st))) return rc;
    pl.mark(2);
    for (int l = L - 1; l >= 2; --l)
        if ((rc = m2m<T, D>(Wm + off[l + 1], Wm + off[l], 1L << (D * l), pl.d_tab.as<T>(), n, P, st))) return rc;
    pl.mark(3);

    const T *base = nullptr;
    if (L >= 2) {
        if ((rc = pl.l2p_out.ensure((size_t)nt * sizeof(T)))) return rc;
        fmma_m2p_args ma{D, pl.dtype, n, P, pl.d_nodes.as<double>(), (long)nbox, t_start, t_pos, nt,
                         pl.geom.as<Geom<T>>()->origin, &pl.geom.as<Geom<T>>()->Len, L, 0, Wm, pl.l2p_out.p};
        if ((rc = pl.ops.m2p(&ma, pl.ops.ctx, st))) return rc;
        base = pl.l2p_out.as<T>();
    }
    pl.mark(4);
    if ((rc = p2p_stage<T, D>(pl, L, 0, nbox, t_start, t_pos, t_perm, nt, ns, shared, base, ans, 0, -1))) return rc;
    pl.mark(5);
    return FMMA_OK;
}

// The nrnmm solver (src/nrnmm.cpp:13-100): one grid of nrn_N^DIM boxes,
// row-major; P2M on it, M2P from every box at Chebyshev distance >= 2 and
// P2P from the others; ans is overwritten.  Stage marks: 0 tree, 1 P2M,
// 2 M2P, 3 P2P, 4 end.
template <typename T, int D>
static int run_nrnmm(Plan &pl, long nt, const T *tg, const T *w, long ns, const T *sr, T *ans)
{
    pl.view.valid = false;
    pl.last_L = -1;
    const int g = pl.nrn_N;
    long nb = 1;
    for (int d = 0; d < D; ++d) nb *= g;
    if (g < 1 || nb > (1L << 30))
        return fail(FMMA_ERR_UNSUPPORTED, "nrn_N " + std::to_string(g) + " gives more than 2^30 boxes");
    const unsigned nbox = (unsigned)nb;
    const int n = pl.poly_ord + 1, P = pl.P();
    if (P > kMaxP) return fail(FMMA_ERR_UNSUPPORTED, "(poly_ord+1)^DIM > 512 is not supported");
    int rc;
    if ((rc = setup_basis(pl))) return rc;
    if ((rc = pl.Wm.ensure((size_t)nb * P * pl.tsize()))) return rc;
    hipStream_t st = pl.stream;
    const bool shared = (const void *)tg == (const void *)sr && nt == ns;

    pl.mark(0);
    if ((rc = pl.geom.ensure(sizeof(Geom<T>))) || (rc = pl.bbox_part.ensure(kBBoxParts * kBBoxStride<D> * sizeof(T))))
        return rc;
    if ((rc = bbox<T, D>(tg, nt, sr, ns, pl.bbox_part.as<T>(), pl.geom.as<Geom<T>>(), g, st))) return rc;
    if ((rc = check_geom<T>(pl))) return rc;
    const T *t_pos;
    const unsigned *t_start, *t_perm;
    if ((rc = sort_points<T, D>(pl, nt, tg, w, ns, sr, 0, g, t_pos, t_start, t_perm))) return rc;
    pl.mark(1);

    T *Wm = pl.Wm.as<T>();
    P2MArgs<T> pa{pl.s_start.as<unsigned>(), pl.s_pos.as<T>(), pl.s_w.as<T>(), ns, pl.geom.as<Geom<T>>(), 0,
                  pl.d_coef.as<T>(), n, P, Wm, nullptr, nullptr, nullptr, g};
    if ((rc = p2m<T, D>(pa, nbox, st))) return rc;
    pl.mark(2);

    if ((rc = pl.l2p_out.ensure((size_t)nt * sizeof(T)))) return rc;
    fmma_m2p_args ma{D, pl.dtype, n, P, pl.d_nodes.as<double>(), (long)nbox, t_start, t_pos, nt,
                     pl.geom.as<Geom<T>>()->origin, &pl.geom.as<Geom<T>>()->Len, 0, g, Wm, pl.l2p_out.p};
    if ((rc = pl.ops.m2p(&ma, pl.ops.ctx, st))) return rc;
    pl.mark(3);
    if ((rc = p2p_stage<T, D>(pl, 0, g, nbox, t_start, t_pos, t_perm, nt, ns, shared, pl.l2p_out.as<T>(), ans, 0,
                              -1)))
        return rc;
    pl.mark(4);
    return FMMA_OK;
}

enum class Op { Solve, Gradient, Transpose };

// out = the solve's ans [n_t], the gradient [n_t][D], or the transpose [n_s].
template <typename T, int D>
static int op_typed(Plan &pl, Op op, long nt, const void *tg_, const void *w_, long ns, const void *sr_, void *out_,
                    bool is_fmm, bool reuse)
{
    const T *tg = (const T *)tg_, *w = (const T *)w_, *sr = (const T *)sr_;
    T *out = (T *)out_;
    int rc;
    if (op == Op::Solve) {
        if (pl.solver == "tree") return run_tree<T, D>(pl, nt, tg, w, ns, sr, out);
        if (pl.solver == "nrnmm") return run_nrnmm<T, D>(pl, nt, tg, w, ns, sr, out);
        return is_fmm ? run_fmm<T, D>(pl, nt, tg, w, ns, sr, out) : run_exact<T, D>(pl, nt, tg, w, ns, sr, out);
    }
    if (op == Op::Gradient) {
        if (!is_fmm) {
            fmma_exact_grad_args ea{D, pl.dtype, nt, ns, tg, w, sr, out};
            return pl.ops.exact_grad(&ea, pl.ops.ctx, pl.stream);
        }
        const Plan::View &v = pl.view;
        const bool same = reuse && v.valid && v.tg == tg_ && v.w == w_ && v.sr == sr_ && v.nt == nt && v.ns == ns &&
                          v.gen == pl.gen && v.L == pl.depth - 1;
        if (!same && (rc = fmm_passes<T, D>(pl, nt, tg, w, ns, sr))) return rc;
        return fmm_eval_grad<T, D>(pl, out);
    }
    // transpose: targets and sources swap roles, f becomes f^T
    const fmma_kernel_ops saved = pl.ops;
    T sgn = T(1);
    if (saved.symmetry == 1 || saved.symmetry == -1)
        sgn = (T)saved.symmetry;
    else if (saved.transposed)
        pl.ops = *saved.transposed;
    else
        return fail(FMMA_ERR_UNSUPPORTED, "the kernel declares no symmetry and has no transposed launcher table");
    if (is_fmm && pl.depth - 1 < 2 && (rc = scale<T>(out, ns, T(0), pl.stream))) return rc;  // P2P adds onto out
    rc = is_fmm ? run_fmm<T, D>(pl, ns, sr, w, nt, tg, out) : run_exact<T, D>(pl, ns, sr, w, nt, tg, out);
    pl.ops = saved;
    if (!rc && sgn != T(1)) rc = scale<T>(out, ns, sgn, pl.stream);
    return rc;
}

static int run_op(Plan &pl, Op op, long nt, const void *tg, const void *w, long ns, const void *sr, void *out,
                  bool reuse)
{
    const std::string &s = pl.solver;
    const bool is_fmm = s == "fmm", is_tree = s == "tree", is_nrnmm = s == "nrnmm";
    if (!is_fmm && !is_tree && !is_nrnmm && s != "exact" && s != "exact_matvec")
        return fail(FMMA_ERR_SOLVER, "solve type " + s + " not undefined");
    if ((is_tree || is_nrnmm) && op != Op::Solve)
        return fail(FMMA_ERR_UNSUPPORTED, "the gradient and the transpose run with the fmm and exact solvers");
    if ((is_tree || is_nrnmm) && !pl.ops.m2p)
        return fail(FMMA_ERR_UNSUPPORTED, "the kernel's launcher table has no M2P (the tree and nrnmm solvers)");
    if (nt < 0 || ns < 0) return fail(FMMA_ERR_INVALID, "negative size");
    if (is_nrnmm && pl.nrn_N <= 0) {
        // (int)(sqrt(3) M^(1/(2 DIM))) + 1 from the source count in TYPE, kept afterwards (src/nrnmm.cpp:20-22)
        const double m = pl.dtype == FMMA_F64 ? (double)ns : (double)(float)ns;
        pl.nrn_N = (int)((int)(std::sqrt(3.0) * std::pow(m, 1.0 / (2.0 * pl.dim))) + 1.0);
    }
    if ((is_fmm || is_tree) && pl.depth <= 0) {
        // auto Depth from the target count, kept afterwards (src/fmm.cpp:23-28); the
        // transpose and the gradient use the Depth of the solve they belong to
        pl.depth = (int)(std::log((double)nt) / pl.dim) + 1;
        if (pl.depth <= 0) pl.depth = 1;
    }
    if (op == Op::Gradient && (!pl.ops.p2p_grad || !pl.ops.exact_grad))
        return fail(FMMA_ERR_UNSUPPORTED, "the kernel has no gradient");
    if (nt == 0 || ns == 0) {
        // an empty sum: the gradient and the transpose are zero (the solve leaves ans alone)
        if (op == Op::Solve) return FMMA_OK;
        const long m = op == Op::Gradient ? nt * pl.dim : ns;
        if (m == 0) return FMMA_OK;
        if (!out) return fail(FMMA_ERR_INVALID, "null pointer");
        const int rc = pl.dtype == FMMA_F64 ? scale<double>((double *)out, m, 0.0, pl.stream)
                                            : scale<float>((float *)out, m, 0.f, pl.stream);
        if (rc) return rc;
        FMMA_HIP(hipStreamSynchronize(pl.stream));
        return FMMA_OK;
    }
    if (!tg || !sr || !out) return fail(FMMA_ERR_INVALID, "null pointer");
    if (!pl.ops.p2p || pl.ops.dim != pl.dim || pl.ops.dtype != pl.dtype)
        return fail(FMMA_ERR_INVALID, "kernel launchers do not match the plan's DIM/TYPE");
    if (pl.timing && !pl.events_ready) {
        for (auto &e : pl.ev) FMMA_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        pl.events_ready = true;
    }
    // the fmm, tree and nrnmm solves record their own stage marks
    const bool staged = (is_fmm || is_tree || is_nrnmm) && op == Op::Solve;
    const char *const *names = is_tree ? kTreeStageNames : is_nrnmm ? kNrnStageNames : kStageNames;
    const int nst = !staged ? kStages : is_tree ? 5 : is_nrnmm ? 4 : kStages;
    if (!staged) pl.mark(0);
    int rc;
    if (pl.dtype == FMMA_F64) {
        if (pl.dim == 1) rc = op_typed<double, 1>(pl, op, nt, tg, w, ns, sr, out, is_fmm, reuse);
        else if (pl.dim == 2) rc = op_typed<double, 2>(pl, op, nt, tg, w, ns, sr, out, is_fmm, reuse);
        else rc = op_typed<double, 3>(pl, op, nt, tg, w, ns, sr, out, is_fmm, reuse);
    } else {
        if (pl.dim == 1) rc = op_typed<float, 1>(pl, op, nt, tg, w, ns, sr, out, is_fmm, reuse);
        else if (pl.dim == 2) rc = op_typed<float, 2>(pl, op, nt, tg, w, ns, sr, out, is_fmm, reuse);
        else rc = op_typed<float, 3>(pl, op, nt, tg, w, ns, sr, out, is_fmm, reuse);
    }
    if (rc) return rc;
    if (!staged) pl.mark(kStages);
    FMMA_HIP(hipStreamSynchronize(pl.stream));
    if (pl.timing && pl.events_ready) {
        float total = 0.f;
        FMMA_HIP(hipEventElapsedTime(&total, pl.ev[0], pl.ev[nst]));
        if (staged) {
            for (int i = 0; i < nst; ++i) {
                float ms = 0.f;
                FMMA_HIP(hipEventElapsedTime(&ms, pl.ev[i], pl.ev[i + 1]));
                pl.time_log[names[i]] += ms;
            }
        }
        const char *what = op == Op::Solve ? "" : op == Op::Gradient ? " gradient" : " transpose";
        pl.time_log["total time(" + s + what + ")"] = total;
    }
    return FMMA_OK;
}

// ------------------------------------------------------------------ batches
//
// fmma_plan_solve_batch: many independent problems of one configuration in
// one pass of batched kernels.  Problem p is solved exactly as
// fmma_plan_solve would solve it alone (its own bounding cube and Depth, the
// auto rule applied to its own target count), so results agree to the last
// bit with a loop of single solves.  Storage: the leaf multipoles/locals of
// all problems sit in global leaf-box order (rows of P), then the coarser
// levels of each problem; M2M/L2L run over item lists (child block, parent)
// and M2L over tasks that carry their problem's row and operator offsets.

// Tile width by rule: the narrowest pipelined width that covers a level's
// parents, at most 64.
static int rule_tb(int D, int KP, int mode, long npar)
{
    for (int tb : {16, 32, 64})
        if (tb >= npar && m2l_pipe_supported(D, KP, mode, tb)) return tb;
    return 64;
}

template <int D>
static int batch_layout(Plan &pl, int nprob, const long *tc, const long *sc, bool shared)
{
    BatchPlan *bp = pl.batch.get();
    const int mode = pl.mode(), P = pl.P();
    if (bp && bp->gen == pl.gen && bp->mode == mode && bp->shared == shared && bp->nprob == nprob &&
        std::equal(tc, tc + nprob, bp->tc.begin()) && std::equal(sc, sc + nprob, bp->sc.begin()))
        return FMMA_OK;
    pl.batch.reset(new BatchPlan());
    bp = pl.batch.get();
    bp->tc.assign(tc, tc + nprob);
    bp->sc.assign(sc, sc + nprob);
    bp->shared = shared;
    bp->gen = pl.gen;
    bp->mode = mode;
    bp->nprob = nprob;
    std::vector<long> toff(nprob + 1, 0), soff(nprob + 1, 0);
    for (int p = 0; p < nprob; ++p) {
        toff[p + 1] = toff[p] + tc[p];
        soff[p + 1] = soff[p] + sc[p];
    }
    bp->nt = toff[nprob];
    bp->ns = soff[nprob];
    bp->L.resize(nprob);
    bp->bbase.assign(nprob + 1, 0);
    long nb = 0;
    for (int p = 0; p < nprob; ++p) {
        int depth = pl.depth > 0 ? pl.depth : (int)(std::log((double)tc[p]) / D) + 1;
        if (depth <= 0) depth = 1;
        const int L = depth - 1;
        if (L < 2 || D * L > 30 || L >= kMaxLevels)
            return fail(FMMA_ERR_UNSUPPORTED, "batch problem " + std::to_string(p) + ": Depth " +
                                                  std::to_string(depth) + "; a batch needs 3 <= Depth and at most "
                                                  "2^30 leaf boxes per problem");
        bp->L[p] = L;
        bp->bbase[p] = (unsigned)nb;
        nb += 1L << (D * L);
        bp->Lmax = std::max(bp->Lmax, L);
        if (nb >= (1L << 31)) return fail(FMMA_ERR_UNSUPPORTED, "batch: more than 2^31 leaf boxes");
    }
    bp->bbase[nprob] = (unsigned)nb;
    bp->NB = nb;
    while ((1L << bp->bits) < nb) ++bp->bits;
    bp->row_off.assign((size_t)nprob * kMaxLevels, -1);
    long next = nb;
    for (int p = 0; p < nprob; ++p) {
        bp->row_off[(size_t)p * kMaxLevels + bp->L[p]] = bp->bbase[p];
        for (int l = 2; l < bp->L[p]; ++l) {
            bp->row_off[(size_t)p * kMaxLevels + l] = next;
            next += 1L << (D * l);
        }
    }
    bp->rows = next;
    // M2M steps (parent level L_p - s), then L2L steps (parent level 2 + t)
    std::vector<long> items;
    auto add = [&](int p, int l) {
        const long nq = 1L << (D * l);
        const long crow = bp->row_off[(size_t)p * kMaxLevels + l + 1], prow = bp->row_off[(size_t)p * kMaxLevels + l];
        for (long q = 0; q < nq; ++q) {
            items.push_back((crow + (q << D)) * P);
            items.push_back((prow + q) * P);
        }
    };
    for (int st = 1; st <= bp->Lmax - 2; ++st) {
        bp->step_first.push_back((long)items.size() / 2);
        for (int p = 0; p < nprob; ++p)
            if (bp->L[p] - st >= 2) add(p, bp->L[p] - st);
        bp->step_count.push_back((long)items.size() / 2 - bp->step_first.back());
    }
    bp->n_m2m_steps = (int)bp->step_first.size();
    for (int l = 2; l + 1 <= bp->Lmax; ++l) {
        bp->step_first.push_back((long)items.size() / 2);
        for (int p = 0; p < nprob; ++p)
            if (l + 1 <= bp->L[p]) add(p, l);
        bp->step_count.push_back((long)items.size() / 2 - bp->step_first.back());
    }
    for (int p = 0; p < nprob; ++p)
        for (int l = 2; l <= bp->L[p]; ++l) {
            bp->ent_level.push_back(l);
            bp->ent_prob.push_back(p);
        }
    std::vector<unsigned> box_prob((size_t)nb);
    for (int p = 0; p < nprob; ++p)
        std::fill(box_prob.begin() + bp->bbase[p], box_prob.begin() + bp->bbase[p + 1], (unsigned)p);
    hipStream_t st = pl.stream;
    int rc;
    if ((rc = upload(bp->d_toff, toff.data(), toff.size() * sizeof(long), st)) ||
        (rc = upload(bp->d_soff, soff.data(), soff.size() * sizeof(long), st)) ||
        (rc = upload(bp->d_levels, bp->L.data(), bp->L.size() * sizeof(int), st)) ||
        (rc = upload(bp->d_bbase, bp->bbase.data(), bp->bbase.size() * sizeof(unsigned), st)) ||
        (rc = upload(bp->d_box_prob, box_prob.data(), box_prob.size() * sizeof(unsigned), st)) ||
        (rc = upload(bp->d_items, items.data(), items.size() * sizeof(long), st)) ||
        (rc = upload(bp->d_ent_level, bp->ent_level.data(), bp->ent_level.size() * sizeof(int), st)) ||
        (rc = upload(bp->d_ent_prob, bp->ent_prob.data(), bp->ent_prob.size() * sizeof(int), st)))
        return rc;
    FMMA_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
    return FMMA_OK;
}

// M2L tasks of the batch; tb_of maps parents-per-level to tile width.  With
// only_npar >= 0, just the tasks of levels with that many parents.
static std::vector<M2LTask> batch_tasks(const Plan &pl, const BatchPlan &bp, int D, std::vector<M2LGroup> &groups,
                                        const std::map<long, int> &tb_of, long only_npar = -1)
{
    const int P = pl.P(), mode = pl.mode();
    int MP, KP;
    m2l_shape(P, mode, MP, KP);
    const long fe = (long)pl.offs.n_far * MP * KP;
    std::map<int, std::vector<M2LTask>> by_tb;
    for (size_t e = 0; e < bp.ent_level.size(); ++e) {
        const int l = bp.ent_level[e], p = bp.ent_prob[e];
        const long npar = 1L << (D * (l - 1));
        if (only_npar >= 0 && npar != only_npar) continue;
        const int tb = tb_of.at(npar);
        const long row = bp.row_off[(size_t)p * kMaxLevels + l];
        for (long p0 = 0; p0 < npar; p0 += tb)
            for (int c = 0; c < (1 << D); ++c)
                by_tb[tb].push_back(M2LTask{l, c, (unsigned)p0, (int)std::min<long>(tb, npar - p0), row * P, row * KP,
                                            (long)e * fe});
    }
    groups.clear();
    std::vector<M2LTask> all;
    for (auto &kv : by_tb) {
        groups.push_back(M2LGroup{kv.first, (int)all.size(), (int)kv.second.size(),
                                  m2l_pipe_supported(D, KP, mode, kv.first)});
        all.insert(all.end(), kv.second.begin(), kv.second.end());
    }
    return all;
}

template <typename T, int D>
static int batch_buffers(Plan &pl, BatchPlan &bp)
{
    const int P = pl.P(), mode = pl.mode();
    int MP, KP;
    m2l_shape(P, mode, MP, KP);
    const size_t ts = sizeof(T);
    int rc;
    auto tree_bufs = [&](long n, DBuf &k, DBuf &v, DBuf &ka, DBuf &va, DBuf &h, DBuf &pos, DBuf *wb,
                         DBuf &start) -> int {
        int r;
        if ((r = k.ensure(n * 4)) || (r = v.ensure(n * 4)) || (r = ka.ensure(n * 4)) || (r = va.ensure(n * 4)) ||
            (r = h.ensure(radix_hist_words(n) * 4)) || (r = pos.ensure((size_t)n * D * ts)) ||
            (r = start.ensure(((size_t)bp.NB + 1) * 4)))
            return r;
        if (wb && (r = wb->ensure((size_t)n * ts))) return r;
        return FMMA_OK;
    };
    if ((rc = tree_bufs(bp.ns, pl.s_keys, pl.s_vals, pl.s_keys_alt, pl.s_vals_alt, pl.s_hist, pl.s_pos, &pl.s_w,
                        pl.s_start)))
        return rc;
    if (!bp.shared && (rc = tree_bufs(bp.nt, pl.t_keys, pl.t_vals, pl.t_keys_alt, pl.t_vals_alt, pl.t_hist, pl.t_pos,
                                      nullptr, pl.t_start)))
        return rc;
    if ((rc = pl.pos_alt.ensure((size_t)(D + 1) * std::max(bp.ns, bp.nt) * ts))) return rc;
    const size_t opsz = mode == 0 ? ts : mode == 1 ? 4 : 2;
    const size_t esz = mode == 0 ? 8 : mode == 1 ? 4 : 2;
    const size_t fe = (size_t)pl.offs.n_far * MP * KP;
    constexpr int NU = ((D == 1 ? 3 : D == 2 ? 9 : 27) - 1) / 2;
    if ((rc = bp.d_geoms.ensure((size_t)bp.nprob * sizeof(Geom<T>))) || (rc = pl.Wm.ensure(bp.rows * P * ts)) ||
        (rc = pl.Wl.ensure(bp.rows * P * ts)) || (rc = pl.Wc.ensure(bp.rows * KP * opsz)) ||
        (mode == 2 && (rc = pl.Wc_lo.ensure(bp.rows * KP * 2))) ||
        (rc = pl.Kf.ensure(bp.ent_level.size() * fe * esz)) ||
        (mode == 2 && (rc = pl.Kf_lo.ensure(bp.ent_level.size() * fe * esz))) ||
        (rc = pl.l2p_out.ensure((size_t)bp.nt * ts)) || (rc = pl.p2p_slots.ensure((size_t)(NU + 1) * bp.nt * ts)))
        return rc;
    return FMMA_OK;
}

// Sort one point set of the batch into its global leaf boxes.
template <typename T, int D>
static int batch_tree(Plan &pl, BatchPlan &bp, const T *pts, const T *w, long n, const long *d_off, bool with_w,
                      DBuf &keys, DBuf &vals, DBuf &keys_alt, DBuf &vals_alt, DBuf &hist, DBuf &pos, DBuf *wbuf,
                      DBuf &start, hipStream_t st)
{
    int rc;
    if ((rc = batch_keys<T, D>(pts, n, d_off, bp.nprob, bp.d_geoms.as<Geom<T>>(), bp.d_levels.as<int>(),
                               bp.d_bbase.as<unsigned>(), keys.as<unsigned>(), st)))
        return rc;
    PointSortBuffers<T> sb{keys.as<unsigned>(), vals.as<unsigned>(), keys_alt.as<unsigned>(),
                           vals_alt.as<unsigned>(), hist.as<unsigned>(), pos.as<T>(),
                           with_w ? wbuf->as<T>() : nullptr, pl.pos_alt.as<T>(), pl.pos_alt.as<T>() + (size_t)D * n};
    if ((rc = sort_points<T, D>(sb, n, bp.bits, pts, w, with_w, st))) return rc;
    return box_ranges(keys.as<unsigned>(), n, (unsigned)bp.NB, start.as<unsigned>(), st);
}

// The batched solve as a sequence of kernels on st (capturable: no
// allocation, copy from the host or synchronisation, except while tuning).
template <typename T, int D>
static int batch_kernels(Plan &pl, BatchPlan &bp, const T *tg, const T *w, const T *sr, T *ans, hipStream_t st,
                         bool eager, bool tune)
{
    const int n = pl.poly_ord + 1, P = pl.P(), mode = pl.mode();
    int rc;
    auto mark = [&](int i) {
        if (eager) pl.mark(i);
    };
    Geom<T> *geoms = bp.d_geoms.as<Geom<T>>();
    mark(0);
    if ((rc = batch_bbox<T, D>(tg, bp.d_toff.as<long>(), bp.shared ? nullptr : sr, bp.d_soff.as<long>(),
                               bp.d_levels.as<int>(), bp.nprob, geoms, st)))
        return rc;
    if ((rc = batch_tree<T, D>(pl, bp, sr, w, bp.ns, bp.d_soff.as<long>(), true, pl.s_keys, pl.s_vals, pl.s_keys_alt,
                               pl.s_vals_alt, pl.s_hist, pl.s_pos, &pl.s_w, pl.s_start, st)))
        return rc;
    const T *t_pos = pl.s_pos.as<T>();
    const unsigned *t_start = pl.s_start.as<unsigned>(), *t_perm = pl.s_vals.as<unsigned>(),
                   *t_keys = pl.s_keys.as<unsigned>();
    if (!bp.shared) {
        if ((rc = batch_tree<T, D>(pl, bp, tg, nullptr, bp.nt, bp.d_toff.as<long>(), false, pl.t_keys, pl.t_vals,
                                   pl.t_keys_alt, pl.t_vals_alt, pl.t_hist, pl.t_pos, nullptr, pl.t_start, st)))
            return rc;
        t_pos = pl.t_pos.as<T>();
        t_start = pl.t_start.as<unsigned>();
        t_perm = pl.t_vals.as<unsigned>();
        t_keys = pl.t_keys.as<unsigned>();
    }
    mark(1);
    T *Wm = pl.Wm.as<T>(), *Wl = pl.Wl.as<T>();
    const unsigned *box_prob = bp.d_box_prob.as<unsigned>();
    P2MArgs<T> pa{pl.s_start.as<unsigned>(), pl.s_pos.as<T>(), pl.s_w.as<T>(), bp.ns, nullptr, 0,
                  pl.d_coef.as<T>(), n, P, Wm, box_prob, geoms, bp.d_levels.as<int>()};
    if ((rc = p2m<T, D>(pa, bp.NB, st))) return rc;
    mark(2);
    const long *items = bp.d_items.as<long>();
    for (int s = 0; s < bp.n_m2m_steps; ++s)
        if ((rc = m2m<T, D>(Wm, Wm, bp.step_count[s], pl.d_tab.as<T>(), n, P, st, items + 2 * bp.step_first[s])))
            return rc;
    mark(3);
    // M2L: every (problem, level) operator set straight into fragment order, then the contraction
    fmma_m2l_op_args oa{};
    oa.dim = D;
    oa.dtype = pl.dtype;
    oa.n_nodes = n;
    oa.P = P;
    oa.nodes = pl.d_nodes.as<double>();
    oa.Len = &geoms->Len;
    oa.len_stride = (int)(sizeof(Geom<T>) / sizeof(T));
    oa.offsets = pl.d_far.as<int>();
    oa.n_off = pl.offs.n_far;
    oa.format = m2l_frag_format(P, mode);
    m2l_shape(P, mode, oa.MP, oa.KP);
    oa.n_batch = (int)bp.ent_level.size();
    oa.batch_level = bp.d_ent_level.as<int>();
    oa.batch_problem = bp.d_ent_prob.as<int>();
    oa.Kf = pl.Kf.p;
    oa.Kf_lo = pl.Kf_lo.p;
    oa.pairs = pl.d_pairs.as<int>();
    oa.n_pairs = pl.n_pairs;
    if ((rc = pl.ops.m2l_op(&oa, pl.ops.ctx, st)))
        return rc == FMMA_ERR_UNSUPPORTED ? fail(rc, "batch: the kernel's launcher table has no fragment form") : rc;
    if ((rc = m2l_convert<T>(Wm, bp.rows, P, oa.KP, mode, pl.Wc.p, pl.Wc_lo.p, st))) return rc;
    M2LKArgs ka{};
    ka.cls_f = pl.d_cls_f.as<int>();
    ka.cls_o = pl.d_cls_o.as<int>();
    ka.n_cls = pl.offs.n_cls;
    ka.P = P;
    m2l_shape(P, mode, ka.MP, ka.KP);
    ka.Wm = Wm;
    ka.Wl = Wl;
    ka.Kf = pl.Kf.p;
    ka.Kf_lo = pl.Kf_lo.p;
    ka.Wc = pl.Wc.p;
    ka.Wc_lo = pl.Wc_lo.p;
    if (tune) {
        // the autotuner: for each number of parents per level, time every
        // pipelined tile width on this batch's own tasks and keep the fastest
        std::map<long, int> choice;
        for (int l = 2; l <= bp.Lmax; ++l) {
            const long npar = 1L << (D * (l - 1));
            const long key = tune_key(D, P, mode, npar);
            if (pl.tuned_tb.count(key)) {
                choice[npar] = pl.tuned_tb[key];
                continue;
            }
            float best = 1e30f;
            int best_tb = rule_tb(D, ka.KP, mode, npar);
            for (int tb : {16, 32, 64, 128, 256}) {
                if (!m2l_pipe_supported(D, ka.KP, mode, tb) || tb > 4 * npar) continue;
                std::map<long, int> one{{npar, tb}};
                std::vector<M2LGroup> g;
                std::vector<M2LTask> t = batch_tasks(pl, bp, D, g, one, npar);
                if (t.empty()) break;
                DBuf dt;
                if ((rc = upload(dt, t.data(), t.size() * sizeof(M2LTask), st))) return rc;
                M2LKArgs a2 = ka;
                a2.tasks = dt.as<M2LTask>();
                if ((rc = m2l_apply<T, D>(a2, g, mode, st))) return rc;  // warm-up
                hipEvent_t e0, e1;
                FMMA_HIP(hipEventCreate(&e0));
                FMMA_HIP(hipEventCreate(&e1));
                FMMA_HIP(hipEventRecord(e0, st));
                for (int r = 0; r < 3 && !rc; ++r) rc = m2l_apply<T, D>(a2, g, mode, st);
                FMMA_HIP(hipEventRecord(e1, st));
                FMMA_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                FMMA_HIP(hipEventElapsedTime(&ms, e0, e1));
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                if (rc) return rc;
                if (ms < best) {
                    best = ms;
                    best_tb = tb;
                }
            }
            pl.tuned_tb[key] = best_tb;
            choice[npar] = best_tb;
        }
        std::vector<M2LTask> tasks = batch_tasks(pl, bp, D, bp.groups, choice);
        if ((rc = upload(bp.d_tasks, tasks.data(), tasks.size() * sizeof(M2LTask), st))) return rc;
        FMMA_HIP(hipStreamSynchronize(st));
        bp.tuned = true;
    }
    ka.tasks = bp.d_tasks.as<M2LTask>();
    if ((rc = m2l_apply<T, D>(ka, bp.groups, mode, st))) return rc;
    mark(4);
    for (size_t s = bp.n_m2m_steps; s < bp.step_first.size(); ++s)
        if ((rc = l2l<T, D>(Wl, Wl, bp.step_count[s], pl.d_tab.as<T>(), n, P, st, items + 2 * bp.step_first[s])))
            return rc;
    mark(5);
    L2PArgs<T> la{t_start, t_keys, t_pos, bp.nt, nullptr, 0, pl.d_coef.as<T>(), n, P, Wl, pl.l2p_out.as<T>(),
                  box_prob, geoms, bp.d_levels.as<int>()};
    if ((rc = l2p<T, D>(la, bp.NB, st))) return rc;
    mark(6);
    const bool mutual = bp.shared && pl.p2p_mutual && (pl.ops.symmetry == 1 || pl.ops.symmetry == -1);
    fmma_p2p_args qa{};
    qa.dim = D;
    qa.dtype = pl.dtype;
    qa.n_target_boxes = bp.NB;
    qa.leaf_level = 0;
    qa.t_start = t_start;
    qa.s_start = pl.s_start.as<unsigned>();
    qa.t_pos = t_pos;
    qa.s_pos = pl.s_pos.as<T>();
    qa.s_w = pl.s_w.as<T>();
    qa.n_t = bp.nt;
    qa.n_s = bp.ns;
    qa.base = pl.l2p_out.as<T>();
    qa.t_perm = t_perm;
    qa.ans = ans;
    qa.accumulate = 0;
    qa.slots = mutual ? pl.p2p_slots.p : nullptr;
    qa.symmetry = mutual ? pl.ops.symmetry : 0;
    qa.box_problem = box_prob;
    qa.problem_box_base = bp.d_bbase.as<unsigned>();
    qa.problem_leaf_level = bp.d_levels.as<int>();
    if ((rc = pl.ops.p2p(&qa, pl.ops.ctx, st))) return rc;
    mark(7);
    if (mutual) {
        P2PCombineArgs<T> ca{pl.p2p_slots.as<T>(), bp.nt, pl.l2p_out.as<T>(), t_perm, ans, 0};
        if ((rc = p2p_combine<T, D>(ca, st))) return rc;
    }
    mark(8);
    return FMMA_OK;
}

static std::vector<const void *> batch_buffer_ptrs(const Plan &pl, const void *tg, const void *w, const void *sr,
                                                   const void *ans)
{
    const BatchPlan &b = *pl.batch;
    return {tg, w, sr, ans, pl.s_keys.p, pl.s_vals.p, pl.s_keys_alt.p, pl.s_vals_alt.p, pl.s_hist.p, pl.s_pos.p,
            pl.s_w.p, pl.s_start.p, pl.t_keys.p, pl.t_vals.p, pl.t_keys_alt.p, pl.t_vals_alt.p, pl.t_hist.p,
            pl.t_pos.p, pl.t_start.p, pl.Wm.p, pl.Wl.p, pl.Wc.p, pl.Wc_lo.p, pl.Kf.p, pl.Kf_lo.p, pl.l2p_out.p,
            pl.p2p_slots.p, pl.d_nodes.p, pl.d_coef.p, pl.d_tab.p, pl.d_far.p, pl.d_cls_f.p, pl.d_cls_o.p,
            pl.d_pairs.p, pl.pos_alt.p,
            b.d_toff.p, b.d_soff.p, b.d_levels.p, b.d_bbase.p, b.d_box_prob.p, b.d_geoms.p, b.d_items.p,
            b.d_tasks.p, b.d_ent_level.p, b.d_ent_prob.p, pl.ops.ctx};
}

// The batch entry: layout, buffers, then an eager pass (which also runs the
// autotuner the first time) or a replay of the captured graph.
template <typename T, int D>
static int batch_typed(Plan &pl, int nprob, const long *tc, const void *tg_, const void *w_, const long *sc,
                       const void *sr_, void *ans_)
{
    const T *tg = (const T *)tg_, *w = (const T *)w_, *sr = (const T *)sr_;
    T *ans = (T *)ans_;
    const bool shared = tg_ == sr_ && std::equal(tc, tc + nprob, sc);
    int rc;
    if ((rc = setup_basis(pl))) return rc;
    if ((rc = batch_layout<D>(pl, nprob, tc, sc, shared))) return rc;
    BatchPlan &bp = *pl.batch;
    if ((rc = batch_buffers<T, D>(pl, bp))) return rc;
    if (!bp.tuned) {
        if (!pl.m2l_autotune || pl.forced_tb > 0) {
            std::map<long, int> choice;
            int MP, KP;
            m2l_shape(pl.P(), pl.mode(), MP, KP);
            for (int l = 2; l <= bp.Lmax; ++l) {
                const long npar = 1L << (D * (l - 1));
                const bool forced = pl.forced_tb > 0 && m2l_pipe_supported(D, KP, pl.mode(), pl.forced_tb);
                choice[npar] = forced ? pl.forced_tb : rule_tb(D, KP, pl.mode(), npar);
            }
            std::vector<M2LTask> tasks = batch_tasks(pl, bp, D, bp.groups, choice);
            if ((rc = upload(bp.d_tasks, tasks.data(), tasks.size() * sizeof(M2LTask), pl.stream))) return rc;
            FMMA_HIP(hipStreamSynchronize(pl.stream));
            bp.tuned = true;
        } else {
            pl.batch_eager = true;
            return batch_kernels<T, D>(pl, bp, tg, w, sr, ans, pl.stream, true, true);
        }
    }
    pl.batch_eager = !pl.batch_graph;
    if (!pl.batch_graph) return batch_kernels<T, D>(pl, bp, tg, w, sr, ans, pl.stream, true, false);
    // the graph holds every buffer address: capture again when any moved
    const std::vector<const void *> bufs = batch_buffer_ptrs(pl, tg_, w_, sr_, ans_);
    if (!bp.exec || bufs != bp.g_bufs) {
        bp.drop_graph();
        if (!pl.cap_stream) FMMA_HIP(hipStreamCreateWithFlags(&pl.cap_stream, hipStreamNonBlocking));
        FMMA_HIP(hipStreamBeginCapture(pl.cap_stream, hipStreamCaptureModeThreadLocal));
        rc = batch_kernels<T, D>(pl, bp, tg, w, sr, ans, pl.cap_stream, false, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(pl.cap_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "capture of the batched solve");
        bp.graph = g;
        FMMA_HIP(hipGraphInstantiateWithFlags(&bp.exec, bp.graph, 0));
        bp.g_bufs = bufs;
    }
    pl.mark(0);
    FMMA_HIP(hipGraphLaunch(bp.exec, pl.stream));
    pl.mark(kStages);
    return FMMA_OK;
}

static int run_batch(Plan &pl, int nprob, const long *tc, const void *tg, const void *w, const long *sc,
                     const void *sr, void *ans)
{
    if (pl.solver != "fmm") return fail(FMMA_ERR_UNSUPPORTED, "batches run the fmm solver only");
    if (nprob <= 0) return FMMA_OK;
    if (!tc || !sc || !tg || !sr || !ans) return fail(FMMA_ERR_INVALID, "null pointer");
    for (int p = 0; p < nprob; ++p)
        if (tc[p] <= 0 || sc[p] <= 0) return fail(FMMA_ERR_INVALID, "every problem of a batch needs points");
    if (!(pl.trans_sym == 1 || pl.ops.translation_invariant))
        return fail(FMMA_ERR_UNSUPPORTED, "batches need per-offset M2L operators (trans_sym_flag 1 or a "
                                          "translation-invariant kernel)");
    if (!pl.ops.p2p || pl.ops.dim != pl.dim || pl.ops.dtype != pl.dtype)
        return fail(FMMA_ERR_INVALID, "kernel launchers do not match the plan's DIM/TYPE");
    if (pl.timing && !pl.events_ready) {
        for (auto &e : pl.ev) FMMA_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        pl.events_ready = true;
    }
    pl.view.valid = false;  // the batch reuses the tree buffers of single solves
    pl.last_L = -1;
    int rc;
    if (pl.dtype == FMMA_F64) {
        if (pl.dim == 1) rc = batch_typed<double, 1>(pl, nprob, tc, tg, w, sc, sr, ans);
        else if (pl.dim == 2) rc = batch_typed<double, 2>(pl, nprob, tc, tg, w, sc, sr, ans);
        else rc = batch_typed<double, 3>(pl, nprob, tc, tg, w, sc, sr, ans);
    } else {
        if (pl.dim == 1) rc = batch_typed<float, 1>(pl, nprob, tc, tg, w, sc, sr, ans);
        else if (pl.dim == 2) rc = batch_typed<float, 2>(pl, nprob, tc, tg, w, sc, sr, ans);
        else rc = batch_typed<float, 3>(pl, nprob, tc, tg, w, sc, sr, ans);
    }
    if (rc) return rc;
    FMMA_HIP(hipStreamSynchronize(pl.stream));
    // The batch runs as one captured sequence, so the cubes are checked after
    // it: a problem the reference cannot bin fails the call (ans is then
    // undefined; the kernels clamp box coordinates and stay in bounds).
    {
        const size_t gsz = pl.dtype == FMMA_F64 ? sizeof(Geom<double>) : sizeof(Geom<float>);
        const size_t soff = pl.dtype == FMMA_F64 ? offsetof(Geom<double>, status) : offsetof(Geom<float>, status);
        std::vector<unsigned char> hg((size_t)nprob * gsz);
        FMMA_HIP(hipMemcpy(hg.data(), pl.batch->d_geoms.p, hg.size(), hipMemcpyDeviceToHost));
        for (int p = 0; p < nprob; ++p) {
            int status;
            std::memcpy(&status, hg.data() + (size_t)p * gsz + soff, sizeof(int));
            if ((rc = geom_status_error(status)))
                return fail(rc, "batch problem " + std::to_string(p) + ": " + g_err);
        }
    }
    if (pl.timing && pl.events_ready) {
        float total = 0.f;
        FMMA_HIP(hipEventElapsedTime(&total, pl.ev[0], pl.ev[kStages]));
        pl.time_log["total time(fmm batch)"] = total;
        if (pl.batch_eager) {  // an eager pass marked its stages (a graph replay does not)
            for (int i = 0; i < kStages; ++i) {
                float ms = 0.f;
                FMMA_HIP(hipEventElapsedTime(&ms, pl.ev[i], pl.ev[i + 1]));
                pl.time_log[std::string("batch ") + kStageNames[i]] += ms;
            }
        }
    }
    return FMMA_OK;
}

static int solve_device(Plan &pl, long nt, const void *tg, const void *w, long ns, const void *sr, void *ans)
{
    return run_op(pl, Op::Solve, nt, tg, w, ns, sr, ans, false);
}

// This GPU's scratch address, from hipGetSymbolAddress on the first no-alloc
// call with that device current.
static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_ptr;  // per device

static int scratch_ptr(void **out)
{
    int dev = 0;
    FMMA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        FMMA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(FMMA_ERR_HIP, "hipGetSymbolAddress gave a null address for the no-alloc scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FMMA_OK;
}

struct ArenaScope {
    explicit ArenaScope(Arena *a) { t_arena = a; }
    ~ArenaScope() { t_arena = nullptr; }
};

// fmma_plan_solve with every device buffer carved from the scratch: a
// private plan with the caller's settings (kernel, solver, order, Depth,
// trans_sym, M2L form, stream) runs the solve, with the rule's M2L tile
// widths (or the forced one) instead of the autotuner and without the stage
// timing, so that it creates no events and allocates no device memory.  The
// auto Depth / nrn_N it picks is kept by the caller's plan, as a solve would.
static int solve_noalloc(Plan &user, long nt, const void *tg, const void *w, long ns, const void *sr, void *ans)
{
    void *base = nullptr;
    int rc;
    if ((rc = scratch_ptr(&base))) return rc;
    std::unique_ptr<Plan> tmp(new Plan());
    Plan &pl = *tmp;
    pl.dim = user.dim;
    pl.dtype = user.dtype;
    pl.solver = user.solver;
    pl.poly_ord = user.poly_ord;
    pl.depth = user.depth;
    pl.trans_sym = user.trans_sym;
    pl.nrn_N = user.nrn_N;
    pl.kernel_id = user.kernel_id;
    pl.kparams = user.kparams;
    pl.ops = user.ops;  // a built-in's ctx points at user.kparams, alive for this call
    pl.precision = user.precision;
    pl.stream = user.stream;
    pl.p2p_mutual = user.p2p_mutual;
    pl.forced_tb = user.forced_tb;
    pl.timing = false;
    pl.m2l_autotune = false;
    Arena ar{static_cast<char *>(base), kNoAllocScratchBytes, 0};
    {
        ArenaScope scope(&ar);
        rc = run_op(pl, Op::Solve, nt, tg, w, ns, sr, ans, false);
    }
    if (rc) (void)hipStreamSynchronize(pl.stream);  // no kernel of this call still uses the scratch
    if (user.depth <= 0 && pl.depth > 0 && (pl.solver == "fmm" || pl.solver == "tree")) user.depth = pl.depth;
    if (user.nrn_N <= 0 && pl.nrn_N > 0 && pl.solver == "nrnmm") user.nrn_N = pl.nrn_N;
    return rc;
}

}  // namespace lib
}  // namespace fmma

using fmma::lib::Plan;

struct fmma_plan_s {
    Plan pl;
};

#define FMMA_GUARD(...)                                                          \
    try {                                                                        \
        __VA_ARGS__;                                                             \
    } catch (const std::bad_alloc &) {                                           \
        return fmma::lib::fail_oom();                                            \
    } catch (...) {                                                              \
        fmma::lib::set_error("unexpected C++ exception");                         \
        return FMMA_ERR_INVALID;                                                 \
    }

namespace fmma {
namespace lib {
int fail_oom() { return fail(FMMA_ERR_INVALID, "host allocation failed"); }
}  // namespace lib
}  // namespace fmma

extern "C" {

int fmma_plan_create(fmma_plan *plan, int dim, int dtype)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan pointer");
    *plan = nullptr;
    if (dim < 1 || dim > 3) return fmma::lib::fail(FMMA_ERR_INVALID, "DIM must be 1, 2 or 3");
    if (dtype != FMMA_F64 && dtype != FMMA_F32) return fmma::lib::fail(FMMA_ERR_INVALID, "dtype must be F64 or F32");
    FMMA_GUARD({
        fmma_plan p = new fmma_plan_s();
        p->pl.dim = dim;
        p->pl.dtype = dtype;
        fmma::lib::builtin_kernel_ops(FMMA_KERNEL_INV_R, dim, dtype, &p->pl.kparams, &p->pl.ops);
        *plan = p;
    });
    return FMMA_OK;
}

int fmma_plan_destroy(fmma_plan plan)
{
    delete plan;
    return FMMA_OK;
}

int fmma_plan_set_solver_type(fmma_plan plan, const char *type)
{
    if (!plan || !type) return fmma::lib::fail(FMMA_ERR_INVALID, "null argument");
    FMMA_GUARD(plan->pl.solver = type);
    return FMMA_OK;
}

int fmma_plan_set_poly_ord(fmma_plan plan, int poly_ord)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    if (poly_ord < 0 || poly_ord + 1 > fmma::lib::kMaxNodes)
        return fmma::lib::fail(FMMA_ERR_INVALID, "poly_ord must be in [0, 15]");
    plan->pl.poly_ord = poly_ord;
    ++plan->pl.gen;
    return FMMA_OK;
}

int fmma_plan_set_depth(fmma_plan plan, int depth)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    plan->pl.depth = depth;
    ++plan->pl.gen;
    return FMMA_OK;
}

int fmma_plan_get_depth(fmma_plan plan, int *depth)
{
    if (!plan || !depth) return fmma::lib::fail(FMMA_ERR_INVALID, "null argument");
    *depth = plan->pl.depth;
    return FMMA_OK;
}

int fmma_plan_set_trans_sym_flag(fmma_plan plan, int flag)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    plan->pl.trans_sym = flag;
    ++plan->pl.gen;
    return FMMA_OK;
}

int fmma_plan_set_nrn_N(fmma_plan plan, int nrn_N)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    plan->pl.nrn_N = nrn_N;
    return FMMA_OK;
}

int fmma_plan_get_nrn_N(fmma_plan plan, int *nrn_N)
{
    if (!plan || !nrn_N) return fmma::lib::fail(FMMA_ERR_INVALID, "null pointer");
    *nrn_N = plan->pl.nrn_N;
    return FMMA_OK;
}

int fmma_plan_set_kernel(fmma_plan plan, int kernel_id, const double *params, int nparams)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    if (nparams > 0 && !params) return fmma::lib::fail(FMMA_ERR_INVALID, "null params");
    fmma::lib::BuiltinParams bp;
    if (nparams > 0) bp.p0 = params[0];
    if (kernel_id == FMMA_KERNEL_GAUSS) {
        if (!(nparams > 0 && bp.p0 > 0.0)) return fmma::lib::fail(FMMA_ERR_INVALID, "gauss needs h > 0");
        bp.p1 = 1.0 / (2.0 * bp.p0 * bp.p0);
    }
    if (kernel_id == FMMA_KERNEL_YUKAWA && nparams < 1)
        return fmma::lib::fail(FMMA_ERR_INVALID, "yukawa needs kappa");
    fmma_kernel_ops o{};
    plan->pl.kparams = bp;
    const int rc = fmma::lib::builtin_kernel_ops(kernel_id, plan->pl.dim, plan->pl.dtype, &plan->pl.kparams, &o);
    if (rc) return fmma::lib::fail(rc, "unknown kernel id " + std::to_string(kernel_id));
    plan->pl.ops = o;
    plan->pl.kernel_id = kernel_id;
    ++plan->pl.gen;
    return FMMA_OK;
}

int fmma_plan_set_kernel_ops(fmma_plan plan, const fmma_kernel_ops *ops)
{
    if (!plan || !ops) return fmma::lib::fail(FMMA_ERR_INVALID, "null argument");
    if (ops->dim != plan->pl.dim || ops->dtype != plan->pl.dtype || !ops->p2p || !ops->exact || !ops->m2l_op ||
        !ops->m2l_general)
        return fmma::lib::fail(FMMA_ERR_INVALID, "kernel table incomplete or of another DIM/TYPE");
    plan->pl.ops = *ops;
    plan->pl.kernel_id = FMMA_KERNEL_CUSTOM;
    ++plan->pl.gen;
    return FMMA_OK;
}

int fmma_plan_set_m2l_precision(fmma_plan plan, int precision)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    if (precision < FMMA_M2L_NATIVE || precision > FMMA_M2L_BF16X3)
        return fmma::lib::fail(FMMA_ERR_INVALID, "unknown M2L precision");
    plan->pl.precision = precision;
    ++plan->pl.gen;
    return FMMA_OK;
}

int fmma_plan_set_stream(fmma_plan plan, void *stream)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    plan->pl.stream = (hipStream_t)stream;
    return FMMA_OK;
}

int fmma_plan_set_timing(fmma_plan plan, int enable)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    plan->pl.timing = enable != 0;
    return FMMA_OK;
}

int fmma_plan_set_p2p_mutual(fmma_plan plan, int enable)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    plan->pl.p2p_mutual = enable != 0;
    return FMMA_OK;
}

int fmma_plan_set_p2p_balance(fmma_plan plan, int enable)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    plan->pl.p2p_balance = enable != 0;
    return FMMA_OK;
}

int fmma_plan_set_batch_graph(fmma_plan plan, int enable)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    plan->pl.batch_graph = enable != 0;
    return FMMA_OK;
}

int fmma_plan_set_m2l_autotune(fmma_plan plan, int enable)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    plan->pl.m2l_autotune = enable != 0;
    return FMMA_OK;
}

int fmma_plan_set_m2l_tile(fmma_plan plan, int tile)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    if (tile != 0 && tile != 16 && tile != 32 && tile != 64 && tile != 128 && tile != 256)
        return fmma::lib::fail(FMMA_ERR_INVALID, "M2L tile width must be 0, 16, 32, 64, 128 or 256");
    fmma::lib::Plan &pl = plan->pl;
    pl.forced_tb = tile;
    pl.set_L = -1;  // rebuild the single-solve tasks
    pl.tuned_set_L = -1;
    if (pl.batch) {
        pl.batch->tuned = false;
        pl.batch->drop_graph();
    }
    return FMMA_OK;
}

int fmma_plan_m2l_tiles(fmma_plan plan, int batch, long *parents, int *tile, int max)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    const Plan &pl = plan->pl;
    int i = 0;
    for (const auto &kv : batch ? pl.tuned_tb : pl.used_tb_single) {
        if (i < max && parents && tile) {
            parents[i] = kv.first & 0xffffffffL;
            tile[i] = kv.second;
        }
        ++i;
    }
    return i;
}

int fmma_plan_solve_batch(fmma_plan plan, int n_problems, const long *t_counts, const void *targets,
                          const void *weights, const long *s_counts, const void *sources, void *ans)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    FMMA_GUARD(return fmma::lib::run_batch(plan->pl, n_problems, t_counts, targets, weights, s_counts, sources,
                                           ans));
}

int fmma_plan_solve(fmma_plan plan, long n_t, const void *targets, const void *weights, long n_s,
                    const void *sources, void *ans)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    FMMA_GUARD(return fmma::lib::solve_device(plan->pl, n_t, targets, weights, n_s, sources, ans));
}

// Host-pointer form of run_op: stages the inputs through device buffers.
// vec is the weights [n_s] (solve, gradient) or the cotangent [n_t]
// (transpose); out is ans [n_t] (read too: a shallow fmm adds onto it),
// the gradient [n_t][dim] or the transpose [n_s].
static int run_op_host(Plan &pl, fmma::lib::Op op, long n_t, const void *targets, const void *vec, long n_s,
                       const void *sources, void *out, bool reuse)
{
    using fmma::lib::Op;
    if (n_t < 0 || n_s < 0) return fmma::lib::fail(FMMA_ERR_INVALID, "negative size");
    const size_t ts = pl.tsize(), D = pl.dim;
    const long n_vec = op == Op::Transpose ? n_t : n_s;
    const long n_out = op == Op::Solve ? n_t : op == Op::Gradient ? n_t * (long)D : n_s;
    if (n_t == 0 || n_s == 0) {
        if (op == Op::Solve) return fmma::lib::run_op(pl, op, n_t, targets, vec, n_s, sources, out, reuse);
        if (n_out > 0) {
            if (!out) return fmma::lib::fail(FMMA_ERR_INVALID, "null pointer");
            std::memset(out, 0, n_out * ts);
        }
        return FMMA_OK;
    }
    if (!targets || !sources || !out) return fmma::lib::fail(FMMA_ERR_INVALID, "null pointer");
    int rc;
    hipStream_t st = pl.stream;
    const bool shared = targets == sources && n_t == n_s;
    if ((rc = fmma::lib::upload(pl.h_s, sources, n_s * D * ts, st))) return rc;
    if (!shared && (rc = fmma::lib::upload(pl.h_t, targets, n_t * D * ts, st))) return rc;
    if (vec && (rc = fmma::lib::upload(pl.h_w, vec, n_vec * ts, st))) return rc;
    if (op == Op::Solve) {
        if ((rc = fmma::lib::upload(pl.h_ans, out, n_out * ts, st))) return rc;
    } else if ((rc = pl.h_ans.ensure(n_out * ts))) {
        return rc;
    }
    const void *dt = shared ? pl.h_s.p : pl.h_t.p;
    FMMA_GUARD(rc = fmma::lib::run_op(pl, op, n_t, dt, vec ? pl.h_w.p : nullptr, n_s, pl.h_s.p, pl.h_ans.p, reuse));
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(out, pl.h_ans.p, n_out * ts, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fmma::lib::hip_fail(e, "copy the result to the host");
    return FMMA_OK;
}

int fmma_plan_solve_noalloc(fmma_plan plan, long n_t, const void *targets, const void *weights, long n_s,
                            const void *sources, void *ans)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    FMMA_GUARD(return fmma::lib::solve_noalloc(plan->pl, n_t, targets, weights, n_s, sources, ans));
}

int fmma_plan_solve_host(fmma_plan plan, long n_t, const void *targets, const void *weights, long n_s,
                         const void *sources, void *ans)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    return run_op_host(plan->pl, fmma::lib::Op::Solve, n_t, targets, weights, n_s, sources, ans, false);
}

int fmma_plan_gradient(fmma_plan plan, long n_t, const void *targets, const void *weights, long n_s,
                       const void *sources, void *grad, int reuse)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    FMMA_GUARD(return fmma::lib::run_op(plan->pl, fmma::lib::Op::Gradient, n_t, targets, weights, n_s, sources, grad,
                                        reuse != 0));
}

int fmma_plan_gradient_host(fmma_plan plan, long n_t, const void *targets, const void *weights, long n_s,
                            const void *sources, void *grad)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    return run_op_host(plan->pl, fmma::lib::Op::Gradient, n_t, targets, weights, n_s, sources, grad, false);
}

int fmma_plan_solve_transpose(fmma_plan plan, long n_t, const void *targets, const void *cot, long n_s,
                              const void *sources, void *out)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    FMMA_GUARD(return fmma::lib::run_op(plan->pl, fmma::lib::Op::Transpose, n_t, targets, cot, n_s, sources, out,
                                        false));
}

int fmma_plan_solve_transpose_host(fmma_plan plan, long n_t, const void *targets, const void *cot, long n_s,
                                   const void *sources, void *out)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    return run_op_host(plan->pl, fmma::lib::Op::Transpose, n_t, targets, cot, n_s, sources, out, false);
}

int fmma_plan_time_log(fmma_plan plan, const char **names, double *ms, int max)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    Plan &pl = plan->pl;
    pl.names_out.clear();
    int i = 0;
    for (const auto &kv : pl.time_log) {
        pl.names_out.push_back(kv.first);
        ++i;
    }
    const int n = i;
    i = 0;
    for (const auto &kv : pl.time_log) {
        if (i >= max) break;
        if (names) names[i] = pl.names_out[i].c_str();
        if (ms) ms[i] = kv.second;
        ++i;
    }
    return n;
}

int fmma_plan_clear_time_log(fmma_plan plan)
{
    if (!plan) return fmma::lib::fail(FMMA_ERR_INVALID, "null plan");
    plan->pl.time_log.clear();
    return FMMA_OK;
}

int fmma_plan_level_coeffs(fmma_plan plan, int which, int level, const void **ptr, long *n_boxes, int *P)
{
    if (!plan || !ptr || !n_boxes || !P) return fmma::lib::fail(FMMA_ERR_INVALID, "null argument");
    Plan &pl = plan->pl;
    if (pl.last_L < 0 || level < 0 || level > pl.last_L)
        return fmma::lib::fail(FMMA_ERR_INVALID, "no such level in the last fmm solve");
    const char *base = (which == 0 ? pl.Wm : pl.Wl).as<const char>();
    *ptr = base + pl.lvl_off[level] * pl.tsize();
    *n_boxes = 1L << (pl.dim * level);
    *P = pl.P();
    return FMMA_OK;
}

int fmma_plan_point_boxes(fmma_plan plan, int which, long n, long *boxes)
{
    if (!plan || !boxes) return fmma::lib::fail(FMMA_ERR_INVALID, "null argument");
    Plan &pl = plan->pl;
    if (pl.last_L < 0) return fmma::lib::fail(FMMA_ERR_INVALID, "no fmm solve yet");
    const bool tgt = which == 1 && pl.last_shared == 0;
    const long cnt = which == 1 ? pl.last_nt : pl.last_ns;
    if (n != cnt) return fmma::lib::fail(FMMA_ERR_INVALID, "n does not match the last solve");
    FMMA_GUARD({
        std::vector<unsigned> keys(n), perm(n);
        const fmma::lib::DBuf &kb = tgt ? pl.t_keys : pl.s_keys, &vb = tgt ? pl.t_vals : pl.s_vals;
        FMMA_HIP(hipStreamSynchronize(pl.stream));
        FMMA_HIP(hipMemcpy(keys.data(), kb.p, n * 4, hipMemcpyDeviceToHost));
        FMMA_HIP(hipMemcpy(perm.data(), vb.p, n * 4, hipMemcpyDeviceToHost));
        for (long i = 0; i < n; ++i) boxes[perm[i]] = fmma_morton_to_rowmajor(pl.dim, pl.last_L, keys[i]);
    });
    return FMMA_OK;
}

long fmma_morton_to_rowmajor(int dim, int level, long morton)
{
    int b[3] = {0, 0, 0};
    if (dim == 1) fmma::detail::morton_decode<1>((unsigned)morton, b);
    else if (dim == 2) fmma::detail::morton_decode<2>((unsigned)morton, b);
    else fmma::detail::morton_decode<3>((unsigned)morton, b);
    long id = 0;
    for (int d = 0; d < dim; ++d) id = (id << level) + b[d];
    return id;
}

const char *fmma_last_error(void) { return fmma::lib::g_err.c_str(); }

const char *fmma_version(void) { return "fmma-mi355x 0.1.0"; }

}  // extern "C"
