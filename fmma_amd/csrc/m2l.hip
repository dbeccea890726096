// m2l.hip -- M2L, the one dense contraction of the FMM, on MFMA.
//
// The reference applies, for every box b and every box m of its interaction
// list, Wl[b] += K_{m-b} Wm[m] as a P x P matvec (src/fmm_core.cpp:419-455,
// matvec src/math.cpp:263-294), with K_o[kl][km] = f(X_kl, Y_km) built once per
// level and offset (src/fmm_core.cpp:388-418).
//
// Here a workgroup owns 64 target boxes that share a child class sigma (the
// sigma-children of 64 consecutive parents in Morton order) and every row
// chunk of 64 coefficients of them.  For each of the class's 6^D - 3^D
// offsets o it gathers the 64 partner multipoles Wm[b + o] into LDS (zero
// columns where b + o leaves the grid) and multiplies them by the operator
// K_o with MFMA: C[kl][b] += K_o[kl][km] * Wm[b+o][km].  The accumulators stay
// in registers across all offsets, so Wl is written exactly once; K_o is read
// straight from global memory (L2) in fragment order.  All levels run in one
// launch.
//
// Lane layouts (CDNA3/CDNA4 MFMA, wave64):
//   16x16x4 (f64, f32): A[i][k]: lane = i + 16k;  B[k][j]: lane = j + 16k;
//                       f32: C[i][j]: lane = j + 16*(i/4), element i%4;
//                       f64: C[i][j]: lane = j + 16*(i%4), element i/4.
//   16x16x32 bf16:      A[i][k]: lane = i + 16*(k/8), element k%8;
//                       B[k][j]: lane = j + 16*(k/8), element k%8; C as above.
#include "m2l.hpp"
#include "../../include/fmma/detail/m2l_layout.hpp"

namespace fmma {
namespace lib {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------- host tables

void OffsetTables::build(int dim)
{
    D = dim;
    far.clear();
    cls_f.clear();
    cls_o.clear();
    int n7 = 1;
    for (int d = 0; d < D; ++d) n7 *= 7;
    std::vector<int> far_index(n7, -1);
    for (int s = 0; s < n7; ++s) {
        int o[3], ss = s, mx = 0;
        for (int d = D - 1; d >= 0; --d) {
            o[d] = ss % 7 - 3;
            ss /= 7;
            mx = std::max(mx, std::abs(o[d]));
        }
        if (mx <= 1) continue;
        far_index[s] = (int)(far.size() / D);
        for (int d = 0; d < D; ++d) far.push_back(o[d]);
    }
    n_far = (int)(far.size() / D);
    n_cls = 0;
    for (int c = 0; c < (1 << D); ++c) {
        int cnt = 0;
        for (int s = 0; s < n7; ++s) {
            if (far_index[s] < 0) continue;
            int o[3], ss = s;
            for (int d = D - 1; d >= 0; --d) {
                o[d] = ss % 7 - 3;
                ss /= 7;
            }
            bool ok = true;
            for (int d = 0; d < D; ++d) {
                const int par = (c >> (D - 1 - d)) & 1;
                const int lo = par ? -3 : -2, hi = par ? 2 : 3;
                ok = ok && o[d] >= lo && o[d] <= hi;
            }
            if (!ok) continue;
            cls_f.push_back(far_index[s]);
            for (int d = 0; d < D; ++d) cls_o.push_back(o[d]);
            ++cnt;
        }
        n_cls = cnt;  // 6^D - 3^D for every class
    }
}

// The contraction chunk of a padded length: all of it below 128 (and at 128
// for FP32 and BF16x3); FP64 at 128 in chunks of 32 (the smaller operand
// tiles and fragments in flight: 6.0 ms in one chunk, 4.6 ms in 64s, 4.4 ms
// in 32s at 5 nodes, N = 4M);
// beyond 128, 32 (FP64, FP32: KP is then a multiple of 32) or 128 (BF16x3: a
// multiple of 128), so that a step's MFMA work still covers the next step's
// loads.
constexpr int m2l_chunk(int KP, int mode)
{
    return KP < 128 ? KP : mode == 2 ? 128 : KP == 128 ? (mode == 0 ? 32 : 128) : 32;
}

// Row blocks of 16 per wave: 2 for BF16x3 at a padded length of 128 (one
// workgroup covers all MP <= 128 rows; measured 1.15 -> 1.10 ms at 5 nodes,
// N = 4M), else 1 (FP32 at 2 measured 2.7 -> 3.8 ms, FP64 unchanged).
constexpr int m2l_row_blocks(int KP, int mode) { return (KP == 128 && mode == 2) || (KP == 224 && mode == 0) ? 2 : 1; }

// Reorders the tasks [first, first + count) of one pipelined launch, made of
// runs of unequal cost (levels; a split level's tasks are shorter), so that
// the contiguous range m2l_pipe gives each XCD (below) holds an equal share
// of every run, each share contiguous in the original order (neighbouring
// tiles on one XCD share partners in its L2).  Run remainders are dealt to
// the XCDs cyclically, which gives XCD x the nt8 + (x < rem) tasks the
// kernel's mapping assigns it when every task is one workgroup (more row
// chunks per task shift the boundaries by at most one task).
static void xcd_balance(std::vector<M2LTask> &t, int first, int count, const std::vector<int> &runs)
{
    std::vector<M2LTask> per[8];
    int ptr = 0;
    for (size_t r = 0; r < runs.size(); ++r) {
        const int a = runs[r], b = r + 1 < runs.size() ? runs[r + 1] : first + count;
        const int n = b - a, base = n / 8, rem = n % 8;
        int pos = a;
        for (int x = 0; x < 8; ++x) {
            const int size = base + ((x - ptr + 8) % 8 < rem ? 1 : 0);
            per[x].insert(per[x].end(), t.begin() + pos, t.begin() + pos + size);
            pos += size;
        }
        ptr = (ptr + rem) % 8;
    }
    int o = first;
    for (int x = 0; x < 8; ++x)
        for (const M2LTask &k : per[x]) t[o++] = k;
}

// Tasks level by level, deepest first (inside a level in the order chosen
// below).  A level's tile width is the widest
// of 256/128/64 boxes that still gives >= 512 workgroups (two per CU), when a
// pipelined kernel exists for it; tasks of equal width form one launch.  A
// split level's parts of one tile are consecutive tasks (the same XCD).
std::vector<M2LTask> make_m2l_tasks(int D, int leaf_level, int P, int mode, std::vector<M2LGroup> &groups,
                                    const long *w_off, const long *c_off, const long *k_off, const int *tb_of_level,
                                    int only_level, std::vector<M2LSplit> *splits, long split_target)
{
    std::vector<M2LTask> t;
    groups.clear();
    if (splits) splits->clear();
    int MP, KP;
    m2l_shape(P, mode, MP, KP);
    int n_cls = 1, n3 = 1;
    for (int d = 0; d < D; ++d) {
        n_cls *= 6;
        n3 *= 3;
    }
    n_cls -= n3;  // interaction offsets per class
    const int rb = m2l_row_blocks(KP, mode);
    // workgroups per task (row chunks); 128 rows each for the 32x32 forms (4 waves x 32 rows)
    const long gy = m2l_wide(mode, KP) ? (MP + 127) / 128 : (MP / 16 + 4 * rb - 1) / (4 * rb);
    // Task order inside a level: tile-major (a tile's 2^D classes together)
    // or class-major (all tiles of one class together, so that each XCD's
    // contiguous share holds few classes and its tasks read the same
    // operator rows at about the same time).  Class-major halves the BF16x3
    // contraction's fabric reads at 5 nodes, N = 4M (6.5 -> 3.3 GB; 1.18 ->
    // 1.01 ms); FP64 and FP32 lose with it (4.39 -> 4.54 ms, 2.32 -> 2.33 ms).
    const bool class_major = mode == 2;
    long part_off = 0;
    std::vector<std::vector<int>> runs;  // per group: the first task of each level
    for (int l = leaf_level; l >= 2; --l) {
        if (only_level >= 0 && l != only_level) continue;
        const unsigned npar = 1u << (D * (l - 1));
        int tb = 64;
        for (int cand : {256, 128}) {
            if ((long)(npar / cand) * (1 << D) >= 512 && m2l_pipe_supported(D, KP, mode, cand)) {
                tb = cand;
                break;
            }
        }
        if (tb_of_level && tb_of_level[l] > 0) tb = tb_of_level[l];
        const bool piped = m2l_pipe_supported(D, KP, mode, tb);
        if (groups.empty() || groups.back().tb != tb || groups.back().pipelined != piped) {
            groups.push_back(M2LGroup{tb, (int)t.size(), 0, piped});
            runs.emplace_back();
        }
        runs.back().push_back((int)t.size());
        int parts = 1;
        if (splits && split_target > 0 && piped) {
            // in workgroups of 64 rows (one of the 32x32 forms covers two)
            const long wg = (long)((npar + tb - 1) / tb) * (1L << D) * (m2l_wide(mode, KP) ? 2 * gy : gy);
            long p = split_target / wg;
            if (p > 8) p = 8;
            if (p > n_cls / 16) p = n_cls / 16;
            parts = p >= 2 ? (int)p : 1;
        }
        const long count = (long)P << (D * l);
        const int nbg = (int)((npar + tb - 1) / tb);  // tiles per class
        for (int it = 0; it < nbg << D; ++it) {
            const int c = class_major ? it / nbg : it % (1 << D);
            const unsigned p0 = (unsigned)(class_major ? it % nbg : it >> D) * tb;
            const unsigned np = npar - p0 < (unsigned)tb ? npar - p0 : (unsigned)tb;
            for (int q = 0; q < parts; ++q) {
                M2LTask k{l, c, p0, (int)np, w_off[l], c_off[l], k_off[l]};
                if (parts > 1) {
                    k.i0 = q * n_cls / parts;
                    k.i1 = (q + 1) * n_cls / parts;
                    k.o_off = part_off + q * count;
                }
                t.push_back(k);
                ++groups.back().count;
            }
        }
        if (parts > 1) {
            splits->push_back(M2LSplit{l, parts, w_off[l], count, part_off});
            part_off += parts * count;
        }
    }
    for (size_t g = 0; g < groups.size(); ++g)
        if (groups[g].pipelined && runs[g].size() > 1) xcd_balance(t, groups[g].first, groups[g].count, runs[g]);
    return t;
}

long m2l_split_elems(const std::vector<M2LSplit> &splits)
{
    long e = 0;
    for (const M2LSplit &s : splits) e = std::max(e, s.part_off + s.parts * s.count);
    return e;
}

bool m2l_wide(int mode, int KP)
{
    return (mode == 1 && (KP == 128 || KP == 224)) || (mode == 2 && (KP == 128 || KP == 256));
}

int m2l_frag_format(int P, int mode)
{
    int MP, KP;
    m2l_shape(P, mode, MP, KP);
    const bool wide = m2l_wide(mode, KP);
    return mode == 0 ? FMMA_FRAG_F64 : mode == 2 ? (wide ? FMMA_FRAG_BF16W : FMMA_FRAG_BF16)
                                                 : (wide ? FMMA_FRAG_F32W : FMMA_FRAG_F32);
}

void m2l_shape(int P, int mode, int &MP, int &KP)
{
    MP = (P + 15) / 16 * 16;
    // Padded to a length with a pipelined kernel: 32, 48, 64, 96 or 128
    // (FP64, FP32; 48 and 96 serve 2-D orders such as 6 nodes, P = 36), a
    // multiple of 32 (BF16x3).  Beyond 128 the pipelined contraction takes
    // chunks of 32 (FP64, FP32) or 128 (BF16x3) of the padded length.  Up to
    // 16 coefficients (low orders, 1-D) the length is only rounded to the
    // MFMA's k-step of 4.
    if (mode == 2) {
        KP = P <= 128 ? (P + 31) / 32 * 32 : (P + 127) / 128 * 128;
    } else if (P <= 16) {
        KP = (P + 3) / 4 * 4;
    } else if (P <= 128) {
        KP = 128;
        for (int c : {32, 48, 64, 96})
            if (P <= c) {
                KP = c;
                break;
            }
    } else {
        KP = (P + 31) / 32 * 32;
    }
    if (m2l_wide(mode, KP)) MP = (P + 31) / 32 * 32;  // whole 32-row blocks
}

size_t m2l_frag_elems(int n_far, int P, int mode)
{
    int MP, KP;
    m2l_shape(P, mode, MP, KP);
    return (size_t)n_far * MP * KP;
}

// --------------------------------------------------------------- fragments

template <typename T, typename CT>
__global__ __launch_bounds__(256) void frag4_kernel(const T *Kraw, CT *Kf, int P, int MP, int KP, long total,
                                                    int format)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const long per = (long)MP * KP;
    const long f = e / per;
    int row, k;
    detail::frag_rowcol(format, KP, e - f * per, row, k);
    Kf[e] = (row < P && k < P) ? (CT)Kraw[(f * P + row) * P + k] : CT(0);
}

template <typename T>
__global__ __launch_bounds__(256) void frag_bf16_kernel(const T *Kraw, __bf16 *Kh, __bf16 *Kl, int P, int MP,
                                                        int KP, long total, int format)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const long per = (long)MP * KP;
    const long f = e / per;
    int row, k;
    detail::frag_rowcol(format, KP, e - f * per, row, k);
    const float v = (row < P && k < P) ? (float)Kraw[(f * P + row) * P + k] : 0.0f;
    const __bf16 h = (__bf16)v;
    Kh[e] = h;
    Kl[e] = (__bf16)(v - (float)h);
}

template <typename T>
int m2l_fragments(const T *Kraw, int n_far, int P, int mode, void *Kf, void *Kf_lo, hipStream_t st)
{
    int MP, KP;
    m2l_shape(P, mode, MP, KP);
    const long total = (long)n_far * MP * KP;
    const unsigned nb = (unsigned)((total + 255) / 256);
    if (mode == 0)
        hipLaunchKernelGGL((frag4_kernel<T, double>), dim3(nb), dim3(256), 0, st, Kraw, (double *)Kf, P, MP, KP, total,
                           (int)FMMA_FRAG_F64);
    else if (mode == 1)
        hipLaunchKernelGGL((frag4_kernel<T, float>), dim3(nb), dim3(256), 0, st, Kraw, (float *)Kf, P, MP, KP, total,
                           m2l_frag_format(P, mode));
    else
        hipLaunchKernelGGL((frag_bf16_kernel<T>), dim3(nb), dim3(256), 0, st, Kraw, (__bf16 *)Kf, (__bf16 *)Kf_lo, P,
                           MP, KP, total, m2l_frag_format(P, mode));
    FMMA_LAUNCHED("m2l_fragments");
    return FMMA_OK;
}

// ---------------------------------------------------------------- contraction

// Target column c of the task: sigma-child of parent p0 + c.  Writes the
// partner box index of offset i (or -1) for the 64 columns.
template <int D>
__device__ __forceinline__ void partner_indices(const M2LKArgs &a, int cls, int i, int N, const int (&tc)[D],
                                                bool tvalid, int *pidx, int tid)
{
    if (tid < 64) {
        int m[D];
        bool ok = tvalid;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            m[d] = tc[d] + a.cls_o[(cls * a.n_cls + i) * D + d];
            ok = ok && m[d] >= 0 && m[d] < N;
        }
        pidx[tid] = ok ? (int)morton_encode<D>(m) : -1;
    }
}

template <int D>
__device__ __forceinline__ void target_coords(const M2LTask &t, int tid, int (&tc)[D], bool &tvalid)
{
    tvalid = tid < t.np;
    int pc[D];
    morton_decode<D>(t.p0 + (tvalid ? tid : 0), pc);
#pragma unroll
    for (int d = 0; d < D; ++d) tc[d] = 2 * pc[d] + ((t.cls >> (D - 1 - d)) & 1);
}

// LDS a pipelined M2L workgroup may take for two operand tiles.
constexpr int kM2LDoubleLds = 160 * 1024 - 4096;

// Row padding (elements) of an LDS operand tile [column][k] so that the
// lanes of an MFMA B-operand read (column = lane % 16, k = lane / 16 + a
// step) hit distinct banks of the 32: one element for FP64 (8-byte reads of
// 16 lanes at a time: word 2 * column, the 32 banks; KC + 2 measured 5x the
// conflicts), two for FP32 (half a wave, columns 0..15 at k and k + 1:
// bank 2 * column + k), 8 for BF16 (16-byte reads of 8 lanes: 4 banks each).
template <int ESIZE>
constexpr int kLdsPad = ESIZE == 2 ? 8 : ESIZE == 4 ? 2 : 1;

// Operand rows staged per LDS pass of m2l_mfma4: 64 boxes x (256 + 1) FP64
// values = 132 KB, inside the 160 KB of a CU for every P <= 512.
constexpr int kM2LChunk = 256;

// MODE 0: FP64 MFMA on FP64 data.  MODE 1: FP32 MFMA, per-offset FP32 sums
// added into TW accumulators.  TW is the storage type of Wm/Wl.
template <typename TW, int D, int MODE>
__global__ __launch_bounds__(256) void m2l_mfma4(M2LKArgs a)
{
    using CT = typename std::conditional<MODE == 0, double, float>::type;
    using V4 = typename std::conditional<MODE == 0, d4_t, f4_t>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    CT *Bs = reinterpret_cast<CT *>(smem);
    __shared__ int pidx[64];

    const M2LTask task = a.tasks[blockIdx.x];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int rb = blockIdx.y * 4 + wave;
    const bool active = rb * 16 < a.MP;
    const int N = 1 << task.level, P = a.P, KP = a.KP;
    const int KC = KP < kM2LChunk ? KP : kM2LChunk, ldb = KC + kLdsPad<sizeof(CT)>;
    int tc[D];
    bool tvalid;
    target_coords<D>(task, tid, tc, tvalid);

    const TW *Wm = static_cast<const TW *>(a.Wm) + task.w_off;
    const CT *Kf = static_cast<const CT *>(a.Kf) + task.k_off;
    const int nks = KP / 4, nrb = a.MP / 16;

    TW acc[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[c][q] = TW(0);

    for (int i = 0; i < a.n_cls; ++i) {
        const int f = a.cls_f[task.cls * a.n_cls + i];
        partner_indices<D>(a, task.cls, i, N, tc, tvalid, pidx, tid);
        __syncthreads();
        const CT *A = Kf + ((size_t)f * nrb + rb) * nks * 64;
        V4 c[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) c[cb] = V4{0, 0, 0, 0};
        // the contraction in chunks of KC operand rows staged in LDS
        for (int k0 = 0; k0 < KP; k0 += KC) {
            const int kc = KP - k0 < KC ? KP - k0 : KC;
            for (int e = tid; e < 64 * kc; e += 256) {
                const int col = e / kc, k = e - col * kc;
                const int pi = pidx[col];
                CT v = CT(0);
                if (pi >= 0 && k0 + k < P) v = (CT)Wm[(size_t)pi * P + k0 + k];
                Bs[col * ldb + k] = v;
            }
            __syncthreads();
            if (active) {
                for (int s = k0 / 4; s < (k0 + kc) / 4; ++s) {
                    const CT av = A[s * 64 + lane];
                    const int kk = 4 * s + (lane >> 4) - k0;
#pragma unroll
                    for (int cb = 0; cb < 4; ++cb) {
                        const CT bv = Bs[(16 * cb + (lane & 15)) * ldb + kk];
                        if constexpr (MODE == 0)
                            c[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, c[cb], 0, 0, 0);
                        else
                            c[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, c[cb], 0, 0, 0);
                    }
                }
            }
            __syncthreads();  // the chunk is consumed before the next is staged
        }
        if (active) {
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[cb][q] += (TW)c[cb][q];
        }
    }
    if (!active) return;
    TW *Wl = static_cast<TW *>(a.Wl) + task.w_off;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const int col = 16 * cb + (lane & 15);
        if (col >= task.np) continue;
        const size_t box = ((size_t)(task.p0 + col) << D) | (unsigned)task.cls;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = MODE == 0 ? 16 * rb + (lane >> 4) + 4 * q : 16 * rb + 4 * (lane >> 4) + q;
            if (row < P) Wl[box * P + row] = acc[cb][q];
        }
    }
}

// Split-BF16 x3: K = Kh + Kl, W = Wh + Wl (each bf16); per offset
// C = Kl*Wh + Kh*Wl + Kh*Wh with FP32 accumulation, then added into TW.
template <typename TW, int D>
__global__ __launch_bounds__(256) void m2l_bf16x3(M2LKArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int pidx[64];
    const M2LTask task = a.tasks[blockIdx.x];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int rb = blockIdx.y * 4 + wave;
    const bool active = rb * 16 < a.MP;
    const int N = 1 << task.level, P = a.P, KP = a.KP, ldb = KP + 8;
    __bf16 *Bh = reinterpret_cast<__bf16 *>(smem);
    __bf16 *Bl = Bh + 64 * ldb;
    int tc[D];
    bool tvalid;
    target_coords<D>(task, tid, tc, tvalid);

    const TW *Wm = static_cast<const TW *>(a.Wm) + task.w_off;
    const bf8_t *Kh = reinterpret_cast<const bf8_t *>(static_cast<const __bf16 *>(a.Kf) + task.k_off);
    const bf8_t *Kl = reinterpret_cast<const bf8_t *>(static_cast<const __bf16 *>(a.Kf_lo) + task.k_off);
    const int nks = KP / 32, nrb = a.MP / 16;

    TW acc[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[c][q] = TW(0);

    for (int i = 0; i < a.n_cls; ++i) {
        const int f = a.cls_f[task.cls * a.n_cls + i];
        partner_indices<D>(a, task.cls, i, N, tc, tvalid, pidx, tid);
        __syncthreads();
        for (int e = tid; e < 64 * KP; e += 256) {
            const int col = e / KP, k = e - col * KP;
            const int pi = pidx[col];
            float v = 0.0f;
            if (pi >= 0 && k < P) v = (float)Wm[(size_t)pi * P + k];
            const __bf16 h = (__bf16)v;
            Bh[col * ldb + k] = h;
            Bl[col * ldb + k] = (__bf16)(v - (float)h);
        }
        __syncthreads();
        if (active) {
            const size_t abase = ((size_t)f * nrb + rb) * nks * 64;
            f4_t c[4];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) c[cb] = f4_t{0, 0, 0, 0};
            for (int s = 0; s < nks; ++s) {
                const bf8_t ah = Kh[abase + s * 64 + lane];
                const bf8_t al = Kl[abase + s * 64 + lane];
                const int kk = 32 * s + 8 * (lane >> 4);
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) {
                    const int off = (16 * cb + (lane & 15)) * ldb + kk;
                    const bf8_t bh = *reinterpret_cast<const bf8_t *>(Bh + off);
                    const bf8_t bl = *reinterpret_cast<const bf8_t *>(Bl + off);
                    c[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c[cb], 0, 0, 0);
                    c[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c[cb], 0, 0, 0);
                    c[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c[cb], 0, 0, 0);
                }
            }
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[cb][q] += (TW)c[cb][q];
        }
    }
    if (!active) return;
    TW *Wl = static_cast<TW *>(a.Wl) + task.w_off;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const int col = 16 * cb + (lane & 15);
        if (col >= task.np) continue;
        const size_t box = ((size_t)(task.p0 + col) << D) | (unsigned)task.cls;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 16 * rb + 4 * (lane >> 4) + q;
            if (row < P) Wl[box * P + row] = acc[cb][q];
        }
    }
}

// Software-pipelined contraction, compile-time padded length KP, tile width
// TB (target boxes per workgroup, 16..256) and contraction chunk KC (KP
// itself up to 128; for longer rows the contraction is cut into KP/KC steps
// per offset so that operand tiles and operator fragments stay small).  A
// step is (offset, chunk).  Thread t gathers TPC = 256/TB threads' share of
// one target's partner chunk: column t / TPC, elements [(t % TPC) * EPT,
// +EPT) of the chunk of the padded operand copy, with 16-byte loads, for step
// j+1 while the MFMAs of step j read LDS; the operator fragments of step j+1
// are prefetched the same way.  Accumulators stay in registers across all
// steps: FP64 in MODE 0; in MODES 1-2 the MFMA's FP32 accumulator sums one
// offset and is then added into FP64 registers, so rounding in FP32 covers
// KP products, not the whole interaction list.  Each wave owns RB row blocks
// of 16 rows: every B fragment it reads from LDS feeds RB MFMAs, and a
// workgroup covers 64 RB rows, so fewer workgroups gather each partner tile.
template <typename TW, int D, int MODE, int KP, int TB, int KC, int RB>
__global__ __launch_bounds__(256) void m2l_pipe(M2LKArgs a)
{
    static_assert(KP % KC == 0, "the chunk divides the padded length");
    constexpr int NCH = KP / KC;                               // chunks per offset
    constexpr int NKS = MODE == 2 ? KP / 32 : KP / 4;          // k-steps per offset
    constexpr int NKC = MODE == 2 ? KC / 32 : KC / 4;          // k-steps per chunk
    constexpr int TPC = 256 / TB, EPT = KC / TPC, NCB = TB / 16;
    using CT = typename std::conditional<MODE == 0, double, float>::type;
    using V4 = typename std::conditional<MODE == 0, d4_t, f4_t>::type;
    using ST = typename std::conditional<MODE == 2, __bf16, CT>::type;
    using AT = typename std::conditional<MODE == 2, bf8_t, CT>::type;
    constexpr int LDB = KC + kLdsPad<sizeof(ST)>;
    constexpr int NPART = MODE == 2 ? 2 : 1;
    constexpr int VEC = 16 / (int)sizeof(ST);  // elements per 16-byte load
    static_assert(EPT % VEC == 0, "gather chunks must be whole 16-byte vectors");
    // Two tiles when they fit: step j+1's tile is written while step j's
    // MFMAs run, one barrier per step instead of two.
    constexpr int BUF = NPART * TB * LDB;
    constexpr bool DB = 2 * BUF * (int)sizeof(ST) <= kM2LDoubleLds;
    __shared__ __align__(16) ST Bs[(DB ? 2 : 1) * BUF];

    // Workgroups are dealt to the 8 XCDs round-robin; give each XCD a
    // contiguous run of (task, row chunk) pairs, the row chunks of a task
    // next to each other, so that its L2 serves the partner tiles the chunks
    // of one task share and the operator rows the tasks of one class share.
    const unsigned gy = (unsigned)((a.MP / 16 + 4 * RB - 1) / (4 * RB));  // row chunks per task
    const unsigned n8 = gridDim.x / 8u, rem = gridDim.x % 8u, xcd = blockIdx.x % 8u;
    const unsigned v = xcd * n8 + (xcd < rem ? xcd : rem) + blockIdx.x / 8u;
    const unsigned tix = v / gy, chunk = v - tix * gy;
    const M2LTask task = a.tasks[tix];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int rb0 = ((int)chunk * 4 + wave) * RB;
    bool act[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) act[r] = (rb0 + r) * 16 < a.MP;
    const bool active = act[0];
    const int N = 1 << task.level, P = a.P, nrb = a.MP / 16;
    const int col = tid / TPC, part = tid % TPC;

    int tc[D];
    const bool cvalid = col < task.np;
    {
        int pc[D];
        morton_decode<D>(task.p0 + (cvalid ? col : 0), pc);
#pragma unroll
        for (int d = 0; d < D; ++d) tc[d] = 2 * pc[d] + ((task.cls >> (D - 1 - d)) & 1);
    }
    const ST *Wg = static_cast<const ST *>(a.Wc) + task.c_off;
    [[maybe_unused]] const ST *Wg_lo = static_cast<const ST *>(a.Wc_lo) + task.c_off;
    // a split level's task takes offsets [i0, i1) of its class
    const int i0 = task.i1 ? task.i0 : 0, i1 = task.i1 ? task.i1 : a.n_cls;
    const int *cls_f = a.cls_f + task.cls * a.n_cls + i0;
    const int *cls_o = a.cls_o + (task.cls * a.n_cls + i0) * D;

    uint4 g[EPT / VEC];
    [[maybe_unused]] uint4 g_lo[MODE == 2 ? EPT / VEC : 1];
    auto gather = [&](int j) {
        const int i = NCH == 1 ? j : j / NCH, c = NCH == 1 ? 0 : j - i * NCH;
        int m[D];
        bool ok = cvalid;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            m[d] = tc[d] + cls_o[i * D + d];
            ok = ok && m[d] >= 0 && m[d] < N;
        }
        const size_t off = (size_t)(ok ? morton_encode<D>(m) : 0u) * KP + c * KC + part * EPT;
        const uint4 *src = reinterpret_cast<const uint4 *>(Wg + off);
#pragma unroll
        for (int e = 0; e < EPT / VEC; ++e) g[e] = ok ? src[e] : uint4{0, 0, 0, 0};
        if constexpr (MODE == 2) {
            const uint4 *sl = reinterpret_cast<const uint4 *>(Wg_lo + off);
#pragma unroll
            for (int e = 0; e < EPT / VEC; ++e) g_lo[e] = ok ? sl[e] : uint4{0, 0, 0, 0};
        }
    };
    AT afr[RB][NKC];
    [[maybe_unused]] AT afr_lo[RB][MODE == 2 ? NKC : 1];
    auto load_a = [&](int j) {
        const int i = NCH == 1 ? j : j / NCH, c = NCH == 1 ? 0 : j - i * NCH;
        const size_t f = (size_t)cls_f[i];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            if (!act[r]) continue;
            const size_t base = ((f * nrb + rb0 + r) * NKS + (size_t)c * NKC) * 64;
            if constexpr (MODE == 2) {
                const bf8_t *Kh = reinterpret_cast<const bf8_t *>(static_cast<const __bf16 *>(a.Kf) + task.k_off);
                const bf8_t *Kl =
                    reinterpret_cast<const bf8_t *>(static_cast<const __bf16 *>(a.Kf_lo) + task.k_off);
#pragma unroll
                for (int s = 0; s < NKC; ++s) {
                    afr[r][s] = Kh[base + s * 64 + lane];
                    afr_lo[r][s] = Kl[base + s * 64 + lane];
                }
            } else {
                const CT *Kf = static_cast<const CT *>(a.Kf) + task.k_off;
#pragma unroll
                for (int s = 0; s < NKC; ++s) afr[r][s] = Kf[base + s * 64 + lane];
            }
        }
    };

    V4 acc[RB][NCB];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) acc[r][cb] = V4{0, 0, 0, 0};
    [[maybe_unused]] double accd[MODE == 0 ? 1 : RB][MODE == 0 ? 1 : NCB][4];
    if constexpr (MODE != 0) {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                for (int q = 0; q < 4; ++q) accd[r][cb][q] = 0.0;
    }

    const int n = (i1 - i0) * NCH;
    auto store_tile = [&](ST *tile) {  // registers -> LDS tile
        ST *dst = tile + col * LDB + part * EPT;
        if constexpr (MODE == 2) {
#pragma unroll
            for (int e = 0; e < EPT / VEC; ++e) {
                reinterpret_cast<uint4 *>(dst)[e] = g[e];
                reinterpret_cast<uint4 *>(dst + TB * LDB)[e] = g_lo[e];
            }
        } else {
            const ST *v = reinterpret_cast<const ST *>(g);
#pragma unroll
            for (int e = 0; e < EPT; ++e) dst[e] = v[e];
        }
    };
    gather(0);
    load_a(0);
    if constexpr (DB) {
        store_tile(Bs);
        __syncthreads();
    }
    for (int i = 0; i < n; ++i) {
        const ST *cur = DB ? Bs + (i & 1) * BUF : Bs;
        if constexpr (!DB) {
            store_tile(Bs);
            __syncthreads();
        }
        AT acur[RB][NKC];
        [[maybe_unused]] AT acur_lo[RB][MODE == 2 ? NKC : 1];
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int s = 0; s < NKC; ++s) {
                acur[r][s] = afr[r][s];
                if constexpr (MODE == 2) acur_lo[r][s] = afr_lo[r][s];
            }
        if (i + 1 < n) {
            gather(i + 1);
            load_a(i + 1);
        }
        if (active) {
            if constexpr (MODE == 2) {
#pragma unroll
                for (int s = 0; s < NKC; ++s) {
                    const int kk = 32 * s + 8 * (lane >> 4);
#pragma unroll
                    for (int cb = 0; cb < NCB; ++cb) {
                        const int off = (16 * cb + (lane & 15)) * LDB + kk;
                        const bf8_t bh = *reinterpret_cast<const bf8_t *>(cur + off);
                        const bf8_t bl = *reinterpret_cast<const bf8_t *>(cur + TB * LDB + off);
#pragma unroll
                        for (int r = 0; r < RB; ++r) {
                            if (!act[r]) continue;
                            acc[r][cb] =
                                __builtin_amdgcn_mfma_f32_16x16x32_bf16(acur_lo[r][s], bh, acc[r][cb], 0, 0, 0);
                            acc[r][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(acur[r][s], bl, acc[r][cb], 0, 0, 0);
                            acc[r][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(acur[r][s], bh, acc[r][cb], 0, 0, 0);
                        }
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < NKC; ++s) {
                    const int kk = 4 * s + (lane >> 4);
#pragma unroll
                    for (int cb = 0; cb < NCB; ++cb) {
                        const CT bv = cur[(16 * cb + (lane & 15)) * LDB + kk];
#pragma unroll
                        for (int r = 0; r < RB; ++r) {
                            if (!act[r]) continue;
                            if constexpr (MODE == 0)
                                acc[r][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(acur[r][s], bv, acc[r][cb], 0, 0, 0);
                            else
                                acc[r][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(acur[r][s], bv, acc[r][cb], 0, 0, 0);
                        }
                    }
                }
            }
            if constexpr (MODE != 0) {
                if (NCH == 1 || (i + 1) % NCH == 0) {  // one offset summed: into FP64
#pragma unroll
                    for (int r = 0; r < RB; ++r)
#pragma unroll
                        for (int cb = 0; cb < NCB; ++cb) {
#pragma unroll
                            for (int q = 0; q < 4; ++q) accd[r][cb][q] += (double)acc[r][cb][q];
                            acc[r][cb] = V4{0, 0, 0, 0};
                        }
                }
            }
        }
        if constexpr (DB) {
            if (i + 1 < n) store_tile(Bs + ((i + 1) & 1) * BUF);
        }
        __syncthreads();
    }
    if (!active) return;
    TW *Wl = task.i1 ? static_cast<TW *>(a.Wpart) + task.o_off : static_cast<TW *>(a.Wl) + task.w_off;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int cc = 16 * cb + (lane & 15);
        if (cc >= task.np) continue;
        const size_t box = ((size_t)(task.p0 + cc) << D) | (unsigned)task.cls;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int rb = rb0 + r;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = MODE == 0 ? 16 * rb + (lane >> 4) + 4 * q : 16 * rb + 4 * (lane >> 4) + q;
                if constexpr (MODE == 0) {
                    if (row < P) Wl[box * P + row] = (TW)acc[r][cb][q];
                } else {
                    if (row < P) Wl[box * P + row] = (TW)accd[r][cb][q];
                }
            }
        }
    }
}

// Tile widths with a pipelined kernel for (MODE, KP, TB); the others use
// m2l_mfma4 / m2l_bf16x3.
template <typename T, int D, int MODE, int KP, int TB>
static bool launch_pipe_tb(const M2LKArgs &a, dim3 grid, hipStream_t st)
{
    constexpr int VEC = MODE == 2 ? 8 : MODE == 1 ? 4 : 2;
    constexpr int KC = m2l_chunk(KP, MODE);
    constexpr int RB = m2l_row_blocks(KP, MODE);
    if constexpr ((KC / (256 / TB)) % VEC != 0 || (MODE == 2 && KP % 32 != 0) || (MODE == 0 && TB > 128) ||
                  (MODE == 2 && KP > 128 && TB > 128)) {
        return false;
    } else {
        const dim3 g1(grid.x * (unsigned)((a.MP / 16 + 4 * RB - 1) / (4 * RB)));
        hipLaunchKernelGGL((m2l_pipe<T, D, MODE, KP, TB, KC, RB>), g1, dim3(256), 0, st, a);
        return true;
    }
}

template <typename T, int D, int MODE, int KP>
static bool launch_pipe_kp(const M2LKArgs &a, int tb, dim3 grid, hipStream_t st)
{
    if (tb == 64) return launch_pipe_tb<T, D, MODE, KP, 64>(a, grid, st);
    if constexpr (D == 3) {
        if constexpr (KP <= 128) {
            if (tb == 16) return launch_pipe_tb<T, D, MODE, KP, 16>(a, grid, st);
            if (tb == 32) return launch_pipe_tb<T, D, MODE, KP, 32>(a, grid, st);
        }
        if (tb == 128) return launch_pipe_tb<T, D, MODE, KP, 128>(a, grid, st);
        if (tb == 256) return launch_pipe_tb<T, D, MODE, KP, 256>(a, grid, st);
    }
    return false;
}

template <typename T, int D, int MODE>
static bool launch_pipe(const M2LKArgs &a, int tb, dim3 grid, hipStream_t st)
{
    if constexpr (D == 1) {
        return false;  // at most 16 coefficients per box
    } else {
        switch (a.KP) {
        case 32: return launch_pipe_kp<T, D, MODE, 32>(a, tb, grid, st);
        case 64: return launch_pipe_kp<T, D, MODE, 64>(a, tb, grid, st);
        case 128: return launch_pipe_kp<T, D, MODE, 128>(a, tb, grid, st);
        }
    }
    if constexpr (D == 2 && MODE != 2) {
        switch (a.KP) {
        case 48: return launch_pipe_kp<T, D, MODE, 48>(a, tb, grid, st);
        case 96: return launch_pipe_kp<T, D, MODE, 96>(a, tb, grid, st);
        }
    }
    if constexpr (D == 3) {
        // 6, 7 and 8 nodes per dimension
        if constexpr (MODE == 2) {
            switch (a.KP) {
            case 256: return launch_pipe_kp<T, D, MODE, 256>(a, tb, grid, st);
            case 384: return launch_pipe_kp<T, D, MODE, 384>(a, tb, grid, st);
            case 512: return launch_pipe_kp<T, D, MODE, 512>(a, tb, grid, st);
            }
        } else {
            switch (a.KP) {
            case 224: return launch_pipe_kp<T, D, MODE, 224>(a, tb, grid, st);
            case 352: return launch_pipe_kp<T, D, MODE, 352>(a, tb, grid, st);
            case 512: return launch_pipe_kp<T, D, MODE, 512>(a, tb, grid, st);
            }
        }
    }
    return false;
}

// FP32 form at a padded length of 128 or 224 on the 32x32x2 MFMA: a
// workgroup of 4 waves covers 128 rows (wave w: rows 32 w .. 32 w + 31; at
// 224, 7 blocks of 32, two workgroups per task) of TB target boxes, so a
// partner tile is gathered once per 128 rows and every operator
// value and every staged operand value feeds twice the flops of the 16x16x4
// shape.  The contraction runs in chunks of KC (64; 32 for 64 boxes, which
// then fits three waves per SIMD: M2L 2.16 -> 2.13 ms at 5 nodes, N = 4M)
// per offset, pipelined as in m2l_pipe (gather and operator loads of step
// j + 1 during the MFMAs of step j, two LDS tiles); an offset's FP32 sum is
// added into FP64 registers.
// Operator fragments are FMMA_FRAG_F32W: one 16-byte load per lane brings 4
// steps.
template <typename TW, int D, int TB, int KC = 64, int KP = 128>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void m2l_pipe_wide(M2LKArgs a)
{
    static_assert(KP % KC == 0, "the chunk divides the padded length");
    constexpr int NCH = KP / KC;
    constexpr int NG = KC / 8;                    // 16-byte operator groups per chunk (4 steps of 2)
    constexpr int TPC = 256 / TB, EPT = KC / TPC, NCB = TB / 32;
    constexpr int LDB = KC + 1;                   // odd: a 32-lane read of one k hits 32 banks
    typedef float f16v __attribute__((ext_vector_type(16)));
    static_assert(EPT % 4 == 0, "gather chunks must be whole 16-byte vectors");
    constexpr int BUF = TB * LDB;
    __shared__ __align__(16) float Bs[2 * BUF];

    // row chunks of 128 (KP > 128: several per task) on neighbouring workgroups of one XCD
    constexpr unsigned GY = (unsigned)(KP + 127) / 128;
    const unsigned n8 = gridDim.x / 8u, rem = gridDim.x % 8u, xcd = blockIdx.x % 8u;
    const unsigned v = xcd * n8 + (xcd < rem ? xcd : rem) + blockIdx.x / 8u;
    const unsigned tix = v / GY;
    const M2LTask task = a.tasks[tix];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int rb = (int)(v - tix * GY) * 4 + wave;  // the wave's 32-row block
    const bool active = rb * 32 < a.MP;
    const int N = 1 << task.level, P = a.P;
    const int col = tid / TPC, part = tid % TPC;

    int tc[D];
    const bool cvalid = col < task.np;
    {
        int pc[D];
        morton_decode<D>(task.p0 + (cvalid ? col : 0), pc);
#pragma unroll
        for (int d = 0; d < D; ++d) tc[d] = 2 * pc[d] + ((task.cls >> (D - 1 - d)) & 1);
    }
    const float *Wg = static_cast<const float *>(a.Wc) + task.c_off;
    const int i0 = task.i1 ? task.i0 : 0, i1 = task.i1 ? task.i1 : a.n_cls;
    const int *cls_f = a.cls_f + task.cls * a.n_cls + i0;
    const int *cls_o = a.cls_o + (task.cls * a.n_cls + i0) * D;
    const float4 *Kf = reinterpret_cast<const float4 *>(static_cast<const float *>(a.Kf) + task.k_off);

    float4 g[EPT / 4];
    auto gather = [&](int j) {
        const int i = j / NCH, c = j - i * NCH;
        int m[D];
        bool ok = cvalid;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            m[d] = tc[d] + cls_o[i * D + d];
            ok = ok && m[d] >= 0 && m[d] < N;
        }
        const size_t off = (size_t)(ok ? morton_encode<D>(m) : 0u) * KP + c * KC + part * EPT;
        const float4 *src = reinterpret_cast<const float4 *>(Wg + off);
#pragma unroll
        for (int e = 0; e < EPT / 4; ++e) g[e] = ok ? src[e] : float4{0, 0, 0, 0};
    };
    float4 afr[NG];
    auto load_a = [&](int j) {
        if (!active) return;
        const int i = j / NCH, c = j - i * NCH;
        const size_t base = (((size_t)cls_f[i] * (a.MP / 32) + rb) * (KP / 8) + (size_t)c * NG) * 64 + lane;
#pragma unroll
        for (int q = 0; q < NG; ++q) afr[q] = Kf[base + (size_t)q * 64];
    };
    auto store_tile = [&](float *tile) {
        float *dst = tile + col * LDB + part * EPT;
#pragma unroll
        for (int e = 0; e < EPT / 4; ++e) {
            dst[4 * e] = g[e].x;
            dst[4 * e + 1] = g[e].y;
            dst[4 * e + 2] = g[e].z;
            dst[4 * e + 3] = g[e].w;
        }
    };

    f16v acc[NCB];
    double accd[NCB][16];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            acc[cb][r] = 0.0f;
            accd[cb][r] = 0.0;
        }
    }
    const int n = (i1 - i0) * NCH;
    gather(0);
    load_a(0);
    store_tile(Bs);
    __syncthreads();
    for (int i = 0; i < n; ++i) {
        const float *cur = Bs + (i & 1) * BUF;
        float4 acur[NG];
#pragma unroll
        for (int q = 0; q < NG; ++q) acur[q] = afr[q];
        if (i + 1 < n) {
            gather(i + 1);
            load_a(i + 1);
        }
        if (active) {
#pragma unroll
            for (int q = 0; q < NG; ++q) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float av = j == 0 ? acur[q].x : j == 1 ? acur[q].y : j == 2 ? acur[q].z : acur[q].w;
                    const int kk = 2 * (4 * q + j) + (lane >> 5);
#pragma unroll
                    for (int cb = 0; cb < NCB; ++cb) {
                        const float bv = cur[(32 * cb + (lane & 31)) * LDB + kk];
                        acc[cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc[cb], 0, 0, 0);
                    }
                }
            }
        }
        if ((i + 1) % NCH == 0) {  // one offset summed: into FP64
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    accd[cb][r] += (double)acc[cb][r];
                    acc[cb][r] = 0.0f;
                }
        }
        if (i + 1 < n) {
            store_tile(Bs + ((i + 1) & 1) * BUF);
            __syncthreads();
        }
    }
    if (!active) return;
    TW *Wl = task.i1 ? static_cast<TW *>(a.Wpart) + task.o_off : static_cast<TW *>(a.Wl) + task.w_off;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int c = 32 * cb + (lane & 31);
        if (c >= task.np) continue;
        const size_t box = ((size_t)(task.p0 + c) << D) | (unsigned)task.cls;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 32 * rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (row < P) Wl[box * P + row] = (TW)accd[cb][r];
        }
    }
}

template <typename T, int D>
static bool launch_pipe_wide(const M2LKArgs &a, int tb, long ntasks, hipStream_t st)
{
    if constexpr (D == 1) {
        return false;
    } else {
        const unsigned gy = (unsigned)(a.KP + 127) / 128;
        const dim3 grid((unsigned)ntasks * gy);
        if (a.KP == 128 && tb == 32)
            hipLaunchKernelGGL((m2l_pipe_wide<T, D, 32>), grid, dim3(256), 0, st, a);
        else if (a.KP == 128 && tb == 64)
            hipLaunchKernelGGL((m2l_pipe_wide<T, D, 64, 32>), grid, dim3(256), 0, st, a);
        else if (a.KP == 224 && tb == 32)
            hipLaunchKernelGGL((m2l_pipe_wide<T, D, 32, 32, 224>), grid, dim3(256), 0, st, a);
        else if (a.KP == 224 && tb == 64)
            hipLaunchKernelGGL((m2l_pipe_wide<T, D, 64, 32, 224>), grid, dim3(256), 0, st, a);
        else
            return false;
        return true;
    }
}

// BF16x3 form at a padded length of 128 or 256 on the 32x32x16 BF16 MFMA,
// laid out as m2l_pipe_wide (4 waves x 32 rows, TB boxes, chunks of 64 per
// offset, operator and operand loads of step j + 1 during step j).  Per
// 16-wide k-step and column block the three products lo x hi, hi x lo,
// hi x hi go into the FP32 accumulator (as in m2l_pipe); an offset's sum is
// added into FP64 registers.  Operator fragments are FMMA_FRAG_BF16W.
template <typename TW, int D, int TB, int KP = 128>
__global__ __launch_bounds__(256) void m2l_pipe_bfw(M2LKArgs a)
{
    constexpr int KC = 64, NCH = KP / KC, NKC = KC / 16;
    constexpr int TPC = 256 / TB, EPT = KC / TPC, NCB = TB / 32;
    constexpr int LDB = KC + kLdsPad<2>;
    typedef float f16v __attribute__((ext_vector_type(16)));
    static_assert(EPT % 8 == 0, "gather chunks must be whole 16-byte vectors");
    constexpr int BUF = 2 * TB * LDB;  // hi tile, then lo tile
    __shared__ __align__(16) __bf16 Bs[2 * BUF];

    // row chunks of 128 (more than 128 rows: several per task) on neighbouring workgroups of one XCD
    const unsigned gy = (unsigned)(a.MP + 127) / 128;
    const unsigned n8 = gridDim.x / 8u, rem = gridDim.x % 8u, xcd = blockIdx.x % 8u;
    const unsigned v = xcd * n8 + (xcd < rem ? xcd : rem) + blockIdx.x / 8u;
    const unsigned tix = v / gy;
    const M2LTask task = a.tasks[tix];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int rb = (int)(v - tix * gy) * 4 + wave;  // the wave's 32-row block
    const bool active = rb * 32 < a.MP;
    const int N = 1 << task.level, P = a.P;
    const int col = tid / TPC, part = tid % TPC;

    int tc[D];
    const bool cvalid = col < task.np;
    {
        int pc[D];
        morton_decode<D>(task.p0 + (cvalid ? col : 0), pc);
#pragma unroll
        for (int d = 0; d < D; ++d) tc[d] = 2 * pc[d] + ((task.cls >> (D - 1 - d)) & 1);
    }
    const __bf16 *Wh = static_cast<const __bf16 *>(a.Wc) + task.c_off;
    const __bf16 *Wlo = static_cast<const __bf16 *>(a.Wc_lo) + task.c_off;
    const int i0 = task.i1 ? task.i0 : 0, i1 = task.i1 ? task.i1 : a.n_cls;
    const int *cls_f = a.cls_f + task.cls * a.n_cls + i0;
    const int *cls_o = a.cls_o + (task.cls * a.n_cls + i0) * D;
    const bf8_t *Kh = reinterpret_cast<const bf8_t *>(static_cast<const __bf16 *>(a.Kf) + task.k_off);
    const bf8_t *Kl = reinterpret_cast<const bf8_t *>(static_cast<const __bf16 *>(a.Kf_lo) + task.k_off);

    uint4 g[EPT / 8], g_lo[EPT / 8];
    auto gather = [&](int j) {
        const int i = j / NCH, c = j - i * NCH;
        int m[D];
        bool ok = cvalid;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            m[d] = tc[d] + cls_o[i * D + d];
            ok = ok && m[d] >= 0 && m[d] < N;
        }
        const size_t off = (size_t)(ok ? morton_encode<D>(m) : 0u) * KP + c * KC + part * EPT;
        const uint4 *sh = reinterpret_cast<const uint4 *>(Wh + off);
        const uint4 *sl = reinterpret_cast<const uint4 *>(Wlo + off);
#pragma unroll
        for (int e = 0; e < EPT / 8; ++e) {
            g[e] = ok ? sh[e] : uint4{0, 0, 0, 0};
            g_lo[e] = ok ? sl[e] : uint4{0, 0, 0, 0};
        }
    };
    bf8_t ah[NKC], al[NKC];
    auto load_a = [&](int j) {
        if (!active) return;
        const int i = j / NCH, c = j - i * NCH;
        const size_t base = (((size_t)cls_f[i] * (a.MP / 32) + rb) * (KP / 16) + (size_t)c * NKC) * 64 + lane;
#pragma unroll
        for (int s = 0; s < NKC; ++s) {
            ah[s] = Kh[base + (size_t)s * 64];
            al[s] = Kl[base + (size_t)s * 64];
        }
    };
    auto store_tile = [&](__bf16 *tile) {
        __bf16 *dst = tile + col * LDB + part * EPT;
#pragma unroll
        for (int e = 0; e < EPT / 8; ++e) {
            reinterpret_cast<uint4 *>(dst)[e] = g[e];
            reinterpret_cast<uint4 *>(dst + TB * LDB)[e] = g_lo[e];
        }
    };

    f16v acc[NCB];
    double accd[NCB][16];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            acc[cb][r] = 0.0f;
            accd[cb][r] = 0.0;
        }
    }
    const int n = (i1 - i0) * NCH;
    gather(0);
    load_a(0);
    store_tile(Bs);
    __syncthreads();
    for (int i = 0; i < n; ++i) {
        const __bf16 *cur = Bs + (i & 1) * BUF;
        bf8_t ch[NKC], cl[NKC];
#pragma unroll
        for (int s = 0; s < NKC; ++s) {
            ch[s] = ah[s];
            cl[s] = al[s];
        }
        if (i + 1 < n) {
            gather(i + 1);
            load_a(i + 1);
        }
        if (active) {
#pragma unroll
            for (int s = 0; s < NKC; ++s) {
                const int kk = 16 * s + 8 * (lane >> 5);
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) {
                    const int off = (32 * cb + (lane & 31)) * LDB + kk;
                    const bf8_t bh = *reinterpret_cast<const bf8_t *>(cur + off);
                    const bf8_t bl = *reinterpret_cast<const bf8_t *>(cur + TB * LDB + off);
                    acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cl[s], bh, acc[cb], 0, 0, 0);
                    acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ch[s], bl, acc[cb], 0, 0, 0);
                    acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ch[s], bh, acc[cb], 0, 0, 0);
                }
            }
        }
        if ((i + 1) % NCH == 0) {  // one offset summed: into FP64
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    accd[cb][r] += (double)acc[cb][r];
                    acc[cb][r] = 0.0f;
                }
        }
        if (i + 1 < n) {
            store_tile(Bs + ((i + 1) & 1) * BUF);
            __syncthreads();
        }
    }
    if (!active) return;
    TW *Wl = task.i1 ? static_cast<TW *>(a.Wpart) + task.o_off : static_cast<TW *>(a.Wl) + task.w_off;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int c = 32 * cb + (lane & 31);
        if (c >= task.np) continue;
        const size_t box = ((size_t)(task.p0 + c) << D) | (unsigned)task.cls;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 32 * rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (row < P) Wl[box * P + row] = (TW)accd[cb][r];
        }
    }
}

template <typename T, int D>
static bool launch_pipe_bfw(const M2LKArgs &a, int tb, long ntasks, hipStream_t st)
{
    if constexpr (D == 1) {
        return false;
    } else {
        const dim3 grid((unsigned)ntasks * (unsigned)((a.MP + 127) / 128));
        if (a.KP == 128 && tb == 32)
            hipLaunchKernelGGL((m2l_pipe_bfw<T, D, 32>), grid, dim3(256), 0, st, a);
        else if (a.KP == 128 && tb == 64)
            hipLaunchKernelGGL((m2l_pipe_bfw<T, D, 64>), grid, dim3(256), 0, st, a);
        else if (a.KP == 256 && tb == 32)
            hipLaunchKernelGGL((m2l_pipe_bfw<T, D, 32, 256>), grid, dim3(256), 0, st, a);
        else if (a.KP == 256 && tb == 64)
            hipLaunchKernelGGL((m2l_pipe_bfw<T, D, 64, 256>), grid, dim3(256), 0, st, a);
        else
            return false;
        return true;
    }
}

bool m2l_pipe_supported(int D, int KP, int mode, int tb)
{
    if (m2l_wide(mode, KP)) return D >= 2 && (tb == 32 || tb == 64);  // m2l_pipe_wide, m2l_pipe_bfw
    const int vec = mode == 2 ? 8 : mode == 1 ? 4 : 2;
    const bool small = D >= 2 && (KP == 32 || KP == 64 || KP == 128 || (D == 2 && mode != 2 && (KP == 48 || KP == 96)));
    const bool big = D == 3 && (mode == 2 ? (KP == 256 || KP == 384 || KP == 512)
                                          : (KP == 224 || KP == 352 || KP == 512));
    if (!small && !big) return false;
    const int kc = m2l_chunk(KP, mode);
    if ((mode == 2 && KP % 32) || (kc / (256 / tb)) % vec) return false;
    if (tb == 64) return true;
    if (D != 3) return false;
    if (tb == 16 || tb == 32) return small;
    return mode == 0 || (mode == 2 && !small) ? tb == 128 : (tb == 128 || tb == 256);
}

template <typename T, int D>
int m2l_apply(const M2LKArgs &a0, const std::vector<M2LGroup> &groups, int mode, hipStream_t st)
{
    for (const M2LGroup &g : groups) {
        if (g.count <= 0) continue;
        M2LKArgs a = a0;
        a.tasks = a0.tasks + g.first;
        const dim3 grid((unsigned)g.count, (unsigned)((a.MP + 63) / 64));
        bool done = false;
        if (mode == 0) {
            if constexpr (std::is_same<T, double>::value) {
                done = g.pipelined && launch_pipe<T, D, 0>(a, g.tb, grid, st);
                if (!done) {
                    const size_t lds =
                        (size_t)64 * ((a.KP < kM2LChunk ? a.KP : kM2LChunk) + kLdsPad<8>) * sizeof(double);
                    hipLaunchKernelGGL((m2l_mfma4<T, D, 0>), grid, dim3(256), lds, st, a);
                }
            } else {
                return FMMA_ERR_INVALID;
            }
        } else if (mode == 1) {
            if (m2l_wide(mode, a.KP)) {  // the operators are in the 32x32x2 order: no other kernel reads them
                if (!g.pipelined || !launch_pipe_wide<T, D>(a, g.tb, g.count, st)) return FMMA_ERR_UNSUPPORTED;
                FMMA_LAUNCHED("m2l_apply");
                continue;
            }
            done = g.pipelined && launch_pipe<T, D, 1>(a, g.tb, grid, st);
            if (!done) {
                const size_t lds =
                    (size_t)64 * ((a.KP < kM2LChunk ? a.KP : kM2LChunk) + kLdsPad<4>) * sizeof(float);
                hipLaunchKernelGGL((m2l_mfma4<T, D, 1>), grid, dim3(256), lds, st, a);
            }
        } else {
            if (m2l_wide(mode, a.KP)) {  // the operators are in the 32x32x16 order: no other kernel reads them
                if (!g.pipelined || !launch_pipe_bfw<T, D>(a, g.tb, g.count, st)) return FMMA_ERR_UNSUPPORTED;
                FMMA_LAUNCHED("m2l_apply");
                continue;
            }
            done = g.pipelined && launch_pipe<T, D, 2>(a, g.tb, grid, st);
            if (!done) {
                const size_t lds = (size_t)2 * 64 * (a.KP + 8) * sizeof(__bf16);
                hipLaunchKernelGGL((m2l_bf16x3<T, D>), grid, dim3(256), lds, st, a);
            }
        }
        FMMA_LAUNCHED("m2l_apply");
    }
    return FMMA_OK;
}

struct M2LSumArgs {
    int n;
    long start[kMaxLevels + 1];  // element ranges of the split levels, concatenated
    long w_off[kMaxLevels], count[kMaxLevels], part_off[kMaxLevels];
    int parts[kMaxLevels];
};

template <typename T>
__global__ __launch_bounds__(256) void m2l_sum_parts_kernel(T *Wl, const T *Wpart, M2LSumArgs s)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= s.start[s.n]) return;
    int k = 0;
    while (e >= s.start[k + 1]) ++k;
    const long r = e - s.start[k];
    const T *p = Wpart + s.part_off[k] + r;
    T v = p[0];
    for (int q = 1; q < s.parts[k]; ++q) v += p[(long)q * s.count[k]];
    Wl[s.w_off[k] + r] = v;
}

template <typename T>
int m2l_sum_parts(T *Wl, const T *Wpart, const std::vector<M2LSplit> &splits, hipStream_t st)
{
    if (splits.empty()) return FMMA_OK;
    if ((int)splits.size() > kMaxLevels) return FMMA_ERR_INVALID;
    M2LSumArgs s{};
    s.n = (int)splits.size();
    for (int k = 0; k < s.n; ++k) {
        s.start[k + 1] = s.start[k] + splits[k].count;
        s.w_off[k] = splits[k].w_off;
        s.count[k] = splits[k].count;
        s.part_off[k] = splits[k].part_off;
        s.parts[k] = splits[k].parts;
    }
    const unsigned nb = (unsigned)((s.start[s.n] + 255) / 256);
    hipLaunchKernelGGL((m2l_sum_parts_kernel<T>), dim3(nb), dim3(256), 0, st, Wl, Wpart, s);
    FMMA_LAUNCHED("m2l_sum_parts");
    return FMMA_OK;
}

// Multipoles of levels 2..L as MFMA operands, one row of KP (zero padded)
// per box: TYPE (mode 0), FP32 (mode 1), or split BF16 hi = bf16(v),
// lo = bf16(v - hi) of v = (float)W (mode 2).
template <typename T>
__global__ __launch_bounds__(256) void convert_operands(const T *W, long nbox, int P, int KP, int mode, void *out,
                                                        __bf16 *lo)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= nbox * KP) return;
    const long b = e / KP;
    const int k = (int)(e - b * KP);
    const T w = k < P ? W[b * P + k] : T(0);
    if (mode == 0) {
        static_cast<T *>(out)[e] = w;
    } else if (mode == 1) {
        static_cast<float *>(out)[e] = (float)w;
    } else {
        const float v = (float)w;
        const __bf16 h = (__bf16)v;
        static_cast<__bf16 *>(out)[e] = h;
        lo[e] = (__bf16)(v - (float)h);
    }
}

template <typename T>
int m2l_convert(const T *W, long nbox, int P, int KP, int mode, void *Wc, void *Wc_lo, hipStream_t st)
{
    if (nbox <= 0) return FMMA_OK;
    const unsigned nb = (unsigned)((nbox * KP + 255) / 256);
    hipLaunchKernelGGL((convert_operands<T>), dim3(nb), dim3(256), 0, st, W, nbox, P, KP, mode, Wc, (__bf16 *)Wc_lo);
    FMMA_LAUNCHED("m2l_convert");
    return FMMA_OK;
}

template int m2l_sum_parts<double>(double *, const double *, const std::vector<M2LSplit> &, hipStream_t);
template int m2l_sum_parts<float>(float *, const float *, const std::vector<M2LSplit> &, hipStream_t);
template int m2l_convert<double>(const double *, long, int, int, int, void *, void *, hipStream_t);
template int m2l_convert<float>(const float *, long, int, int, int, void *, void *, hipStream_t);
template int m2l_fragments<double>(const double *, int, int, int, void *, void *, hipStream_t);
template int m2l_fragments<float>(const float *, int, int, int, void *, void *, hipStream_t);
template int m2l_apply<double, 1>(const M2LKArgs &, const std::vector<M2LGroup> &, int, hipStream_t);
template int m2l_apply<double, 2>(const M2LKArgs &, const std::vector<M2LGroup> &, int, hipStream_t);
template int m2l_apply<double, 3>(const M2LKArgs &, const std::vector<M2LGroup> &, int, hipStream_t);
template int m2l_apply<float, 1>(const M2LKArgs &, const std::vector<M2LGroup> &, int, hipStream_t);
template int m2l_apply<float, 2>(const M2LKArgs &, const std::vector<M2LGroup> &, int, hipStream_t);
template int m2l_apply<float, 3>(const M2LKArgs &, const std::vector<M2LGroup> &, int, hipStream_t);

}  // namespace lib
}  // namespace fmma
