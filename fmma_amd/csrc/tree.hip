// tree.hip -- bounding cube, leaf keys, stable radix sort, box ranges.
//
// Replaces the reference's serial binning (src/fmm_core.cpp:255-267, which
// pushes every source index into a std::vector per box) and its
// get_minmax/get_origin_and_length (src/fmm_core.cpp:127-173).  Points are
// sorted by Morton leaf key with a stable LSD radix sort, so inside a box
// they keep ascending original index, the order of the reference's lists.
// Box membership is bit-identical to the reference: the same IEEE
// subtraction, division and truncation (leaf_coord in common.hpp).
#include "tree.hpp"

namespace fmma {
namespace lib {

// ------------------------------------------------------------ bounding cube

// Partial min/max per block, and whether the block saw a non-finite
// coordinate (min/max alone would skip a NaN): kBBoxStride<D> values per block.
template <typename T, int D>
__global__ __launch_bounds__(256) void bbox_partial(const T *a, long na, const T *b, long nb, T *part)
{
    bool bad = false;
    T lo[D], hi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lo[d] = T(INFINITY);
        hi[d] = T(-INFINITY);
    }
    const long stride = (long)gridDim.x * 256;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < na; i += stride)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T v = a[i * D + d];
            lo[d] = v < lo[d] ? v : lo[d];
            hi[d] = hi[d] < v ? v : hi[d];
            bad = bad || !__builtin_isfinite(v);
        }
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < nb; i += stride)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T v = b[i * D + d];
            lo[d] = v < lo[d] ? v : lo[d];
            hi[d] = hi[d] < v ? v : hi[d];
            bad = bad || !__builtin_isfinite(v);
        }
    __shared__ int s_bad;
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    if (bad) s_bad = 1;
#pragma unroll
    for (int d = 0; d < D; ++d)
        for (int o = 32; o > 0; o >>= 1) {
            const T l2 = __shfl_xor(lo[d], o), h2 = __shfl_xor(hi[d], o);
            lo[d] = l2 < lo[d] ? l2 : lo[d];
            hi[d] = hi[d] < h2 ? h2 : hi[d];
        }
    __shared__ T s[4][2 * D];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            s[wave][d] = lo[d];
            s[wave][D + d] = hi[d];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w)
#pragma unroll
            for (int d = 0; d < D; ++d) {
                s[0][d] = s[w][d] < s[0][d] ? s[w][d] : s[0][d];
                s[0][D + d] = s[0][D + d] < s[w][D + d] ? s[w][D + d] : s[0][D + d];
            }
#pragma unroll
        for (int d = 0; d < 2 * D; ++d) part[blockIdx.x * kBBoxStride<D> + d] = s[0][d];
        part[blockIdx.x * kBBoxStride<D> + 2 * D] = s_bad ? T(1) : T(0);
    }
}

// Len = max_d (max_d - min_d), origin_d = (max_d + min_d)/2 - Len/2, in TYPE
// (src/fmm_core.cpp:164-171); len_leaf = Len / (TYPE)n_side, n_side the leaf
// boxes per side (2^leaf_level; nrnmm: nrn_N, src/nrnmm.cpp:37).
template <typename T, int D>
__global__ __launch_bounds__(64) void bbox_final(const T *part, int np, Geom<T> *g, int n_side)
{
    T lo[D], hi[D];
    bool bad = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lo[d] = T(INFINITY);
        hi[d] = T(-INFINITY);
    }
    for (int p = threadIdx.x; p < np; p += 64) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T l = part[p * kBBoxStride<D> + d], h = part[p * kBBoxStride<D> + D + d];
            lo[d] = l < lo[d] ? l : lo[d];
            hi[d] = hi[d] < h ? h : hi[d];
        }
        bad = bad || part[p * kBBoxStride<D> + 2 * D] != T(0);
    }
    bad = __any(bad);
#pragma unroll
    for (int d = 0; d < D; ++d)
        for (int o = 32; o > 0; o >>= 1) {
            const T l2 = __shfl_xor(lo[d], o), h2 = __shfl_xor(hi[d], o);
            lo[d] = l2 < lo[d] ? l2 : lo[d];
            hi[d] = hi[d] < h2 ? h2 : hi[d];
        }
    if (threadIdx.x == 0) {
#pragma clang fp contract(off)
        T Len = T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T e = hi[d] - lo[d];
            Len = Len < e ? e : Len;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) g->origin[d] = (hi[d] + lo[d]) / 2 - Len / 2;
        for (int d = D; d < 3; ++d) g->origin[d] = T(0);
        g->Len = Len;
        g->len_leaf = Len / (T)n_side;
        g->status = bad ? kGeomNonFinite : Len == T(0) ? kGeomZeroLen : kGeomOk;
    }
}

template <typename T, int D>
int bbox(const T *t, long nt, const T *s, long ns, T *part, Geom<T> *g, int n_side, hipStream_t st)
{
    if (t == s && nt == ns) ns = 0;  // one point set: read it once
    const long n = nt > ns ? nt : ns;
    int nb = (int)((n + 255) / 256);
    nb = nb < kBBoxParts ? nb : kBBoxParts;
    nb = nb < 1 ? 1 : nb;
    hipLaunchKernelGGL((bbox_partial<T, D>), dim3(nb), dim3(256), 0, st, t, nt, s, ns, part);
    hipLaunchKernelGGL((bbox_final<T, D>), dim3(1), dim3(64), 0, st, part, nb, g, n_side);
    FMMA_LAUNCHED("bbox");
    return FMMA_OK;
}

// ---------------------------------------------------------------- leaf keys

// --------------------------------------------------------------- radix sort
//
// LSD passes of 8-bit digits over tiles of 4096 keys.  rs_hist counts digits
// per tile (the first pass's counts come from the key kernel), scan_exclusive
// turns the digit-major [256][tiles] table into scatter bases, and
// rs_scatter_points ranks the tile's keys stably and moves every array of
// the points to its sorted place.

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;

template <typename T, int D>
__device__ __forceinline__ unsigned leaf_key(const T *pts, long i, const Geom<T> *g, int NL, bool row_major)
{
    int b[D];
#pragma unroll
    for (int d = 0; d < D; ++d) b[d] = leaf_coord(pts[i * D + d], g->origin[d], g->len_leaf, NL);
    if (row_major) {
        unsigned id = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) id = id * (unsigned)NL + (unsigned)b[d];
        return id;
    }
    return morton_encode<D>(b);
}

// One workgroup per sort tile: the keys of its 4096 points and, with hist,
// the tile's counts of the lowest digit (the first sort pass's histogram).
template <typename T, int D>
__global__ __launch_bounds__(256) void keys_kernel(const T *pts, long n, const Geom<T> *g, int leaf_level,
                                                   int grid_n, unsigned *keys, unsigned *hist, unsigned ntiles)
{
    __shared__ unsigned h[256];
    if (hist) h[threadIdx.x] = 0u;
    __syncthreads();
    const int NL = grid_n > 0 ? grid_n : 1 << leaf_level;
    const long tile0 = (long)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ITEMS; ++r) {
        const long i = tile0 + r * RS_THREADS + threadIdx.x;
        if (i < n) {
            const unsigned k = leaf_key<T, D>(pts, i, g, NL, grid_n > 0);
            keys[i] = k;
            if (hist) atomicAdd(&h[k & 255u], 1u);
        }
    }
    if (hist) {
        __syncthreads();
        hist[(long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
    }
}

template <typename T, int D>
int leaf_keys(const T *pts, long n, const Geom<T> *g, int leaf_level, int grid_n, unsigned *keys, hipStream_t st,
              unsigned *hist0)
{
    if (n <= 0) return FMMA_OK;
    const unsigned ntiles = (unsigned)((n + RS_TILE - 1) / RS_TILE);
    hipLaunchKernelGGL((keys_kernel<T, D>), dim3(ntiles), dim3(256), 0, st, pts, n, g, leaf_level, grid_n, keys,
                       hist0, ntiles);
    FMMA_LAUNCHED("leaf_keys");
    return FMMA_OK;
}

__global__ __launch_bounds__(256) void rs_hist(const unsigned *keys, long n, int shift, unsigned *hist,
                                               unsigned ntiles)
{
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const long base = (long)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ITEMS; ++r) {
        const long i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive prefix sum of m values in place, in three launches: per-chunk
// sums, a scan of those sums, then a scan of every chunk from its offset.
constexpr int SCAN_BLOCKS = 256;

// Inclusive scan of v over a workgroup of 256 threads; returns this thread's
// inclusive value and the block total.
__device__ __forceinline__ unsigned block_scan256(unsigned v, unsigned &total, unsigned *wsum)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    unsigned before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const unsigned ws = wsum[w];
        if (w < wave) before += ws;
        total += ws;
    }
    __syncthreads();
    return before + x;
}

__global__ __launch_bounds__(256) void scan_reduce(const unsigned *a, long m, long chunk, unsigned *part)
{
    __shared__ unsigned wsum[4];
    const long b0 = (long)blockIdx.x * chunk, b1 = b0 + chunk < m ? b0 + chunk : m;
    unsigned s = 0;
    for (long i = b0 + threadIdx.x; i < b1; i += 256) s += a[i];
    unsigned total;
    block_scan256(s, total, wsum);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void scan_parts(unsigned *part, int np)
{
    __shared__ unsigned wsum[4];
    const unsigned v = (int)threadIdx.x < np ? part[threadIdx.x] : 0u;
    unsigned total;
    const unsigned inc = block_scan256(v, total, wsum);
    if ((int)threadIdx.x < np) part[threadIdx.x] = inc - v;
}

__global__ __launch_bounds__(256) void scan_down(unsigned *a, long m, long chunk, const unsigned *part)
{
    __shared__ unsigned wsum[4];
    const long b0 = (long)blockIdx.x * chunk, b1 = b0 + chunk < m ? b0 + chunk : m;
    unsigned carry = part[blockIdx.x];
    for (long base = b0; base < b1; base += 1024) {
        const long i0 = base + threadIdx.x * 4;
        unsigned v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = i0 + q < b1 ? a[i0 + q] : 0u;
        const unsigned s = v[0] + v[1] + v[2] + v[3];
        unsigned total;
        const unsigned inc = block_scan256(s, total, wsum);
        unsigned e = carry + inc - s;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (i0 + q < b1) a[i0 + q] = e;
            e += v[q];
        }
        carry += total;
    }
}

static int scan_exclusive(unsigned *a, long m, unsigned *part, hipStream_t st)
{
    long chunk = (m + SCAN_BLOCKS - 1) / SCAN_BLOCKS;
    chunk = (chunk + 1023) / 1024 * 1024;
    const int nb = (int)((m + chunk - 1) / chunk);
    hipLaunchKernelGGL(scan_reduce, dim3(nb), dim3(256), 0, st, a, m, chunk, part);
    hipLaunchKernelGGL(scan_parts, dim3(1), dim3(256), 0, st, part, nb);
    hipLaunchKernelGGL(scan_down, dim3(nb), dim3(256), 0, st, a, m, chunk, part);
    FMMA_LAUNCHED("scan_exclusive");
    return FMMA_OK;
}

// One pass of the point sort.  Every value of a thread's 16 points (key,
// index, coordinates, weight) is loaded at once, so the loads overlap the
// ranking.  A tile of 4096 points is ranked stably by the 8-bit digit at
// `shift`: wave w ranks its contiguous quarter of the tile in 16 rounds of
// 64 against counters of its own (ballots find the lanes of a round that
// share a digit), so no barrier is needed until the four waves' counts are
// combined.  Every array then goes through one LDS staging tile in sorted
// order, and lane j writes staged element j to its digit's global run: the
// writes of a run are consecutive lanes.  The first pass reads the caller's
// points (AoS coordinates, weights or 1, the index as the value); later
// passes read the previous pass's SoA output.
template <typename T, int D, bool FIRST, bool WITH_W>
__global__ __launch_bounds__(256) void rs_scatter_points(PointPass<T> a)
{
    constexpr int NV = WITH_W ? D + 1 : D;
    __shared__ unsigned wcnt[4][256];
    __shared__ unsigned loc[256], gdel[256], wsum[4];
    __shared__ unsigned skey[RS_TILE];
    __shared__ __align__(16) T stage[RS_TILE];
    unsigned *stage_u = reinterpret_cast<unsigned *>(stage);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // Workgroups are dealt to the 8 XCDs round-robin; XCD x takes a
    // contiguous run of tiles, so the runs of one digit that neighbouring
    // tiles write side by side meet in one L2 and leave it as whole lines.
    const unsigned n8 = gridDim.x / 8u, rem = gridDim.x % 8u, xcd = blockIdx.x % 8u;
    const unsigned tile = xcd * n8 + (xcd < rem ? xcd : rem) + blockIdx.x / 8u;
    const long n = a.n, tile0 = (long)tile * RS_TILE;
    const int cnt = (int)(n - tile0 < RS_TILE ? n - tile0 : RS_TILE);
    const long my0 = tile0 + (long)wave * (RS_TILE / 4) + lane;

    unsigned key[RS_ITEMS], idx[RS_ITEMS];
    T val[NV][RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const long i = my0 + r * 64;
        const bool valid = i < n;
        key[r] = valid ? a.kin[i] : 0u;
        idx[r] = FIRST ? (unsigned)i : valid ? a.vin[i] : 0u;
#pragma unroll
        for (int d = 0; d < NV; ++d) {
            if (!valid)
                val[d][r] = T(0);
            else if (d < D)
                val[d][r] = FIRST ? a.pin[i * D + d] : a.pin[(long)d * n + i];
            else
                val[d][r] = FIRST ? (a.win ? a.win[i] : T(1)) : a.win[i];
        }
    }

    for (int j = lane; j < 256; j += 64) wcnt[wave][j] = 0u;
    const unsigned long long below = __lanemask_lt();
    unsigned lp[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const bool valid = my0 + r * 64 < n;
        const unsigned dig = (key[r] >> a.shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dig >> b) & 1u;
            const unsigned long long bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const unsigned rank = __popcll(peers & below);
        const unsigned pre = wcnt[wave][dig];
        if (valid && rank == 0) wcnt[wave][dig] = pre + (unsigned)__popcll(peers);
        lp[r] = pre + rank;
    }
    __syncthreads();
    {  // thread tid owns digit tid: wave prefixes, the tile-local start, the global start
        const unsigned c0 = wcnt[0][tid], c1 = wcnt[1][tid], c2 = wcnt[2][tid], c3 = wcnt[3][tid];
        const unsigned tot = c0 + c1 + c2 + c3;
        wcnt[0][tid] = 0u;
        wcnt[1][tid] = c0;
        wcnt[2][tid] = c0 + c1;
        wcnt[3][tid] = c0 + c1 + c2;
        unsigned all;
        const unsigned start = block_scan256(tot, all, wsum) - tot;
        loc[tid] = start;
        gdel[tid] = a.offs[(long)tid * a.ntiles + tile] - start;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        if (my0 + r * 64 < n) {
            const unsigned dig = (key[r] >> a.shift) & 255u;
            lp[r] += loc[dig] + wcnt[wave][dig];
            skey[lp[r]] = key[r];
            stage_u[lp[r]] = idx[r];
        }
    }
    __syncthreads();
    unsigned gout[RS_ITEMS];  // global place of staged element r * 256 + tid
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int j = r * 256 + tid;
        if (j < cnt) {
            const unsigned k = skey[j];
            gout[r] = gdel[(k >> a.shift) & 255u] + (unsigned)j;
            a.kout[gout[r]] = k;
            a.vout[gout[r]] = stage_u[j];
        }
    }
#pragma unroll
    for (int d = 0; d < NV; ++d) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RS_ITEMS; ++r)
            if (my0 + r * 64 < n) stage[lp[r]] = val[d][r];
        __syncthreads();
        T *out = d < D ? a.pout + (long)d * n : a.wout;
#pragma unroll
        for (int r = 0; r < RS_ITEMS; ++r)
            if (r * 256 + tid < cnt) out[gout[r]] = stage[r * 256 + tid];
    }
}

__global__ __launch_bounds__(256) void iota_kernel(unsigned *v, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = (unsigned)i;
}

template <typename T, int D>
int sort_points(PointSortBuffers<T> &b, long n, int bits, const T *pts, const T *w, bool with_w, hipStream_t st,
                bool hist0_ready)
{
    if (n <= 0) return FMMA_OK;
    if (bits <= 0) {  // one box: the identity permutation
        hipLaunchKernelGGL(iota_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, b.vals, n);
        FMMA_LAUNCHED("iota");
        return gather<T, D>(pts, w, b.vals, n, b.pos, with_w ? b.ws : nullptr, st);
    }
    const unsigned ntiles = (unsigned)((n + RS_TILE - 1) / RS_TILE);
    const int passes = (bits + 7) / 8;
    // keys and indices ping-pong from (keys, vals) and end in the pair the
    // last pass wrote; coordinates and weights start from the caller's
    // arrays and alternate so that the last pass writes (pos, ws)
    unsigned *kin = b.keys, *kout = b.keys_alt, *vin = nullptr, *vout = b.vals_alt;
    const T *pin = pts, *win = w;
    for (int p = 0; p < passes; ++p) {
        const bool to_final = (passes - 1 - p) % 2 == 0;
        T *pout = to_final ? b.pos : b.pos_alt, *wout = to_final ? b.ws : b.ws_alt;
        const int shift = 8 * p;
        if (p > 0 || !hist0_ready)
            hipLaunchKernelGGL(rs_hist, dim3(ntiles), dim3(256), 0, st, kin, n, shift, b.hist, ntiles);
        int rc = scan_exclusive(b.hist, (long)ntiles * 256, b.hist + radix_hist_words(n) - SCAN_BLOCKS, st);
        if (rc) return rc;
        PointPass<T> a{kin, vin, kout, vout, pin, win, pout, wout, n, shift, b.hist, ntiles};
        if (p == 0) {
            if (with_w)
                hipLaunchKernelGGL((rs_scatter_points<T, D, true, true>), dim3(ntiles), dim3(256), 0, st, a);
            else
                hipLaunchKernelGGL((rs_scatter_points<T, D, true, false>), dim3(ntiles), dim3(256), 0, st, a);
        } else {
            if (with_w)
                hipLaunchKernelGGL((rs_scatter_points<T, D, false, true>), dim3(ntiles), dim3(256), 0, st, a);
            else
                hipLaunchKernelGGL((rs_scatter_points<T, D, false, false>), dim3(ntiles), dim3(256), 0, st, a);
        }
        FMMA_LAUNCHED("sort_points");
        unsigned *kprev = kin;
        kin = kout;
        vin = vout;
        kout = kprev;
        vout = (vout == b.vals_alt) ? b.vals : b.vals_alt;
        pin = pout;
        win = wout;
    }
    if (kin != b.keys) FMMA_HIP(hipMemcpyAsync(b.keys, kin, n * sizeof(unsigned), hipMemcpyDeviceToDevice, st));
    if (vin != b.vals) FMMA_HIP(hipMemcpyAsync(b.vals, vin, n * sizeof(unsigned), hipMemcpyDeviceToDevice, st));
    return FMMA_OK;
}

size_t radix_hist_words(long n)
{
    const long ntiles = (n + RS_TILE - 1) / RS_TILE;
    return (size_t)(ntiles > 0 ? ntiles : 1) * 256 + SCAN_BLOCKS;  // + scan partials
}

// --------------------------------------------------------------- box ranges

// start[k] = first sorted position whose key is >= k, for k in [0, nbox].
__global__ __launch_bounds__(256) void ranges_kernel(const unsigned *keys, long n, unsigned nbox, unsigned *start)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i > n) return;
    const long kprev = i == 0 ? -1 : (long)keys[i - 1];
    const long kcur = i == n ? (long)nbox : (long)keys[i];
    for (long k = kprev + 1; k <= kcur; ++k) start[k] = (unsigned)i;
}

int box_ranges(const unsigned *keys, long n, unsigned nbox, unsigned *start, hipStream_t st)
{
    hipLaunchKernelGGL(ranges_kernel, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, st, keys, n, nbox,
                       start);
    FMMA_LAUNCHED("box_ranges");
    return FMMA_OK;
}

// ------------------------------------------------------------------ gather

template <typename T, int D>
__global__ __launch_bounds__(256) void gather_kernel(const T *pts, const T *w, const unsigned *perm, long n, T *pos,
                                                     T *ws)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const unsigned p = perm[i];
#pragma unroll
    for (int d = 0; d < D; ++d) pos[d * n + i] = pts[(long)p * D + d];
    if (ws) ws[i] = w ? w[p] : T(1);
}

template <typename T, int D>
int gather(const T *pts, const T *w, const unsigned *perm, long n, T *pos, T *ws, hipStream_t st)
{
    if (n <= 0) return FMMA_OK;
    hipLaunchKernelGGL((gather_kernel<T, D>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, pts, w, perm,
                       n, pos, ws);
    FMMA_LAUNCHED("gather");
    return FMMA_OK;
}

#define FMMA_TREE_INST(T, D)                                                                                    \
    template int bbox<T, D>(const T *, long, const T *, long, T *, Geom<T> *, int, hipStream_t);                \
    template int leaf_keys<T, D>(const T *, long, const Geom<T> *, int, int, unsigned *, hipStream_t, unsigned *); \
    template int gather<T, D>(const T *, const T *, const unsigned *, long, T *, T *, hipStream_t);              \
    template int sort_points<T, D>(PointSortBuffers<T> &, long, int, const T *, const T *, bool, hipStream_t, bool);
FMMA_TREE_INST(double, 1)
FMMA_TREE_INST(double, 2)
FMMA_TREE_INST(double, 3)
FMMA_TREE_INST(float, 1)
FMMA_TREE_INST(float, 2)
FMMA_TREE_INST(float, 3)

}  // namespace lib
}  // namespace fmma
