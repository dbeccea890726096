/*
 * fmm_oracle.c -- CPU restatement of the reference black-box FMM.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * fmma_amd/csrc: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it, and never as the thing that is measured or shipped.
 *
 * It restates, in plain C with OpenMP, the algorithm of fockl/FMMA's `fmm` and
 * `exact` solvers (uniform 2^D tree, Chebyshev interpolation, Fong & Darve
 * 2009), loop for loop, so that its results agree with the reference to
 * rounding (<=1e-13 relative) on the same inputs:
 *
 *   oracle_nodes / cheb_T / cheb_S   src/math.cpp:14-50, src/fmm_core.cpp:240-253
 *   bbox                             src/fmm_core.cpp:127-173
 *   leaf binning                     src/fmm_core.cpp:255-267, 595-596; src/fmm.cpp:126-131
 *   P2M                              src/fmm_core.cpp:224-283
 *   M2M                              src/fmm_core.cpp:295-363
 *   M2L (translation-symmetric)      src/fmm_core.cpp:383-455
 *   M2L (general)                    src/fmm_core.cpp:456-492
 *   interaction list                 src/fmm_core.cpp:16-63
 *   neighbour list                   src/fmm_core.cpp:81-115
 *   L2L                              src/fmm_core.cpp:507-568
 *   L2P                              src/fmm_core.cpp:580-615
 *   P2P and stage order              src/fmm.cpp:15-144
 *   exact                            src/fmma.cpp:79-94
 *   tree (M2P per level)             src/tree.cpp:12-90, M2P src/fmm_core.cpp:625-661
 *   nrnmm (one grid, all boxes)      src/nrnmm.cpp:13-100
 *
 * Differentiation (no counterpart in the reference; the gradient is that of
 * the approximation above): oracle_kernel_grad (each kernel's formula
 * differentiated), oracle_cheb_dS (the L2P basis of src/fmm_core.cpp:602-612
 * differentiated), oracle_fmm_grad, oracle_exact_grad.  Pinned by central
 * differences of oracle_kernel, oracle_cheb_S, oracle_exact and oracle_fmm.
 *
 * Pinning (tests/test_oracle.py): Chebyshev closed forms T_0..T_3, discrete
 * orthogonality S_n(c_i,c_j)=delta_ij, interpolation exactness for
 * polynomials of degree < n, M2M(P2M(child)) == P2M(parent), the L2L/M2M
 * adjoint identity, closed-form interaction-list counts, brute-force exact
 * sums, and the compiled reference (oracle/_ref) at 1e-13 on seeded inputs.
 *
 * Everything is double precision.  Box ids here are the reference's row-major
 * ids (dimension 0 most significant, src/fmm_core.cpp:205-212).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fmm_oracle.h"

#define OR_MAXD 3

/* ------------------------------------------------------------------ kernels */

/* The kernel table shared with the HIP library (include/fmma_hip.h).  Each
 * entry is written the way the reference's default kernel is written
 * (include/fmma/fmma.hpp:55-63): squared differences summed in dimension
 * order, then one libm call. */
double oracle_kernel(int id, const double *prm, int dim, const double *x, const double *y)
{
    double r2 = 0.0;
    switch (id) {
    case ORACLE_K_INV_R:
        for (int d = 0; d < dim; ++d) { double t = x[d] - y[d]; r2 += t * t; }
        return 1.0 / sqrt(r2);
    case ORACLE_K_INV_R_REG:
        for (int d = 0; d < dim; ++d) { double t = x[d] - y[d]; r2 += t * t; }
        return r2 == 0.0 ? 0.0 : 1.0 / sqrt(r2);
    case ORACLE_K_INV_DIFF:
        return 1.0 / (x[0] - y[0]);
    case ORACLE_K_LOG_R:
        for (int d = 0; d < dim; ++d) { double t = x[d] - y[d]; r2 += t * t; }
        return r2 == 0.0 ? 0.0 : 0.5 * log(r2);
    case ORACLE_K_GAUSS: {
        for (int d = 0; d < dim; ++d) { double t = x[d] - y[d]; r2 += t * t; }
        double h = prm[0];
        return exp(-r2 / (2.0 * h * h));
    }
    case ORACLE_K_YUKAWA: {
        for (int d = 0; d < dim; ++d) { double t = x[d] - y[d]; r2 += t * t; }
        if (r2 == 0.0) return 0.0;
        double r = sqrt(r2);
        return exp(-prm[0] * r) / r;
    }
    case ORACLE_K_SUM_R2: {
        double s = 0.0;
        for (int d = 0; d < dim; ++d) { double t = x[d] - y[d]; r2 += t * t; s += t; }
        return s / r2;
    }
    case ORACLE_K_SUM2_R2: {
        double s2 = 0.0;
        for (int d = 0; d < dim; ++d) {
            double t = x[d] - y[d], u = x[d] + y[d];
            r2 += t * t; s2 += u * u;
        }
        return s2 / r2;
    }
    case ORACLE_K_INV_R2:
        for (int d = 0; d < dim; ++d) { double t = x[d] - y[d]; r2 += t * t; }
        return 1.0 / r2;
    default:
        fprintf(stderr, "oracle: unknown kernel id %d\n", id);
        abort();
    }
}

/* grad_x f(x, y) of each kernel, written from its formula (differentiation
 * is new; the reference has none).  Kernels with an r = 0 special case take
 * the gradient as 0 there. */
void oracle_kernel_grad(int id, const double *prm, int dim, const double *x, const double *y, double *g)
{
    double t[OR_MAXD], r2 = 0.0;
    for (int d = 0; d < dim; ++d) { t[d] = x[d] - y[d]; r2 += t[d] * t[d]; }
    double c = 0.0;
    switch (id) {
    case ORACLE_K_INV_R:
    case ORACLE_K_INV_R_REG:
        c = r2 == 0.0 ? 0.0 : -1.0 / (r2 * sqrt(r2));
        break;
    case ORACLE_K_INV_DIFF:
        for (int d = 0; d < dim; ++d) g[d] = 0.0;
        g[0] = -1.0 / (t[0] * t[0]);
        return;
    case ORACLE_K_LOG_R:
        c = r2 == 0.0 ? 0.0 : 1.0 / r2;
        break;
    case ORACLE_K_GAUSS: {
        double h = prm[0];
        c = -exp(-r2 / (2.0 * h * h)) / (h * h);
        break;
    }
    case ORACLE_K_YUKAWA: {
        if (r2 == 0.0) { c = 0.0; break; }
        double r = sqrt(r2), k = prm[0];
        c = -(1.0 + k * r) * exp(-k * r) / (r2 * r);
        break;
    }
    case ORACLE_K_SUM_R2: {
        double s = 0.0;
        for (int d = 0; d < dim; ++d) s += t[d];
        for (int d = 0; d < dim; ++d) g[d] = 1.0 / r2 - 2.0 * s * t[d] / (r2 * r2);
        return;
    }
    case ORACLE_K_SUM2_R2: {
        double s2 = 0.0;
        for (int d = 0; d < dim; ++d) { double u = x[d] + y[d]; s2 += u * u; }
        for (int d = 0; d < dim; ++d) g[d] = 2.0 * (x[d] + y[d]) / r2 - 2.0 * s2 * t[d] / (r2 * r2);
        return;
    }
    case ORACLE_K_INV_R2:
        c = -2.0 / (r2 * r2);
        break;
    default:
        fprintf(stderr, "oracle: unknown kernel id %d\n", id);
        abort();
    }
    for (int d = 0; d < dim; ++d) g[d] = c * t[d];
}

/* ---------------------------------------------------------------- Chebyshev */

/* T_j(x) = cos(j acos x), defined on [-1,1] only (src/math.cpp:14-26). */
double oracle_cheb_T(int j, double x)
{
    if (x < -1.0 || x > 1.0 || j < 0) {
        fprintf(stderr, "oracle: Chebyshev argument out of range (%d, %.17g)\n", j, x);
        abort();
    }
    return cos(j * acos(x));
}

/* S_n(x,y) = 1/n + sum_{j=1}^{n-1} (2/n) T_j(x) T_j(y) (src/math.cpp:31-50),
 * accumulated in the same left-to-right order. */
double oracle_cheb_S(int n, double x, double y)
{
    double s = 1.0 / n;
    for (int j = 1; j < n; ++j) s += 2.0 / n * oracle_cheb_T(j, x) * oracle_cheb_T(j, y);
    return s;
}

/* d/dx S_n(x, y) = (2/n) sum_{j=1}^{n-1} T_j'(x) T_j(y), T_j' = j U_{j-1} with
 * U_0 = 1, U_1 = 2x, U_k = 2x U_{k-1} - U_{k-2}: the basis of L2P
 * (src/fmm_core.cpp:602-612) differentiated in its first argument. */
double oracle_cheb_dS(int n, double x, double y)
{
    double s = 0.0, um = 0.0, u = 1.0;  /* U_{j-2}, U_{j-1} */
    for (int j = 1; j < n; ++j) {
        s += 2.0 / n * (j * u) * oracle_cheb_T(j, y);
        double un = 2.0 * x * u - um;
        um = u;
        u = un;
    }
    return s;
}

/* Chebyshev nodes c_k = cos((2k+1) pi / (2n)), k = 0..n-1 (src/fmm_core.cpp:240-243). */
void oracle_nodes(int n, double *c)
{
    int p = n - 1;
    for (int k = 0; k <= p; ++k) c[k] = cos((2.0 * k + 1.0) / (2 * p + 2) * M_PI);
}

/* ------------------------------------------------------------------- tree */

typedef struct {
    int dim, n, P, depth, L, NL;
    double origin[OR_MAXD], Len;
    double c[32];
    int *digit;      /* P x dim: node index of multi-index k along dim d (dim 0 most significant) */
} otree_t;

static long ipow(long b, int e) { long r = 1; while (e-- > 0) r *= b; return r; }

/* get_minmax + get_origin_and_length (src/fmm_core.cpp:127-173): the cube
 * over targets then sources. */
void oracle_bbox(int dim, long nt, const double *t, long ns, const double *s,
                 double *origin, double *Len)
{
    double lo[OR_MAXD], hi[OR_MAXD];
    const double *first = nt > 0 ? t : s;
    for (int d = 0; d < dim; ++d) lo[d] = hi[d] = first[d];
    for (long i = 0; i < nt; ++i)
        for (int d = 0; d < dim; ++d) {
            double v = t[i * dim + d];
            if (v < lo[d]) lo[d] = v;
            if (hi[d] < v) hi[d] = v;
        }
    for (long i = 0; i < ns; ++i)
        for (int d = 0; d < dim; ++d) {
            double v = s[i * dim + d];
            if (v < lo[d]) lo[d] = v;
            if (hi[d] < v) hi[d] = v;
        }
    double L = 0.0;
    for (int d = 0; d < dim; ++d) { double e = hi[d] - lo[d]; if (L < e) L = e; }
    for (int d = 0; d < dim; ++d) origin[d] = (hi[d] + lo[d]) / 2 - L / 2;
    *Len = L;
}

/* Leaf coordinate of one point: b = min((int)((x-o)/len), N-1), and the
 * clamped relative coordinate 2((x-o)/len - b) - 1 (src/fmm_core.cpp:261,264). */
static inline int bin1(double x, double o, double len, int N, double *rel)
{
    double q = (x - o) / len;
    int b = (int)q;
    if (b > N - 1) b = N - 1;
    if (rel) {
        double r = 2.0 * (q - b) - 1.0;
        if (r > 1.0) r = 1.0;
        if (r < -1.0) r = -1.0;
        *rel = r;
    }
    return b;
}

static long box_id(int dim, const int *b, long N)
{
    long id = 0;
    for (int d = 0; d < dim; ++d) id = id * N + b[d];
    return id;
}

static void box_coords(int dim, long id, long N, int *b)
{
    for (int d = dim - 1; d >= 0; --d) { b[d] = (int)(id % N); id /= N; }
}

int oracle_leaf_boxes(int dim, int depth, long nt, const double *t, long ns, const double *s,
                      long np, const double *pts, long *box, double *rel)
{
    double origin[OR_MAXD], Len;
    oracle_bbox(dim, nt, t, ns, s, origin, &Len);
    long NL = 1L << (depth - 1);
    double len = Len / (double)NL;
    for (long i = 0; i < np; ++i) {
        int b[OR_MAXD];
        for (int d = 0; d < dim; ++d)
            b[d] = bin1(pts[i * dim + d], origin[d], len, (int)NL, rel ? &rel[i * dim + d] : NULL);
        box[i] = box_id(dim, b, NL);
    }
    return 0;
}

/* Single-precision version of the binning, so that float inputs can be
 * checked for identical box membership.  The reference's float instantiation
 * computes (x-o)/len in float (src/fmm_core.cpp:261). */
int oracle_leaf_boxes_f32(int dim, int depth, long nt, const float *t, long ns, const float *s,
                          long np, const float *pts, long *box)
{
    float lo[OR_MAXD], hi[OR_MAXD];
    const float *first = nt > 0 ? t : s;
    for (int d = 0; d < dim; ++d) lo[d] = hi[d] = first[d];
    for (long i = 0; i < nt; ++i)
        for (int d = 0; d < dim; ++d) {
            float v = t[i * dim + d];
            if (v < lo[d]) lo[d] = v;
            if (hi[d] < v) hi[d] = v;
        }
    for (long i = 0; i < ns; ++i)
        for (int d = 0; d < dim; ++d) {
            float v = s[i * dim + d];
            if (v < lo[d]) lo[d] = v;
            if (hi[d] < v) hi[d] = v;
        }
    float L = 0.0f;
    for (int d = 0; d < dim; ++d) { float e = hi[d] - lo[d]; if (L < e) L = e; }
    float origin[OR_MAXD];
    for (int d = 0; d < dim; ++d) origin[d] = (hi[d] + lo[d]) / 2 - L / 2;
    long NL = 1L << (depth - 1);
    float len = L / (float)NL;
    for (long i = 0; i < np; ++i) {
        long id = 0;
        for (int d = 0; d < dim; ++d) {
            float q = (pts[i * dim + d] - origin[d]) / len;
            int b = (int)q;
            if (b > (int)NL - 1) b = (int)NL - 1;
            id = id * NL + b;
        }
        box[i] = id;
    }
    return 0;
}

/* ---------------------------------------------------------------- operators */

static void tree_init(otree_t *T, int dim, int poly_ord, int depth)
{
    T->dim = dim; T->n = poly_ord + 1; T->depth = depth; T->L = depth - 1;
    T->NL = 1 << (depth - 1);
    T->P = (int)ipow(T->n, dim);
    oracle_nodes(T->n, T->c);
    T->digit = (int *)malloc(sizeof(int) * T->P * dim);
    for (int k = 0; k < T->P; ++k) {
        int kk = k;
        for (int d = dim - 1; d >= 0; --d) { T->digit[k * dim + d] = kk % T->n; kk /= T->n; }
    }
}

/* P2M at the leaf level (src/fmm_core.cpp:224-283).  Also returns the
 * per-box source lists, in ascending source index, used later by P2P. */
static void p2m(const otree_t *T, long ns, const double *w, const double *s,
                long **lst_start, long **lst, double *Wm)
{
    int dim = T->dim, P = T->P, n = T->n;
    long nbox = ipow(T->NL, dim);
    double len = T->Len / (double)T->NL;
    long *bx = (long *)malloc(sizeof(long) * (ns > 0 ? ns : 1));
    double *rel = (double *)malloc(sizeof(double) * (ns > 0 ? ns : 1) * dim);
    long *cnt = (long *)calloc(nbox + 1, sizeof(long));
    for (long i = 0; i < ns; ++i) {
        int b[OR_MAXD];
        for (int d = 0; d < dim; ++d) b[d] = bin1(s[i * dim + d], T->origin[d], len, T->NL, &rel[i * dim + d]);
        bx[i] = box_id(dim, b, T->NL);
        cnt[bx[i] + 1]++;
    }
    for (long b = 0; b < nbox; ++b) cnt[b + 1] += cnt[b];
    long *fill = (long *)malloc(sizeof(long) * (nbox + 1));
    memcpy(fill, cnt, sizeof(long) * (nbox + 1));
    long *L = (long *)malloc(sizeof(long) * (ns > 0 ? ns : 1));
    for (long i = 0; i < ns; ++i) L[fill[bx[i]]++] = i;
    free(fill);

#pragma omp parallel for schedule(dynamic, 16)
    for (long b = 0; b < nbox; ++b) {
        double *W = Wm + b * P;
        for (int k = 0; k < P; ++k) W[k] = 0.0;
        for (long j = cnt[b]; j < cnt[b + 1]; ++j) {
            long i = L[j];
            for (int k = 0; k < P; ++k) {
                double v = 1.0;
                for (int d = 0; d < dim; ++d)
                    v *= oracle_cheb_S(n, T->c[T->digit[k * dim + d]], rel[i * dim + d]);
                W[k] += w[i] * v;
            }
        }
    }
    free(bx); free(rel);
    *lst_start = cnt; *lst = L;
}

/* Child-to-parent transfer table (src/fmm_core.cpp:319-336), indexed
 * [child parity bit][k_parent][k_child] per dimension. */
static void m2m_table(const otree_t *T, double *tab /* 2*n*n */)
{
    int n = T->n;
    for (int s = 0; s < 2; ++s)
        for (int ko = 0; ko < n; ++ko)
            for (int ki = 0; ki < n; ++ki) {
                int shift = 2 * s - 1;
                tab[(s * n + ko) * n + ki] = oracle_cheb_S(n, T->c[ko], (T->c[ki] + shift) / 2);
            }
}

/* M2M from the level with N boxes per side to N/2 (src/fmm_core.cpp:295-363);
 * children are visited in row-major order like the reference. */
static void m2m(const otree_t *T, long N, const double *tab, const double *Win, double *Wout)
{
    int dim = T->dim, P = T->P, n = T->n;
    long nout = ipow(N / 2, dim), nin = ipow(N, dim);
    memset(Wout, 0, sizeof(double) * nout * P);
    for (long ci = 0; ci < nin; ++ci) {
        int b[OR_MAXD], pb[OR_MAXD], sig[OR_MAXD];
        box_coords(dim, ci, N, b);
        for (int d = 0; d < dim; ++d) { pb[d] = b[d] / 2; sig[d] = b[d] % 2; }
        long po = box_id(dim, pb, N / 2);
        for (int ko = 0; ko < P; ++ko) {
            for (int ki = 0; ki < P; ++ki) {
                double v = 1.0;
                for (int d = dim - 1; d >= 0; --d)
                    v *= tab[(sig[d] * n + T->digit[ko * dim + d]) * n + T->digit[ki * dim + d]];
                double tmp = Win[ci * P + ki];
                tmp *= v;
                Wout[po * P + ko] += tmp;
            }
        }
    }
}

/* L2L from N/2 to N boxes per side (src/fmm_core.cpp:507-568).  The table is
 * S_n((c_ko+shift)/2, c_ki) = m2m_table[s][ki][ko] since S_n is symmetric;
 * it is rebuilt here with the reference's argument order. */
static void l2l(const otree_t *T, long N, const double *Win, double *Wout)
{
    int dim = T->dim, P = T->P, n = T->n;
    double tab[2 * 32 * 32];
    for (int s = 0; s < 2; ++s)
        for (int ko = 0; ko < n; ++ko)
            for (int ki = 0; ki < n; ++ki)
                tab[(s * n + ko) * n + ki] = oracle_cheb_S(n, (T->c[ko] + (2 * s - 1)) / 2, T->c[ki]);
    long nout = ipow(N, dim);
    for (long co = 0; co < nout; ++co) {
        int b[OR_MAXD], pb[OR_MAXD], sig[OR_MAXD];
        box_coords(dim, co, N, b);
        for (int d = 0; d < dim; ++d) { pb[d] = b[d] / 2; sig[d] = b[d] % 2; }
        long pi = box_id(dim, pb, N / 2);
        for (int ko = 0; ko < P; ++ko)
            for (int ki = 0; ki < P; ++ki) {
                double v = 1.0;
                for (int d = dim - 1; d >= 0; --d)
                    v *= tab[(sig[d] * n + T->digit[ko * dim + d]) * n + T->digit[ki * dim + d]];
                double tmp = Win[pi * P + ki];
                tmp *= v;
                Wout[co * P + ko] += tmp;
            }
    }
}

/* Interaction list of box b at a level with N boxes per side: children of
 * the parent's neighbours (clipped) that are not adjacent to b, in row-major
 * order (src/fmm_core.cpp:16-63).  Returns the count. */
int oracle_interaction_list(int dim, const int *b, long N, long *out)
{
    int lo[OR_MAXD], hi[OR_MAXD], shape[OR_MAXD];
    long size = 1;
    for (int d = 0; d < dim; ++d) {
        int p = b[d] / 2;
        lo[d] = (p > 0 ? p - 1 : p) * 2;
        hi[d] = (p + 1 < N / 2 ? p + 1 : p) * 2 + 1;
        shape[d] = hi[d] - lo[d] + 1;
        size *= shape[d];
    }
    int cnt = 0;
    for (long s = 0; s < size; ++s) {
        long ss = s;
        int pos[OR_MAXD], far = 0;
        for (int d = dim - 1; d >= 0; --d) {
            pos[d] = lo[d] + (int)(ss % shape[d]);
            ss /= shape[d];
            int dd = abs(pos[d] - b[d]);
            if (dd > far) far = dd;
        }
        if (far <= 1) continue;
        out[cnt++] = box_id(dim, pos, N);
    }
    return cnt;
}

/* Near-field list: the clipped 3^D block around b, row-major (src/fmm_core.cpp:81-115). */
int oracle_neighbour_list(int dim, const int *b, long N, long *out)
{
    int lo[OR_MAXD], shape[OR_MAXD];
    long size = 1;
    for (int d = 0; d < dim; ++d) {
        lo[d] = b[d] > 0 ? b[d] - 1 : b[d];
        int hi = b[d] + 1 < N ? b[d] + 1 : b[d];
        shape[d] = hi - lo[d] + 1;
        size *= shape[d];
    }
    for (long s = 0; s < size; ++s) {
        long ss = s;
        int pos[OR_MAXD];
        for (int d = dim - 1; d >= 0; --d) { pos[d] = lo[d] + (int)(ss % shape[d]); ss /= shape[d]; }
        out[s] = box_id(dim, pos, N);
    }
    return (int)size;
}

/* M2L at the level with N boxes per side (src/fmm_core.cpp:375-495). */
static void m2l(const otree_t *T, long N, int trans_sym, int kid, const double *prm,
                const double *Wm, double *Wl)
{
    int dim = T->dim, P = T->P;
    double len = T->Len / (double)N;
    long nbox = ipow(N, dim);
    if (trans_sym) {
        long nsh = ipow(7, dim);
        double *K = (double *)calloc((size_t)nsh * P * P, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
        for (long si = 0; si < nsh; ++si) {
            int sh[OR_MAXD], far = 0;
            long ss = si;
            for (int d = dim - 1; d >= 0; --d) {
                sh[d] = (int)(ss % 7) - 3; ss /= 7;
                if (abs(sh[d]) > far) far = abs(sh[d]);
            }
            if (far <= 1) continue;
            for (int kl = 0; kl < P; ++kl) {
                double X[OR_MAXD];
                for (int d = 0; d < dim; ++d) X[d] = len * (T->c[T->digit[kl * dim + d]] / 2);
                for (int km = 0; km < P; ++km) {
                    double Y[OR_MAXD];
                    for (int d = 0; d < dim; ++d) Y[d] = len * ((double)sh[d] + T->c[T->digit[km * dim + d]] / 2);
                    K[(si * P + kl) * P + km] = oracle_kernel(kid, prm, dim, X, Y);
                }
            }
        }
#pragma omp parallel
        {
            long lst[512];
#pragma omp for schedule(dynamic, 16)
            for (long bi = 0; bi < nbox; ++bi) {
                int b[OR_MAXD];
                box_coords(dim, bi, N, b);
                int cnt = oracle_interaction_list(dim, b, N, lst);
                for (int i = 0; i < cnt; ++i) {
                    int m[OR_MAXD];
                    box_coords(dim, lst[i], N, m);
                    long si = 0;
                    for (int d = 0; d < dim; ++d) si = si * 7 + (m[d] + 3 - b[d]);
                    const double *A = K + si * P * P, *x = Wm + lst[i] * P;
                    double *y = Wl + bi * P;
                    for (int kl = 0; kl < P; ++kl)
                        for (int km = 0; km < P; ++km) y[kl] += A[kl * P + km] * x[km];
                }
            }
        }
        free(K);
    } else {
#pragma omp parallel
        {
            long lst[512];
            double *A = (double *)malloc(sizeof(double) * P * P);
            double *Xl = (double *)malloc(sizeof(double) * P * dim);
            double *Ym = (double *)malloc(sizeof(double) * P * dim);
#pragma omp for schedule(dynamic, 4)
            for (long bi = 0; bi < nbox; ++bi) {
                int b[OR_MAXD];
                box_coords(dim, bi, N, b);
                for (int kl = 0; kl < P; ++kl)
                    for (int d = 0; d < dim; ++d)
                        Xl[kl * dim + d] = len * ((double)b[d] + (1.0 + T->c[T->digit[kl * dim + d]]) / 2) + T->origin[d];
                int cnt = oracle_interaction_list(dim, b, N, lst);
                for (int i = 0; i < cnt; ++i) {
                    int m[OR_MAXD];
                    box_coords(dim, lst[i], N, m);
                    for (int km = 0; km < P; ++km)
                        for (int d = 0; d < dim; ++d)
                            Ym[km * dim + d] = len * ((double)m[d] + (1.0 + T->c[T->digit[km * dim + d]]) / 2) + T->origin[d];
                    for (int kl = 0; kl < P; ++kl)
                        for (int km = 0; km < P; ++km)
                            A[kl * P + km] = oracle_kernel(kid, prm, dim, Xl + kl * dim, Ym + km * dim);
                    const double *x = Wm + lst[i] * P;
                    double *y = Wl + bi * P;
                    for (int kl = 0; kl < P; ++kl)
                        for (int km = 0; km < P; ++km) y[kl] += A[kl * P + km] * x[km];
                }
            }
            free(A); free(Xl); free(Ym);
        }
    }
}

/* --------------------------------------------------------------- solvers */

int oracle_exact(int dim, int kid, const double *prm, long nt, const double *t,
                 long ns, const double *w, const double *s, double *ans)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nt; ++i) {
        double a = 0.0;
        for (long j = 0; j < ns; ++j) a += w[j] * oracle_kernel(kid, prm, dim, t + i * dim, s + j * dim);
        ans[i] = a;
    }
    return 0;
}

/* grad[i][d] = sum_j w_j d/dx_d f(x_i, y_j), sources in ascending order. */
int oracle_exact_grad(int dim, int kid, const double *prm, long nt, const double *t,
                      long ns, const double *w, const double *s, double *grad)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nt; ++i) {
        double a[OR_MAXD] = {0.0, 0.0, 0.0}, g[OR_MAXD];
        for (long j = 0; j < ns; ++j) {
            oracle_kernel_grad(kid, prm, dim, t + i * dim, s + j * dim, g);
            for (int d = 0; d < dim; ++d) a[d] += w[j] * g[d];
        }
        for (int d = 0; d < dim; ++d) grad[i * dim + d] = a[d];
    }
    return 0;
}

int oracle_auto_depth(long nt, int dim)
{
    int depth = (int)(log((double)nt) / dim) + 1;
    return depth <= 0 ? 1 : depth;
}

/* The fmm solver (src/fmm.cpp:15-144).  `ans` must hold nt values; it is
 * overwritten by L2P when the leaf level is >= 2 and otherwise accumulated
 * into, as in the reference.  When Wm_out/Wl_out are given they receive the
 * multipole/local coefficients of levels 0..depth-1 packed one after another
 * (row-major box ids; levels never built stay zero). */
static int fmm_run(int dim, int poly_ord, int depth, int trans_sym, int kid, const double *prm,
                   long nt, const double *t, long ns, const double *w, const double *s,
                   double *ans, double *grad, double *Wm_out, double *Wl_out);

int oracle_fmm_levels(int dim, int poly_ord, int depth, int trans_sym, int kid, const double *prm,
                      long nt, const double *t, long ns, const double *w, const double *s,
                      double *ans, double *Wm_out, double *Wl_out)
{
    return fmm_run(dim, poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, ans, NULL, Wm_out, Wl_out);
}

/* The gradient of the fmm result with respect to the target coordinates:
 * the L2P interpolant differentiated (d r_d / d x_d = 2 / len_leaf) plus
 * sum_j w_j grad_x f(x_i, y_j) over the neighbour boxes.  grad is [nt][dim],
 * overwritten (zero for an empty sum). */
int oracle_fmm_grad(int dim, int poly_ord, int depth, int trans_sym, int kid, const double *prm,
                    long nt, const double *t, long ns, const double *w, const double *s, double *grad)
{
    for (long i = 0; i < nt * dim; ++i) grad[i] = 0.0;
    return fmm_run(dim, poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, NULL, grad, NULL, NULL);
}

static int fmm_run(int dim, int poly_ord, int depth, int trans_sym, int kid, const double *prm,
                   long nt, const double *t, long ns, const double *w, const double *s,
                   double *ans, double *grad, double *Wm_out, double *Wl_out)
{
    if (dim < 1 || dim > OR_MAXD || poly_ord < 0 || depth < 1) return -1;
    if (nt == 0 || ns == 0) return 0;
    otree_t T;
    tree_init(&T, dim, poly_ord, depth);
    oracle_bbox(dim, nt, t, ns, s, T.origin, &T.Len);
    int P = T.P, D = depth;

    double **Wm = (double **)calloc(D, sizeof(double *));
    double **Wl = (double **)calloc(D, sizeof(double *));
    for (int l = 0; l < D; ++l) {
        long nb = ipow(1L << l, dim);
        Wm[l] = (double *)calloc((size_t)nb * P, sizeof(double));
        Wl[l] = (double *)calloc((size_t)nb * P, sizeof(double));
    }
    long *lst_start, *lst;
    p2m(&T, ns, w, s, &lst_start, &lst, Wm[D - 1]);

    double tab[2 * 32 * 32];
    m2m_table(&T, tab);
    long N = T.NL;
    for (int step = 0; step + 3 < D; ++step) {
        m2m(&T, N, tab, Wm[D - step - 1], Wm[D - step - 2]);
        N /= 2;
    }
    for (int l = 2; l < D; ++l) m2l(&T, 1L << l, trans_sym, kid, prm, Wm[l], Wl[l]);
    for (int l = 2; l + 1 < D; ++l) l2l(&T, 2L << l, Wl[l], Wl[l + 1]);

    double len = T.Len / (double)T.NL;
    if (grad && D - 1 >= 2) {
        const double *W = Wl[D - 1];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nt; ++i) {
            int b[OR_MAXD];
            double rel[OR_MAXD];
            for (int d = 0; d < dim; ++d) b[d] = bin1(t[i * dim + d], T.origin[d], len, T.NL, &rel[d]);
            long bi = box_id(dim, b, T.NL);
            for (int e = 0; e < dim; ++e) {
                double a = 0.0;
                for (int k = 0; k < P; ++k) {
                    double v = 1.0;
                    for (int d = 0; d < dim; ++d) {
                        double c = T.c[T.digit[k * dim + d]];
                        v *= d == e ? oracle_cheb_dS(T.n, rel[d], c) : oracle_cheb_S(T.n, rel[d], c);
                    }
                    a += W[bi * P + k] * v;
                }
                grad[i * dim + e] = a * (2.0 / len);
            }
        }
    }
    if (grad) {
#pragma omp parallel
        {
            long nb[32];
#pragma omp for schedule(dynamic, 64)
            for (long i = 0; i < nt; ++i) {
                int b[OR_MAXD];
                for (int d = 0; d < dim; ++d) b[d] = bin1(t[i * dim + d], T.origin[d], len, T.NL, NULL);
                int cnt = oracle_neighbour_list(dim, b, T.NL, nb);
                double a[OR_MAXD], g[OR_MAXD];
                for (int d = 0; d < dim; ++d) a[d] = grad[i * dim + d];
                for (int q = 0; q < cnt; ++q)
                    for (long j = lst_start[nb[q]]; j < lst_start[nb[q] + 1]; ++j) {
                        long sj = lst[j];
                        oracle_kernel_grad(kid, prm, dim, t + i * dim, s + sj * dim, g);
                        for (int d = 0; d < dim; ++d) a[d] += w[sj] * g[d];
                    }
                for (int d = 0; d < dim; ++d) grad[i * dim + d] = a[d];
            }
        }
    }
    if (ans && D - 1 >= 2) {
        const double *W = Wl[D - 1];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nt; ++i) {
            int b[OR_MAXD];
            double rel[OR_MAXD];
            for (int d = 0; d < dim; ++d) b[d] = bin1(t[i * dim + d], T.origin[d], len, T.NL, &rel[d]);
            long bi = box_id(dim, b, T.NL);
            double a = 0.0;
            for (int k = 0; k < P; ++k) {
                double v = 1.0;
                for (int d = 0; d < dim; ++d) v *= oracle_cheb_S(T.n, rel[d], T.c[T.digit[k * dim + d]]);
                a += W[bi * P + k] * v;
            }
            ans[i] = 0.0;
            ans[i] += a;
        }
    }

#pragma omp parallel if (ans)
    {
        long nb[32];
#pragma omp for schedule(dynamic, 64)
        for (long i = 0; i < (ans ? nt : 0); ++i) {
            int b[OR_MAXD];
            for (int d = 0; d < dim; ++d) b[d] = bin1(t[i * dim + d], T.origin[d], len, T.NL, NULL);
            int cnt = oracle_neighbour_list(dim, b, T.NL, nb);
            double a = ans[i];
            for (int q = 0; q < cnt; ++q)
                for (long j = lst_start[nb[q]]; j < lst_start[nb[q] + 1]; ++j) {
                    long sj = lst[j];
                    a += w[sj] * oracle_kernel(kid, prm, dim, t + i * dim, s + sj * dim);
                }
            ans[i] = a;
        }
    }

    if (Wm_out || Wl_out) {
        size_t off = 0;
        for (int l = 0; l < D; ++l) {
            size_t cnt = (size_t)ipow(1L << l, dim) * P;
            if (Wm_out) memcpy(Wm_out + off, Wm[l], cnt * sizeof(double));
            if (Wl_out) memcpy(Wl_out + off, Wl[l], cnt * sizeof(double));
            off += cnt;
        }
    }
    for (int l = 0; l < D; ++l) { free(Wm[l]); free(Wl[l]); }
    free(Wm); free(Wl); free(lst_start); free(lst); free(T.digit);
    return 0;
}

/* M2P of one target from box s of a level with N boxes per side
 * (src/fmm_core.cpp:625-661): f(x, node position) * Wm[s][k] for every node,
 * node position = (c_k + 1)/2 * len + (len * s_d + origin_d). */
static double m2p_box(const otree_t *T, long N, int kid, const double *prm, const double *x, long sid,
                      const double *W)
{
    int dim = T->dim, sc[OR_MAXD];
    double len = T->Len / (double)N, rel[OR_MAXD], y[OR_MAXD], a = 0.0;
    box_coords(dim, sid, N, sc);
    for (int d = 0; d < dim; ++d) rel[d] = len * sc[d] + T->origin[d];
    for (int k = 0; k < T->P; ++k) {
        for (int d = 0; d < dim; ++d) y[d] = (T->c[T->digit[k * dim + d]] + 1.0) / 2.0 * len + rel[d];
        a += oracle_kernel(kid, prm, dim, x, y) * W[sid * T->P + k];
    }
    return a;
}

/* The tree solver (src/tree.cpp:12-90): P2M and M2M as in fmm, then M2P from
 * the interaction-list boxes of every level 2..Depth-1, then P2P over the
 * leaf neighbours.  ans is overwritten (zeroed first, src/tree.cpp:64-66). */
int oracle_tree(int dim, int poly_ord, int depth, int kid, const double *prm, long nt, const double *t,
                long ns, const double *w, const double *s, double *ans)
{
    if (dim < 1 || dim > OR_MAXD || poly_ord < 0 || depth < 1) return -1;
    if (nt == 0 || ns == 0) return 0;
    otree_t T;
    tree_init(&T, dim, poly_ord, depth);
    oracle_bbox(dim, nt, t, ns, s, T.origin, &T.Len);
    int D = depth;
    double **Wm = (double **)calloc(D, sizeof(double *));
    for (int l = 0; l < D; ++l) Wm[l] = (double *)calloc((size_t)ipow(1L << l, dim) * T.P, sizeof(double));
    long *lst_start, *lst;
    p2m(&T, ns, w, s, &lst_start, &lst, Wm[D - 1]);
    double tab[2 * 32 * 32];
    m2m_table(&T, tab);
    long N = T.NL;
    for (int step = 0; step + 3 < D; ++step) {
        m2m(&T, N, tab, Wm[D - step - 1], Wm[D - step - 2]);
        N /= 2;
    }
    double len = T.Len / (double)T.NL;
#pragma omp parallel
    {
        long lst_buf[1024], nb[32];
#pragma omp for schedule(dynamic, 64)
        for (long i = 0; i < nt; ++i) {
            const double *x = t + i * dim;
            double a = 0.0;
            long Nl = 1;
            for (int l = 0; l < D; ++l) {
                if (l >= 2) {
                    int b[OR_MAXD];
                    double ll = T.Len / (double)Nl;
                    for (int d = 0; d < dim; ++d) b[d] = bin1(x[d], T.origin[d], ll, (int)Nl, NULL);
                    int cnt = oracle_interaction_list(dim, b, Nl, lst_buf);
                    for (int q = 0; q < cnt; ++q) a += m2p_box(&T, Nl, kid, prm, x, lst_buf[q], Wm[l]);
                }
                Nl *= 2;
            }
            int b[OR_MAXD];
            for (int d = 0; d < dim; ++d) b[d] = bin1(x[d], T.origin[d], len, T.NL, NULL);
            int cnt = oracle_neighbour_list(dim, b, T.NL, nb);
            for (int q = 0; q < cnt; ++q)
                for (long j = lst_start[nb[q]]; j < lst_start[nb[q] + 1]; ++j) {
                    long sj = lst[j];
                    a += w[sj] * oracle_kernel(kid, prm, dim, x, s + sj * dim);
                }
            ans[i] = a;
        }
    }
    for (int l = 0; l < D; ++l) free(Wm[l]);
    free(Wm); free(lst_start); free(lst); free(T.digit);
    return 0;
}

/* nrn_N when it is not set: (int)(sqrt(3) * M^(1/(2 DIM))) + 1 (src/nrnmm.cpp:20-22). */
int oracle_nrn_auto(long ns, int dim)
{
    return (int)(sqrt(3.0) * pow((double)ns, 1.0 / (2.0 * dim))) + 1;
}

/* The nrnmm solver (src/nrnmm.cpp:13-100): one grid of n^DIM boxes; every
 * target takes its neighbour boxes (Chebyshev distance <= 1) directly and
 * every other box through its multipole expansion, boxes in row-major order.
 * ans is overwritten. */
int oracle_nrnmm(int dim, int poly_ord, int n, int kid, const double *prm, long nt, const double *t,
                 long ns, const double *w, const double *s, double *ans)
{
    if (dim < 1 || dim > OR_MAXD || poly_ord < 0) return -1;
    if (n <= 0) n = oracle_nrn_auto(ns, dim);
    if (nt == 0 || ns == 0) return 0;
    otree_t T;
    tree_init(&T, dim, poly_ord, 1);
    T.NL = n;
    oracle_bbox(dim, nt, t, ns, s, T.origin, &T.Len);
    long size = ipow(n, dim);
    double *Wm = (double *)calloc((size_t)size * T.P, sizeof(double));
    long *lst_start, *lst;
    p2m(&T, ns, w, s, &lst_start, &lst, Wm);
    double len = T.Len / (double)n;
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < nt; ++i) {
        const double *x = t + i * dim;
        int tb[OR_MAXD], sb[OR_MAXD];
        for (int d = 0; d < dim; ++d) tb[d] = bin1(x[d], T.origin[d], len, n, NULL);
        double a = 0.0;
        for (long sid = 0; sid < size; ++sid) {
            box_coords(dim, sid, n, sb);
            int far = 0;
            for (int d = 0; d < dim; ++d) {
                int dd = abs(sb[d] - tb[d]);
                if (dd > far) far = dd;
            }
            if (far <= 1) {
                for (long j = lst_start[sid]; j < lst_start[sid + 1]; ++j) {
                    long sj = lst[j];
                    a += w[sj] * oracle_kernel(kid, prm, dim, x, s + sj * dim);
                }
            } else {
                a += m2p_box(&T, n, kid, prm, x, sid, Wm);
            }
        }
        ans[i] = a;
    }
    free(Wm); free(lst_start); free(lst); free(T.digit);
    return 0;
}

int oracle_fmm(int dim, int poly_ord, int depth, int trans_sym, int kid, const double *prm,
               long nt, const double *t, long ns, const double *w, const double *s, double *ans)
{
    return oracle_fmm_levels(dim, poly_ord, depth, trans_sym, kid, prm, nt, t, ns, w, s, ans, NULL, NULL);
}

/* Stand-alone P2M at an arbitrary level and M2M chain, for the
 * M2M(P2M(child)) == P2M(parent) identity (test/src/fmma/m2m.cpp:17-82). */
int oracle_p2m_level(int dim, int poly_ord, int level, long nt, const double *t,
                     long ns, const double *w, const double *s, double *Wm)
{
    otree_t T;
    tree_init(&T, dim, poly_ord, level + 1);
    oracle_bbox(dim, nt, t, ns, s, T.origin, &T.Len);
    long *a, *b;
    p2m(&T, ns, w, s, &a, &b, Wm);
    free(a); free(b); free(T.digit);
    return 0;
}

int oracle_m2m_step(int dim, int poly_ord, int level_in, const double *Win, double *Wout)
{
    otree_t T;
    tree_init(&T, dim, poly_ord, level_in + 1);
    double tab[2 * 32 * 32];
    m2m_table(&T, tab);
    m2m(&T, 1L << level_in, tab, Win, Wout);
    free(T.digit);
    return 0;
}

int oracle_l2l_step(int dim, int poly_ord, int level_out, const double *Win, double *Wout)
{
    otree_t T;
    tree_init(&T, dim, poly_ord, level_out + 1);
    long nb = ipow(1L << level_out, dim);
    memset(Wout, 0, sizeof(double) * nb * T.P);
    l2l(&T, 1L << level_out, Win, Wout);
    free(T.digit);
    return 0;
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
