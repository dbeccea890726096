/*
 * fmm_oracle.h -- CPU restatement of the reference FMM (test checker only).
 * See fmm_oracle.c for the reference file:line each function follows.
 */
#ifndef FMM_ORACLE_H
#define FMM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Kernel ids; the same numbering as FMMA_KERNEL_* in include/fmma_hip.h. */
enum {
    ORACLE_K_INV_R = 0,     /* 1/|x-y|, inf at r=0 (reference default kernel) */
    ORACLE_K_INV_R_REG = 1, /* 1/|x-y|, 0 at r=0 */
    ORACLE_K_INV_DIFF = 2,  /* 1/(x_0-y_0) */
    ORACLE_K_LOG_R = 3,     /* log|x-y| = 0.5 log r^2, 0 at r=0 */
    ORACLE_K_GAUSS = 4,     /* exp(-r^2/(2h^2)), prm[0]=h */
    ORACLE_K_YUKAWA = 5,    /* exp(-kappa r)/r, 0 at r=0, prm[0]=kappa */
    ORACLE_K_SUM_R2 = 6,    /* sum_d (x_d-y_d) / r^2 */
    ORACLE_K_SUM2_R2 = 7,   /* sum_d (x_d+y_d)^2 / r^2 (not translation invariant) */
    ORACLE_K_INV_R2 = 8     /* 1/r^2, inf at r=0: the kernel of the reference's Python tutorials
                               (tutorials/tutorial0{1..5}/*.py), checked there as a Python callable */
};

double oracle_kernel(int id, const double *prm, int dim, const double *x, const double *y);
void oracle_kernel_grad(int id, const double *prm, int dim, const double *x, const double *y, double *g);
double oracle_cheb_T(int j, double x);
double oracle_cheb_dS(int n, double x, double y);
double oracle_cheb_S(int n, double x, double y);
void oracle_nodes(int n, double *c);
void oracle_bbox(int dim, long nt, const double *t, long ns, const double *s, double *origin, double *Len);
int oracle_leaf_boxes(int dim, int depth, long nt, const double *t, long ns, const double *s,
                      long np, const double *pts, long *box, double *rel);
int oracle_leaf_boxes_f32(int dim, int depth, long nt, const float *t, long ns, const float *s,
                          long np, const float *pts, long *box);
int oracle_interaction_list(int dim, const int *b, long N, long *out);
int oracle_neighbour_list(int dim, const int *b, long N, long *out);
int oracle_exact(int dim, int kid, const double *prm, long nt, const double *t,
                 long ns, const double *w, const double *s, double *ans);
int oracle_exact_grad(int dim, int kid, const double *prm, long nt, const double *t,
                      long ns, const double *w, const double *s, double *grad);
int oracle_auto_depth(long nt, int dim);
int oracle_tree(int dim, int poly_ord, int depth, int kid, const double *prm, long nt, const double *t,
                long ns, const double *w, const double *s, double *ans);
int oracle_nrn_auto(long ns, int dim);
int oracle_nrnmm(int dim, int poly_ord, int n, int kid, const double *prm, long nt, const double *t,
                 long ns, const double *w, const double *s, double *ans);
int oracle_fmm_grad(int dim, int poly_ord, int depth, int trans_sym, int kid, const double *prm,
                    long nt, const double *t, long ns, const double *w, const double *s, double *grad);
int oracle_fmm(int dim, int poly_ord, int depth, int trans_sym, int kid, const double *prm,
               long nt, const double *t, long ns, const double *w, const double *s, double *ans);
int oracle_fmm_levels(int dim, int poly_ord, int depth, int trans_sym, int kid, const double *prm,
                      long nt, const double *t, long ns, const double *w, const double *s,
                      double *ans, double *Wm_out, double *Wl_out);
int oracle_p2m_level(int dim, int poly_ord, int level, long nt, const double *t,
                     long ns, const double *w, const double *s, double *Wm);
int oracle_m2m_step(int dim, int poly_ord, int level_in, const double *Win, double *Wout);
int oracle_l2l_step(int dim, int poly_ord, int level_out, const double *Win, double *Wout);
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
