// builtins.hip -- launcher tables (fmma_kernel_ops) for the FMMA_KERNEL_* ids.
#include "../../include/fmma/detail/pair_kernels.hpp"
#include "builtin_kernels.hpp"
#include "builtins.hpp"

namespace fmma {
namespace lib {

template <typename T, int D, class K>
struct builtin_ops {
    static K make(void *ctx)
    {
        const BuiltinParams *c = static_cast<const BuiltinParams *>(ctx);
        K k;
        k.p0 = c->p0;
        k.p1 = c->p1;
        return k;
    }
    static int p2p(const fmma_p2p_args *a, void *ctx, void *s)
    {
        return detail::launch_p2p<T, D, K>(a, make(ctx), (hipStream_t)s);
    }
    static int exact(const fmma_exact_args *a, void *ctx, void *s)
    {
        return detail::launch_exact<T, D, K>(a, make(ctx), (hipStream_t)s);
    }
    static int m2l_op(const fmma_m2l_op_args *a, void *ctx, void *s)
    {
        return detail::launch_m2l_op<T, D, K>(a, make(ctx), (hipStream_t)s);
    }
    static int m2l_general(const fmma_m2l_general_args *a, void *ctx, void *s)
    {
        return detail::launch_m2l_general<T, D, K>(a, make(ctx), (hipStream_t)s);
    }
    static int p2p_grad(const fmma_p2p_grad_args *a, void *ctx, void *s)
    {
        return detail::launch_p2p_grad<T, D, K>(a, make(ctx), (hipStream_t)s);
    }
    static int exact_grad(const fmma_exact_grad_args *a, void *ctx, void *s)
    {
        return detail::launch_exact_grad<T, D, K>(a, make(ctx), (hipStream_t)s);
    }
    static int m2p(const fmma_m2p_args *a, void *ctx, void *s)
    {
        return detail::launch_m2p<T, D, K>(a, make(ctx), (hipStream_t)s);
    }
    static void fill(fmma_kernel_ops *o, void *ctx)
    {
        o->dim = D;
        o->dtype = sizeof(T) == 8 ? FMMA_F64 : FMMA_F32;
        o->translation_invariant = K::translation_invariant ? 1 : 0;
        o->symmetry = detail::kernel_symmetry<K>::value;
        o->ctx = ctx;
        o->p2p = &p2p;
        o->exact = &exact;
        o->m2l_op = &m2l_op;
        o->m2l_general = &m2l_general;
        o->p2p_grad = &p2p_grad;
        o->exact_grad = &exact_grad;
        o->transposed = nullptr;  // every built-in is symmetric or antisymmetric
        o->m2p = &m2p;
    }
};

// Compiled once per (TYPE, DIM) with FMMA_BI_TYPE and FMMA_BI_DIM defined, so
// that the eight kernels' instantiations build in parallel; compiled without
// them it is the dispatcher over those six objects.
template <typename T, int D>
int fill_builtin(int id, void *ctx, fmma_kernel_ops *o)
{
    switch (id) {
    case FMMA_KERNEL_INV_R: builtin_ops<T, D, KInvR>::fill(o, ctx); return FMMA_OK;
    case FMMA_KERNEL_INV_R_REG: builtin_ops<T, D, KInvRReg>::fill(o, ctx); return FMMA_OK;
    case FMMA_KERNEL_INV_DIFF: builtin_ops<T, D, KInvDiff>::fill(o, ctx); return FMMA_OK;
    case FMMA_KERNEL_LOG_R: builtin_ops<T, D, KLogR>::fill(o, ctx); return FMMA_OK;
    case FMMA_KERNEL_GAUSS: builtin_ops<T, D, KGauss>::fill(o, ctx); return FMMA_OK;
    case FMMA_KERNEL_YUKAWA: builtin_ops<T, D, KYukawa>::fill(o, ctx); return FMMA_OK;
    case FMMA_KERNEL_SUM_R2: builtin_ops<T, D, KSumR2>::fill(o, ctx); return FMMA_OK;
    case FMMA_KERNEL_SUM2_R2: builtin_ops<T, D, KSum2R2>::fill(o, ctx); return FMMA_OK;
    }
    return FMMA_ERR_INVALID;
}

#if defined(FMMA_BI_TYPE) && defined(FMMA_BI_DIM)
template int fill_builtin<FMMA_BI_TYPE, FMMA_BI_DIM>(int, void *, fmma_kernel_ops *);
#else
extern template int fill_builtin<double, 1>(int, void *, fmma_kernel_ops *);
extern template int fill_builtin<double, 2>(int, void *, fmma_kernel_ops *);
extern template int fill_builtin<double, 3>(int, void *, fmma_kernel_ops *);
extern template int fill_builtin<float, 1>(int, void *, fmma_kernel_ops *);
extern template int fill_builtin<float, 2>(int, void *, fmma_kernel_ops *);
extern template int fill_builtin<float, 3>(int, void *, fmma_kernel_ops *);

int builtin_kernel_ops(int id, int dim, int dtype, void *ctx, fmma_kernel_ops *out)
{
    if (dtype == FMMA_F64) {
        if (dim == 1) return fill_builtin<double, 1>(id, ctx, out);
        if (dim == 2) return fill_builtin<double, 2>(id, ctx, out);
        if (dim == 3) return fill_builtin<double, 3>(id, ctx, out);
    } else {
        if (dim == 1) return fill_builtin<float, 1>(id, ctx, out);
        if (dim == 2) return fill_builtin<float, 2>(id, ctx, out);
        if (dim == 3) return fill_builtin<float, 3>(id, ctx, out);
    }
    return FMMA_ERR_INVALID;
}
#endif

}  // namespace lib
}  // namespace fmma
