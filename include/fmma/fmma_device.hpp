// fmma/fmma_device.hpp -- FMMA::set_fn for __host__ __device__ functors.
//
// Included by fmma/fmma.hpp when compiling with hipcc.  The functor is
// wrapped so that it is called exactly like the reference calls f
// (include/fmma/fmma.hpp:55-87): with std::array<TYPE, DIM> arguments, as
// f(x, y) or, for a one-argument f, as f(x - y).  The kernel-dependent
// stages (P2P, exact, M2L operators; the gradient's near field when a grad_f
// is given; the transposed kernel's stages when f declares no symmetry) are
// instantiated for it here, in the caller's translation unit, and handed to
// the plan as fmma_kernel_ops.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <type_traits>

#include "detail/pair_kernels.hpp"
#include "fmma.hpp"

namespace fmma {
namespace detail {

struct NoGrad {};

template <typename TYPE, std::size_t DIM, class F, bool TWO_ARG, class G = NoGrad>
struct UserKernel {
    F f;
    G g;  // grad_x f, or NoGrad
    // f(x - y) is translation invariant by construction; f(x, y) is taken as
    // general unless the caller sets trans_sym_flag (src/fmm_core.cpp:383).
    static constexpr bool translation_invariant = !TWO_ARG;
    // a functor type may declare `static constexpr int symmetry = 1` (f(y,x) =
    // f(x,y)) or -1 (f(y,x) = -f(x,y)) to let P2P evaluate each pair once
    static constexpr int symmetry = kernel_symmetry<F>::value;

    template <typename T, int D>
    __device__ __forceinline__ T operator()(const T *x, const T *y) const
    {
        std::array<TYPE, DIM> a, b;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            a[d] = x[d];
            b[d] = y[d];
        }
        if constexpr (TWO_ARG) {
            return (T)f(a, b);
        } else {
            std::array<TYPE, DIM> diff;
#pragma unroll
            for (int d = 0; d < D; ++d) diff[d] = a[d] - b[d];
            return (T)f(diff);
        }
    }

    template <typename T, int D, class GG = G, class = std::enable_if_t<!std::is_same<GG, NoGrad>::value>>
    __device__ __forceinline__ void grad(const T *x, const T *y, T *out) const
    {
        std::array<TYPE, DIM> a, b, r;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            a[d] = x[d];
            b[d] = y[d];
        }
        if constexpr (TWO_ARG) {
            r = g(a, b);
        } else {
            std::array<TYPE, DIM> diff;
#pragma unroll
            for (int d = 0; d < D; ++d) diff[d] = a[d] - b[d];
            r = g(diff);
        }
#pragma unroll
        for (int d = 0; d < D; ++d) out[d] = (T)r[d];
    }
};

// The functor, its transpose and the transpose's launcher table, kept alive
// together with the plan that points at them.
template <class K>
struct UserHolder {
    K k;
    Transposed<K> kt;
    fmma_kernel_ops ops_t;
};

template <typename TYPE, std::size_t DIM, class K>
std::shared_ptr<UserHolder<K>> make_user_ops(const K &k, fmma_kernel_ops &ops)
{
    static_assert(std::is_trivially_copyable<K>::value, "the functor is copied to the GPU by value");
    auto h = std::make_shared<UserHolder<K>>(UserHolder<K>{k, Transposed<K>{k}, fmma_kernel_ops{}});
    ops = ops_for<TYPE, (int)DIM, K>::make(&h->k);
    if constexpr (kernel_symmetry<K>::value == 0) {
        h->ops_t = ops_for<TYPE, (int)DIM, Transposed<K>>::make(&h->kt);
        ops.transposed = &h->ops_t;
    }
    return h;
}

}  // namespace detail

template <typename TYPE, std::size_t DIM>
template <class F>
void FMMA<TYPE, DIM>::set_fn(F f)
{
    using A = std::array<TYPE, DIM>;
    constexpr bool two = std::is_invocable<F, const A &, const A &>::value;
    constexpr bool one = std::is_invocable<F, const A &>::value;
    static_assert(two || one, "set_fn needs f(x, y) or f(x - y) taking std::array<TYPE, DIM>");
    using K = detail::UserKernel<TYPE, DIM, F, two>;
    fmma_kernel_ops ops;
    auto holder = detail::make_user_ops<TYPE, DIM>(K{f, {}}, ops);
    check(fmma_plan_set_kernel_ops(plan_, &ops));
    user_ = holder;
}

template <typename TYPE, std::size_t DIM>
template <class F, class G>
void FMMA<TYPE, DIM>::set_fn(F f, G grad_f)
{
    using A = std::array<TYPE, DIM>;
    constexpr bool two = std::is_invocable<F, const A &, const A &>::value;
    constexpr bool one = std::is_invocable<F, const A &>::value;
    static_assert(two || one, "set_fn needs f(x, y) or f(x - y) taking std::array<TYPE, DIM>");
    if constexpr (two)
        static_assert(std::is_convertible<std::invoke_result_t<G, const A &, const A &>, A>::value,
                      "grad_f(x, y) must return std::array<TYPE, DIM>");
    else
        static_assert(std::is_convertible<std::invoke_result_t<G, const A &>, A>::value,
                      "grad_f(x - y) must return std::array<TYPE, DIM>");
    using K = detail::UserKernel<TYPE, DIM, F, two, G>;
    fmma_kernel_ops ops;
    auto holder = detail::make_user_ops<TYPE, DIM>(K{f, grad_f}, ops);
    check(fmma_plan_set_kernel_ops(plan_, &ops));
    user_ = holder;
}

}  // namespace fmma
