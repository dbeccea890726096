// fmma/fmma_device.hpp -- FMMA::set_fn for __host__ __device__ functors.
//
// Included by fmma/fmma.hpp when compiling with hipcc.  The functor is
// wrapped so that it is called exactly like the reference calls f
// (include/fmma/fmma.hpp:55-87): with std::array<TYPE, DIM> arguments, as
// f(x, y) or, for a one-argument f, as f(x - y).  The kernel-dependent
// stages (P2P, exact, M2L operators) are instantiated for it here, in the
// caller's translation unit, and handed to the plan as fmma_kernel_ops.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <type_traits>

#include "detail/pair_kernels.hpp"
#include "fmma.hpp"

namespace fmma {
namespace detail {

template <typename TYPE, std::size_t DIM, class F, bool TWO_ARG>
struct UserKernel {
    F f;
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
};

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
    static_assert(std::is_trivially_copyable<K>::value, "the functor is copied to the GPU by value");
    auto holder = std::make_shared<K>(K{f});
    fmma_kernel_ops ops = detail::ops_for<TYPE, (int)DIM, K>::make(holder.get());
    check(fmma_plan_set_kernel_ops(plan_, &ops));
    user_ = holder;
}

}  // namespace fmma
