// builtins.hpp -- launcher tables of the built-in kernels (builtins.hip).
#pragma once
#include "common.hpp"

namespace fmma {
namespace lib {

// Parameters of a built-in kernel as the device functors take them:
// p0 = h (gauss) or kappa (yukawa); p1 = 1/(2h^2) for gauss.
struct BuiltinParams {
    double p0 = 0.0, p1 = 0.0;
};

int builtin_kernel_ops(int id, int dim, int dtype, void *ctx, fmma_kernel_ops *out);

}  // namespace lib
}  // namespace fmma
