"""torch.autograd over the FMM: u = FMM(x; w, y), u_i = sum_j w_j f(x_i, y_j).

The backward is written by hand on the library's own entry points:

* dL/dw_j = sum_i g_i f(x_i, y_j) (g = dL/du): Plan.solve_transpose, the FMM
  with targets and sources swapped and f transposed on the same tree and
  Depth, which is the exact adjoint of the forward FMM operator;
* dL/dx_i = g_i grad_x u(x_i): Plan.gradient, the L2P interpolant
  differentiated through the Chebyshev basis plus the near field of grad_x f.

The reference (fockl/FMMA) has no gradients; the forward is its fmm solver
(src/fmm.cpp:15-144).  Gradients with respect to the source coordinates are
not provided.

``recompute=False`` lets the backward reuse the local expansions that the
forward left in the plan (valid while nothing else ran on the plan with other
arrays; otherwise the library recomputes them by itself).  ``recompute=True``
always reruns the upward and downward passes in the backward, so several
calls can share one plan and no per-call state is kept.
"""
import torch

from . import Plan

__all__ = ["FMMFunction", "fmm", "make_plan"]


class FMMFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, targets, weights, sources, plan, recompute):
        ans = torch.zeros(targets.shape[0], dtype=targets.dtype, device=targets.device)
        plan.solve(targets, weights, sources, ans)
        ctx.save_for_backward(targets, weights, sources)
        ctx.plan = plan
        ctx.recompute = bool(recompute)
        return ans

    @staticmethod
    def backward(ctx, g):
        targets, weights, sources = ctx.saved_tensors
        plan = ctx.plan
        if ctx.needs_input_grad[2]:
            raise NotImplementedError("gradients with respect to the source coordinates are not provided")
        g = g.contiguous()
        gx = gw = None
        # the gradient first: it may reuse the forward's expansions, which the
        # transpose solve overwrites
        if ctx.needs_input_grad[0]:
            grad = torch.empty_like(targets)
            plan.gradient(targets, weights, sources, grad, reuse=not ctx.recompute)
            gx = g[:, None] * grad
        if ctx.needs_input_grad[1]:
            gw = torch.empty_like(weights)
            plan.solve_transpose(targets, g, sources, gw)
        return gx, gw, None, None, None


def make_plan(dim, dtype=torch.float64, kernel="inv_r_reg", params=(), poly_ord=3, depth=-1, trans_sym=1,
              m2l_precision="native", solver="fmm"):
    """A Plan configured for FMMFunction."""
    pl = Plan(dim, "float64" if dtype == torch.float64 else "float32")
    pl.set_solver_type(solver)
    pl.set_poly_ord(poly_ord)
    pl.set_Depth(depth)
    pl.set_trans_sym_flag(trans_sym)
    pl.set_kernel(kernel, *params)
    pl.set_m2l_precision(m2l_precision)
    return pl


def fmm(targets, weights, sources, plan, recompute=False):
    """u = FMM(targets; weights, sources) with a hand-written backward for
    targets and weights.  targets [N, D], weights [M], sources [M, D]:
    contiguous GPU tensors of the plan's dtype."""
    return FMMFunction.apply(targets, weights, sources, plan, recompute)
