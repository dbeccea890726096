"""fmma_amd -- a black-box FMM for AMD MI355X (gfx950).

Computes c_i = sum_j w_j f(x_i, y_j) with the Chebyshev black-box FMM of
fockl/FMMA on the GPU.  ``Plan`` mirrors ``fmma::FMMA<TYPE,DIM>``
(include/fmma/fmma.hpp:52-170 of the reference) over the C-ABI of
``libfmma_hip.so``; ``fmma_amd.pyfmma`` is the drop-in for the reference's
``pyfmma`` module.  Arrays may be torch tensors on the GPU (device pointers,
solved in place on torch's current stream) or numpy arrays (staged through
device memory).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FMMAError, KERNELS, M2L_PRECISION  # noqa: F401

# one HIP runtime for torch and this package, whichever is imported first
# (fmma_amd.pyfmma links libfmma_hip.so directly)
_lib.pin_hip_runtime()

__all__ = ["Plan", "FMMAError", "KERNELS", "M2L_PRECISION", "version"]


def version():
    return _lib.lib().fmma_version().decode()


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _dptr(x):
    return x.data_ptr() if x is not None else None


def _hptr(x):
    return x.ctypes.data if x is not None else None


class Plan:
    """One FMM problem configuration: ``fmma::FMMA<TYPE, DIM>``.

    The defaults are the reference's (include/fmma/fmma.hpp:90-127): solver
    "exact", poly_ord 1, Depth -1 (auto, sticky), trans_sym_flag 0, kernel
    1/|x-y|.
    """

    def __init__(self, dim, dtype="float64"):
        self.dim = int(dim)
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float64, np.float32):
            raise ValueError("dtype must be float64 or float32")
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.fmma_plan_create(ctypes.byref(h), self.dim,
                                      _lib.FMMA_F64 if self.dtype == np.float64 else _lib.FMMA_F32))
        self._h = h
        self._L = L

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.fmma_plan_destroy(h)
            self._h = None

    # ---- setters (include/fmma/fmma.hpp:65-127)
    def set_solver_type(self, s):
        _lib.check(self._L.fmma_plan_set_solver_type(self._h, s.encode()))

    def set_poly_ord(self, p):
        _lib.check(self._L.fmma_plan_set_poly_ord(self._h, int(p)))

    def set_Depth(self, d):
        _lib.check(self._L.fmma_plan_set_depth(self._h, int(d)))

    set_depth = set_Depth

    @property
    def depth(self):
        d = ctypes.c_int()
        _lib.check(self._L.fmma_plan_get_depth(self._h, ctypes.byref(d)))
        return d.value

    def set_trans_sym_flag(self, f):
        _lib.check(self._L.fmma_plan_set_trans_sym_flag(self._h, int(f)))

    def set_nrn_N(self, n):
        _lib.check(self._L.fmma_plan_set_nrn_N(self._h, int(n)))

    @property
    def nrn_N(self):
        v = ctypes.c_int()
        _lib.check(self._L.fmma_plan_get_nrn_N(self._h, ctypes.byref(v)))
        return v.value

    def set_kernel(self, name, *params):
        kid = KERNELS[name] if isinstance(name, str) else int(name)
        p = (ctypes.c_double * max(1, len(params)))(*params)
        _lib.check(self._L.fmma_plan_set_kernel(self._h, kid, p, len(params)))

    def set_fn(self, fn, nargs=2):
        """A Python kernel: fn(x, y) (nargs 2, the reference's set_fn2) or
        fn(x - y) (nargs 1, set_fn1), written with arithmetic and numpy
        functions over the coordinates x[0], x[1], ...; traced and compiled
        for the GPU once (fmma_amd.jit), with its gradient."""
        from . import jit
        addr = jit.kernel_ops_address(fn, self.dim, self.dtype, nargs)
        _lib.check(self._L.fmma_plan_set_kernel_ops(self._h, ctypes.c_void_p(addr)))

    def set_fn1(self, fn):
        self.set_fn(fn, 1)

    def set_fn2(self, fn):
        self.set_fn(fn, 2)

    def set_m2l_precision(self, prec):
        v = M2L_PRECISION[prec] if isinstance(prec, str) else int(prec)
        _lib.check(self._L.fmma_plan_set_m2l_precision(self._h, v))

    def set_stream(self, stream_ptr):
        _lib.check(self._L.fmma_plan_set_stream(self._h, ctypes.c_void_p(int(stream_ptr or 0))))

    def set_timing(self, on):
        _lib.check(self._L.fmma_plan_set_timing(self._h, 1 if on else 0))

    def set_p2p_mutual(self, on):
        """Allow the mutual P2P form (default): with targets = sources and a
        symmetric kernel each pair of neighbour boxes is evaluated once."""
        _lib.check(self._L.fmma_plan_set_p2p_mutual(self._h, 1 if on else 0))

    def set_p2p_balance(self, on):
        """Balance the direct P2P over work items when some target boxes have
        a near field far above the average (default): clustered points."""
        _lib.check(self._L.fmma_plan_set_p2p_balance(self._h, 1 if on else 0))

    # ---- solve (src/fmma.cpp:143-207)
    def solve(self, targets, weights, sources, ans):
        """ans <- sum_j w_j f(x_i, y_j).  weights may be None (all ones).

        torch GPU tensors are used in place (device pointers, torch's current
        stream); numpy arrays go through the host entry point.  As in the
        reference, ans is overwritten by fmm when the leaf level is >= 2 and
        added to otherwise."""
        if _is_torch(targets):
            return self._solve_torch(targets, weights, sources, ans)
        return self._solve_numpy(targets, weights, sources, ans)

    def _check_shapes(self, t, w, s, a):
        if t.ndim != 2 or t.shape[1] != self.dim:
            raise ValueError("targets must have shape (N, %d)" % self.dim)
        if s.ndim != 2 or s.shape[1] != self.dim:
            raise ValueError("sources must have shape (M, %d)" % self.dim)
        if w is not None and (w.ndim != 1 or w.shape[0] != s.shape[0]):
            raise ValueError("weights must have shape (M,)")
        if a.ndim != 1 or a.shape[0] != t.shape[0]:
            raise ValueError("ans must have shape (N,)")

    def _solve_torch(self, t, w, s, a):
        import torch
        want = torch.float64 if self.dtype == np.float64 else torch.float32
        for x in (t, w, s, a):
            if x is None:
                continue
            if not x.is_cuda or x.dtype != want or not x.is_contiguous():
                raise ValueError("torch inputs must be contiguous %s tensors on the GPU" % want)
        self._check_shapes(t, w, s, a)
        self.set_stream(torch.cuda.current_stream(a.device).cuda_stream)
        _lib.check(self._L.fmma_plan_solve(self._h, t.shape[0], t.data_ptr(), w.data_ptr() if w is not None else None,
                                           s.shape[0], s.data_ptr(), a.data_ptr()))
        return a

    def _solve_numpy(self, t, w, s, a):
        for x in (t, w, s, a):
            if x is None:
                continue
            if not isinstance(x, np.ndarray) or x.dtype != self.dtype or not x.flags.c_contiguous:
                raise ValueError("numpy inputs must be C-contiguous %s arrays" % self.dtype)
        if not a.flags.writeable:
            raise ValueError("ans must be writeable")
        self._check_shapes(t, w, s, a)
        self.set_stream(0)
        ptr = lambda x: x.ctypes.data if x is not None else None  # noqa: E731
        _lib.check(self._L.fmma_plan_solve_host(self._h, t.shape[0], ptr(t), ptr(w), s.shape[0], ptr(s), ptr(a)))
        return a

    # ---- differentiation (new; the reference has no gradients)
    def gradient(self, targets, weights, sources, grad, reuse=True):
        """grad[i, d] <- d ans_i / d x_{i,d} for ans = solve(targets, weights,
        sources): the derivative of the fmm approximation (L2P interpolant
        differentiated, plus the near field of grad_x f), or of the exact sum
        for the exact solvers.  With reuse the local expansions of this plan's
        previous fmm solve are used when it had the same arrays and settings
        (the caller guarantees their contents are unchanged)."""
        self._check_vec(targets, weights, sources, grad, (targets.shape[0], self.dim), "grad")
        if _is_torch(targets):
            self._torch_ready(targets, weights, sources, grad)
            _lib.check(self._L.fmma_plan_gradient(self._h, targets.shape[0], targets.data_ptr(), _dptr(weights),
                                                  sources.shape[0], sources.data_ptr(), grad.data_ptr(),
                                                  1 if reuse else 0))
        else:
            self._numpy_ready(targets, weights, sources, grad)
            _lib.check(self._L.fmma_plan_gradient_host(self._h, targets.shape[0], _hptr(targets), _hptr(weights),
                                                       sources.shape[0], _hptr(sources), _hptr(grad)))
        return grad

    def solve_transpose(self, targets, cot, sources, out):
        """out[j] <- sum_i cot_i f(x_i, y_j): the adjoint in the weights of
        solve(targets, ., sources), i.e. d(sum_i cot_i ans_i)/d w_j.  The fmm
        solver runs with the roles swapped and f transposed on the same tree
        and Depth, so it is the exact transpose of that solve's operator."""
        if cot is None or cot.ndim != 1 or cot.shape[0] != targets.shape[0]:
            raise ValueError("cot must have shape (N,)")
        self._check_vec(targets, None, sources, out, (sources.shape[0],), "out")
        if _is_torch(targets):
            self._torch_ready(targets, cot, sources, out)
            _lib.check(self._L.fmma_plan_solve_transpose(self._h, targets.shape[0], targets.data_ptr(),
                                                         cot.data_ptr(), sources.shape[0], sources.data_ptr(),
                                                         out.data_ptr()))
        else:
            self._numpy_ready(targets, cot, sources, out)
            _lib.check(self._L.fmma_plan_solve_transpose_host(self._h, targets.shape[0], _hptr(targets), _hptr(cot),
                                                              sources.shape[0], _hptr(sources), _hptr(out)))
        return out

    def _check_vec(self, t, w, s, o, shape, name):
        if t.ndim != 2 or t.shape[1] != self.dim:
            raise ValueError("targets must have shape (N, %d)" % self.dim)
        if s.ndim != 2 or s.shape[1] != self.dim:
            raise ValueError("sources must have shape (M, %d)" % self.dim)
        if w is not None and (w.ndim != 1 or w.shape[0] != s.shape[0]):
            raise ValueError("weights must have shape (M,)")
        if tuple(o.shape) != tuple(shape):
            raise ValueError("%s must have shape %s" % (name, tuple(shape)))

    def _torch_ready(self, *xs):
        import torch
        want = torch.float64 if self.dtype == np.float64 else torch.float32
        for x in xs:
            if x is None:
                continue
            if not x.is_cuda or x.dtype != want or not x.is_contiguous():
                raise ValueError("torch inputs must be contiguous %s tensors on the GPU" % want)
        self.set_stream(torch.cuda.current_stream(xs[-1].device).cuda_stream)

    def _numpy_ready(self, *xs):
        for x in xs:
            if x is None:
                continue
            if not isinstance(x, np.ndarray) or x.dtype != self.dtype or not x.flags.c_contiguous:
                raise ValueError("numpy inputs must be C-contiguous %s arrays" % self.dtype)
        if not xs[-1].flags.writeable:
            raise ValueError("the output must be writeable")
        self.set_stream(0)

    # ---- batches (new; the reference solves one problem per call)
    def solve_batch(self, t_counts, targets, weights, s_counts, sources, ans):
        """Independent problems in one pass: problem p owns t_counts[p]
        consecutive rows of targets / ans and s_counts[p] consecutive rows of
        sources / weights (GPU tensors, concatenated).  Each is solved as
        solve() would solve it alone.  Passing the same tensor as targets and
        sources (equal counts) makes every problem's targets its sources."""
        tc = np.ascontiguousarray(t_counts, dtype=np.int64)
        sc = np.ascontiguousarray(s_counts, dtype=np.int64)
        if tc.ndim != 1 or sc.shape != tc.shape:
            raise ValueError("t_counts and s_counts must be 1-D of equal length")
        if targets.shape[0] != int(tc.sum()) or sources.shape[0] != int(sc.sum()):
            raise ValueError("the counts must add up to the number of targets and sources")
        self._check_vec(targets, weights, sources, ans, (targets.shape[0],), "ans")
        if not _is_torch(targets):
            raise ValueError("solve_batch takes GPU tensors")
        self._torch_ready(targets, weights, sources, ans)
        lp = ctypes.POINTER(ctypes.c_long)
        _lib.check(self._L.fmma_plan_solve_batch(self._h, int(tc.size), tc.ctypes.data_as(lp), targets.data_ptr(),
                                                 _dptr(weights), sc.ctypes.data_as(lp), sources.data_ptr(),
                                                 ans.data_ptr()))
        return ans

    def set_batch_graph(self, on):
        """Replay batched solves from one captured hipGraph (default on)."""
        _lib.check(self._L.fmma_plan_set_batch_graph(self._h, 1 if on else 0))

    def set_m2l_autotune(self, on):
        """Time the M2L tile widths on the first solve of a Depth and form, or
        the first batched solve of a layout (default on)."""
        _lib.check(self._L.fmma_plan_set_m2l_autotune(self._h, 1 if on else 0))

    def set_m2l_tile(self, tile):
        """Force the M2L tile width (target boxes per workgroup: 16, 32, 64,
        128, 256) wherever a pipelined kernel has it; 0 restores the tuner."""
        _lib.check(self._L.fmma_plan_set_m2l_tile(self._h, int(tile)))

    def m2l_tiles(self, batch=True):
        """{parent boxes per level: M2L tile width} chosen by the autotuner
        for batches (batch=True) or single solves."""
        b = 1 if batch else 0
        n = self._L.fmma_plan_m2l_tiles(self._h, b, None, None, 0)
        par = (ctypes.c_long * max(n, 1))()
        tb = (ctypes.c_int * max(n, 1))()
        self._L.fmma_plan_m2l_tiles(self._h, b, par, tb, n)
        return {int(par[i]): int(tb[i]) for i in range(n)}

    # ---- diagnostics
    def time_log(self):
        n = self._L.fmma_plan_time_log(self._h, None, None, 0)
        names = (ctypes.c_char_p * max(n, 1))()
        ms = (ctypes.c_double * max(n, 1))()
        self._L.fmma_plan_time_log(self._h, names, ms, n)
        return {names[i].decode(): ms[i] for i in range(n)}

    def clear_time_log(self):
        _lib.check(self._L.fmma_plan_clear_time_log(self._h))

    def point_boxes(self, which, n):
        """Row-major leaf box id of every source ('s') or target ('t') of the
        last fmm solve, in the caller's point order."""
        out = np.empty(int(n), dtype=np.int64)
        _lib.check(self._L.fmma_plan_point_boxes(self._h, 1 if which == "t" else 0, int(n),
                                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_long))))
        return out

    def level_coeffs(self, which, level, rowmajor=True):
        """Host copy of a level's multipoles (which='Wm') or locals ('Wl')
        from the last fmm solve, [boxes, P], in the reference's row-major box
        order by default."""
        ptr, nb, P = ctypes.c_void_p(), ctypes.c_long(), ctypes.c_int()
        _lib.check(self._L.fmma_plan_level_coeffs(self._h, 0 if which == "Wm" else 1, int(level), ctypes.byref(ptr),
                                                  ctypes.byref(nb), ctypes.byref(P)))
        out = np.empty((nb.value, P.value), dtype=self.dtype)
        _lib.hip().hipDeviceSynchronize()
        _lib.memcpy_d2h(out.ctypes.data, ptr.value, out.nbytes)
        if not rowmajor:
            return out
        rm = np.array([self._L.fmma_morton_to_rowmajor(self.dim, int(level), m) for m in range(nb.value)])
        res = np.empty_like(out)
        res[rm] = out
        return res
