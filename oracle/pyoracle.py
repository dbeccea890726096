"""ctypes access to the CPU checkers (test infrastructure only).

* ``liboracle.so``          -- the C restatement in ``fmm_oracle.c``.
* ``_ref/libfmma_ref.so``   -- the compiled reference (fockl/FMMA ``src/*.cpp``)
  behind ``ref_shim.cpp``; present only where ``make -C oracle`` found the
  reference sources, or where such a build was copied along with this tree.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module.  The product path (``fmma_amd``) never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libfmma_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_lp = ctypes.POINTER(ctypes.c_long)
_ip = ctypes.POINTER(ctypes.c_int)

_lib = None
_ref = None


def _d(a):
    return a.ctypes.data_as(_dp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_ORACLE):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(_ORACLE)
        L.oracle_kernel.restype = ctypes.c_double
        L.oracle_kernel.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.oracle_cheb_T.restype = ctypes.c_double
        L.oracle_cheb_T.argtypes = [ctypes.c_int, ctypes.c_double]
        L.oracle_cheb_S.restype = ctypes.c_double
        L.oracle_cheb_S.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.oracle_nodes.argtypes = [ctypes.c_int, _dp]
        L.oracle_bbox.argtypes = [ctypes.c_int, ctypes.c_long, _dp, ctypes.c_long, _dp, _dp, _dp]
        L.oracle_leaf_boxes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, _dp, ctypes.c_long,
                                        _dp, ctypes.c_long, _dp, _lp, _dp]
        L.oracle_leaf_boxes_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, _fp, ctypes.c_long,
                                            _fp, ctypes.c_long, _fp, _lp]
        L.oracle_interaction_list.argtypes = [ctypes.c_int, _ip, ctypes.c_long, _lp]
        L.oracle_neighbour_list.argtypes = [ctypes.c_int, _ip, ctypes.c_long, _lp]
        L.oracle_exact.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long, _dp, ctypes.c_long,
                                   _dp, _dp, _dp]
        L.oracle_auto_depth.argtypes = [ctypes.c_long, ctypes.c_int]
        L.oracle_fmm_levels.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _dp, ctypes.c_long, _dp, ctypes.c_long, _dp, _dp, _dp, _dp, _dp]
        L.oracle_p2m_level.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long, _dp,
                                       ctypes.c_long, _dp, _dp, _dp]
        L.oracle_m2m_step.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_kernel_grad.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_tree.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long, _dp,
                                  ctypes.c_long, _dp, _dp, _dp]
        L.oracle_nrn_auto.argtypes = [ctypes.c_long, ctypes.c_int]
        L.oracle_nrnmm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long, _dp,
                                   ctypes.c_long, _dp, _dp, _dp]
        L.oracle_cheb_dS.restype = ctypes.c_double
        L.oracle_cheb_dS.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.oracle_exact_grad.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long, _dp, ctypes.c_long,
                                        _dp, _dp, _dp]
        L.oracle_fmm_grad.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      _dp, ctypes.c_long, _dp, ctypes.c_long, _dp, _dp, _dp]
        L.oracle_l2l_step.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


def have_ref():
    return os.path.exists(_REF)


def ref():
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError("oracle/_ref/libfmma_ref.so is not built")
        R = ctypes.CDLL(_REF)
        R.ref_solve_f64.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _dp, ctypes.c_long, _dp, ctypes.c_long, _dp, _dp, _dp]
        R.ref_solve_f32.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _dp, ctypes.c_long, _fp, ctypes.c_long, _fp, _fp, _fp]
        R.ref_levels_f64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     _dp, ctypes.c_long, _dp, ctypes.c_long, _dp, _dp, _dp, _dp]
        R.ref_leaf_boxes_f64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, _dp, ctypes.c_long, _dp,
                                         ctypes.c_long, _dp, _lp]
        R.ref_leaf_boxes_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, _fp, ctypes.c_long, _fp,
                                         ctypes.c_long, _fp, _lp]
        R.ref_set_threads.argtypes = [ctypes.c_int]
        R.ref_set_nrn_N.argtypes = [ctypes.c_int]
        _ref = R
    return _ref


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _prm(params):
    p = np.zeros(4)
    if params:
        p[: len(params)] = params
    return p


def auto_depth(nt, dim):
    return lib().oracle_auto_depth(int(nt), int(dim))


def kernel(kid, x, y, params=()):
    x, y, p = _c(x), _c(y), _prm(params)
    return lib().oracle_kernel(kid, _d(p), x.size, _d(x), _d(y))


def nodes(n):
    c = np.zeros(n)
    lib().oracle_nodes(n, _d(c))
    return c


def cheb_T(j, x):
    return lib().oracle_cheb_T(j, float(x))


def cheb_S(n, x, y):
    return lib().oracle_cheb_S(n, float(x), float(y))


def bbox(targets, sources):
    t, s = _c(targets), _c(sources)
    dim = t.shape[1]
    o, L = np.zeros(dim), np.zeros(1)
    lib().oracle_bbox(dim, t.shape[0], _d(t), s.shape[0], _d(s), _d(o), _d(L))
    return o, L[0]


def leaf_boxes(targets, sources, points, depth, rel=False):
    """Row-major leaf box id (and relative coordinates) of each point."""
    if np.asarray(points).dtype == np.float32:
        t, s, p = (_c(a, np.float32) for a in (targets, sources, points))
        box = np.zeros(p.shape[0], dtype=np.int64)
        fp = lambda a: a.ctypes.data_as(_fp)  # noqa: E731
        lib().oracle_leaf_boxes_f32(p.shape[1], depth, t.shape[0], fp(t), s.shape[0], fp(s),
                                    p.shape[0], fp(p), box.ctypes.data_as(_lp))
        return box
    t, s, p = _c(targets), _c(sources), _c(points)
    box = np.zeros(p.shape[0], dtype=np.int64)
    r = np.zeros_like(p)
    lib().oracle_leaf_boxes(p.shape[1], depth, t.shape[0], _d(t), s.shape[0], _d(s), p.shape[0], _d(p),
                            box.ctypes.data_as(_lp), _d(r))
    return (box, r) if rel else box


def interaction_list(b, N):
    b = np.ascontiguousarray(b, dtype=np.int32)
    out = np.zeros(1024, dtype=np.int64)
    n = lib().oracle_interaction_list(b.size, b.ctypes.data_as(_ip), N, out.ctypes.data_as(_lp))
    return out[:n]


def neighbour_list(b, N):
    b = np.ascontiguousarray(b, dtype=np.int32)
    out = np.zeros(64, dtype=np.int64)
    n = lib().oracle_neighbour_list(b.size, b.ctypes.data_as(_ip), N, out.ctypes.data_as(_lp))
    return out[:n]


def exact(targets, weights, sources, kid=0, params=()):
    t, w, s, p = _c(targets), _c(weights), _c(sources), _prm(params)
    ans = np.zeros(t.shape[0])
    lib().oracle_exact(t.shape[1], kid, _d(p), t.shape[0], _d(t), s.shape[0], _d(w), _d(s), _d(ans))
    return ans


def fmm(targets, weights, sources, poly_ord, depth=-1, trans_sym=0, kid=0, params=(), ans=None,
        levels=False):
    """The oracle fmm solver; ``ans`` (if given) is the in/out buffer whose old
    contents survive when the leaf level is < 2, as in the reference."""
    t, w, s, p = _c(targets), _c(weights), _c(sources), _prm(params)
    dim = t.shape[1]
    if depth <= 0:
        depth = auto_depth(t.shape[0], dim)
    out = np.zeros(t.shape[0]) if ans is None else _c(ans).copy()
    P = (poly_ord + 1) ** dim
    tot = sum((1 << (l * dim)) * P for l in range(depth))
    Wm = np.zeros(tot) if levels else None
    Wl = np.zeros(tot) if levels else None
    rc = lib().oracle_fmm_levels(dim, poly_ord, depth, trans_sym, kid, _d(p), t.shape[0], _d(t), s.shape[0],
                                 _d(w), _d(s), _d(out), _d(Wm) if levels else None, _d(Wl) if levels else None)
    if rc != 0:
        raise RuntimeError("oracle_fmm failed (%d)" % rc)
    if levels:
        return out, split_levels(Wm, dim, P, depth), split_levels(Wl, dim, P, depth)
    return out


def tree(targets, weights, sources, poly_ord, depth=-1, kid=0, params=()):
    """The oracle tree solver (M2P per level + P2P); ans overwritten."""
    t, w, s, p = _c(targets), _c(weights), _c(sources), _prm(params)
    dim = t.shape[1]
    if depth <= 0:
        depth = auto_depth(t.shape[0], dim)
    out = np.zeros(t.shape[0])
    if lib().oracle_tree(dim, poly_ord, depth, kid, _d(p), t.shape[0], _d(t), s.shape[0], _d(w), _d(s), _d(out)):
        raise RuntimeError("oracle_tree failed")
    return out


def nrn_auto(ns, dim):
    return lib().oracle_nrn_auto(ns, dim)


def nrnmm(targets, weights, sources, poly_ord, nrn_N=-1, kid=0, params=()):
    """The oracle nrnmm solver (one grid: neighbours direct, the rest by M2P)."""
    t, w, s, p = _c(targets), _c(weights), _c(sources), _prm(params)
    out = np.zeros(t.shape[0])
    if lib().oracle_nrnmm(t.shape[1], poly_ord, nrn_N, kid, _d(p), t.shape[0], _d(t), s.shape[0], _d(w), _d(s),
                          _d(out)):
        raise RuntimeError("oracle_nrnmm failed")
    return out


def kernel_grad(kid, params, x, y):
    """grad_x f(x, y) of one kernel at one pair of points."""
    x, y, p = _c(x), _c(y), _prm(params)
    g = np.zeros(x.shape[0])
    lib().oracle_kernel_grad(kid, _d(p), x.shape[0], _d(x), _d(y), _d(g))
    return g


def cheb_dS(n, x, y):
    return lib().oracle_cheb_dS(n, float(x), float(y))


def exact_grad(targets, weights, sources, kid=0, params=()):
    """grad[i, d] = sum_j w_j d/dx_d f(x_i, y_j)."""
    t, w, s, p = _c(targets), _c(weights), _c(sources), _prm(params)
    g = np.zeros(t.shape)
    lib().oracle_exact_grad(t.shape[1], kid, _d(p), t.shape[0], _d(t), s.shape[0], _d(w), _d(s), _d(g))
    return g


def fmm_grad(targets, weights, sources, poly_ord, depth=-1, trans_sym=0, kid=0, params=()):
    """Gradient of the oracle fmm result with respect to the target
    coordinates: [N, dim]."""
    t, w, s, p = _c(targets), _c(weights), _c(sources), _prm(params)
    dim = t.shape[1]
    if depth <= 0:
        depth = auto_depth(t.shape[0], dim)
    g = np.zeros(t.shape)
    rc = lib().oracle_fmm_grad(dim, poly_ord, depth, trans_sym, kid, _d(p), t.shape[0], _d(t), s.shape[0], _d(w),
                               _d(s), _d(g))
    if rc != 0:
        raise RuntimeError("oracle_fmm_grad failed (%d)" % rc)
    return g


SYMMETRY = {0: 1, 1: 1, 2: -1, 3: 1, 4: 1, 5: 1, 6: -1, 7: 1, 8: 1}


def fmm_transpose(targets, cot, sources, poly_ord, depth=-1, trans_sym=0, kid=0, params=()):
    """out[j] = sum_i cot_i f(x_i, y_j) by the oracle fmm with targets and
    sources swapped (f^T = symmetry * f for every kernel id here) at the Depth
    of the solve with (targets, sources)."""
    t = _c(targets)
    if depth <= 0:
        depth = auto_depth(t.shape[0], t.shape[1])
    s = _c(sources)
    out = fmm(s, cot, t, poly_ord, depth, trans_sym, kid, params, ans=np.zeros(s.shape[0]))
    return SYMMETRY[kid] * out


def split_levels(flat, dim, P, depth):
    out, off = [], 0
    for l in range(depth):
        nb = 1 << (l * dim)
        out.append(flat[off: off + nb * P].reshape(nb, P))
        off += nb * P
    return out


def p2m_level(targets, weights, sources, poly_ord, level):
    t, w, s = _c(targets), _c(weights), _c(sources)
    dim = t.shape[1]
    P = (poly_ord + 1) ** dim
    W = np.zeros((1 << (level * dim)) * P)
    lib().oracle_p2m_level(dim, poly_ord, level, t.shape[0], _d(t), s.shape[0], _d(w), _d(s), _d(W))
    return W.reshape(-1, P)


def m2m_step(Win, dim, poly_ord, level_in):
    Win = _c(Win)
    P = Win.shape[1]
    Wout = np.zeros(((1 << ((level_in - 1) * dim)), P))
    lib().oracle_m2m_step(dim, poly_ord, level_in, _d(Win), _d(Wout))
    return Wout


def l2l_step(Win, dim, poly_ord, level_out):
    Win = _c(Win)
    P = Win.shape[1]
    Wout = np.zeros(((1 << (level_out * dim)), P))
    lib().oracle_l2l_step(dim, poly_ord, level_out, _d(Win), _d(Wout))
    return Wout


# ----------------------------------------------------------- the reference

def ref_solve(solver, targets, weights, sources, poly_ord=1, depth=-1, trans_sym=0, kid=0, params=(),
              ans=None, threads=0, nrn_N=-1):
    """Run the compiled reference's FMMA<T,DIM>::solve (float32 or float64)."""
    R = ref()
    if threads:
        R.ref_set_threads(threads)
    R.ref_set_nrn_N(int(nrn_N))
    f32 = np.asarray(targets).dtype == np.float32
    dt = np.float32 if f32 else np.float64
    t, w, s, p = _c(targets, dt), _c(weights, dt), _c(sources, dt), _prm(params)
    out = np.zeros(t.shape[0], dtype=dt) if ans is None else _c(ans, dt).copy()
    fn = R.ref_solve_f32 if f32 else R.ref_solve_f64
    cast = (lambda a: a.ctypes.data_as(_fp)) if f32 else _d
    rc = fn(t.shape[1], solver.encode(), poly_ord, depth, trans_sym, kid, _d(p), t.shape[0], cast(t),
            s.shape[0], cast(w), cast(s), cast(out))
    if rc != 0:
        raise RuntimeError("reference solve failed")
    return out


def ref_levels(targets, weights, sources, poly_ord, depth, trans_sym=0, kid=0, params=()):
    t, w, s, p = _c(targets), _c(weights), _c(sources), _prm(params)
    dim = t.shape[1]
    P = (poly_ord + 1) ** dim
    tot = sum((1 << (l * dim)) * P for l in range(depth))
    Wm, Wl = np.zeros(tot), np.zeros(tot)
    ref().ref_levels_f64(dim, poly_ord, depth, trans_sym, kid, _d(p), t.shape[0], _d(t), s.shape[0], _d(w),
                         _d(s), _d(Wm), _d(Wl))
    return split_levels(Wm, dim, P, depth), split_levels(Wl, dim, P, depth)


def ref_leaf_boxes(targets, sources, points, depth):
    """The reference's own row-major leaf box of each point: P2M's binning
    (src/fmm_core.cpp:255-267) of `points` in the cube of targets and sources."""
    f32 = np.asarray(points).dtype == np.float32
    dt = np.float32 if f32 else np.float64
    t, s, p = _c(targets, dt), _c(sources, dt), _c(points, dt)
    box = np.zeros(p.shape[0], dtype=np.int64)
    fn = ref().ref_leaf_boxes_f32 if f32 else ref().ref_leaf_boxes_f64
    cast = (lambda a: a.ctypes.data_as(_fp)) if f32 else _d
    rc = fn(p.shape[1], int(depth), t.shape[0], cast(t), s.shape[0], cast(s), p.shape[0], cast(p),
            box.ctypes.data_as(_lp))
    if rc != 0 or np.any(box < 0):
        raise RuntimeError("reference leaf boxes failed")
    return box


def ref_set_threads(n):
    return ref().ref_set_threads(int(n))
