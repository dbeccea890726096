"""fmma_plan_solve_noalloc (the solve whose device buffers all come from the
library's static per-GPU scratch) against fmma_plan_solve, through the C-ABI:
two calls in a row on one plan, the second smaller (the scratch then holds the
first call's data), in each M2L form and solver; a solve beyond the scratch
fails with FMMA_ERR_SIZE and leaves ans alone."""
import ctypes

import numpy as np
import pytest

from fmma_amd import synth
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _plan(L, dim, solver, poly_ord, depth, prec, kernel=1, params=()):
    from fmma_amd import _lib
    h = ctypes.c_void_p()
    _lib.check(L.fmma_plan_create(ctypes.byref(h), dim, _lib.FMMA_F64))
    _lib.check(L.fmma_plan_set_solver_type(h, solver.encode()))
    _lib.check(L.fmma_plan_set_poly_ord(h, poly_ord))
    _lib.check(L.fmma_plan_set_depth(h, depth))
    _lib.check(L.fmma_plan_set_trans_sym_flag(h, 1))
    p = (ctypes.c_double * max(1, len(params)))(*params)
    _lib.check(L.fmma_plan_set_kernel(h, kernel, p, len(params)))
    _lib.check(L.fmma_plan_set_m2l_precision(h, prec))
    return h


def _call(fn, h, x, w, a):
    import torch
    rc = fn(h, x.shape[0], x.data_ptr(), w.data_ptr(), x.shape[0], x.data_ptr(), a.data_ptr())
    torch.cuda.synchronize()
    return rc


CASES = [
    # (id, dim, solver, poly_ord, depth, M2L form, kernel id, params, sizes of the two calls)
    ("fmm_3d_fp64", 3, "fmm", 3, -1, 0, 1, (), (100000, 30000)),
    ("fmm_3d_fp32", 3, "fmm", 3, 4, 1, 1, (), (60000, 20000)),
    ("fmm_3d_bf16x3_p4", 3, "fmm", 4, 4, 2, 5, (2.0,), (40000, 9000)),
    ("fmm_2d_log", 2, "fmm", 5, -1, 0, 3, (), (80000, 20000)),
    ("exact_3d", 3, "exact", 1, -1, 0, 1, (), (6000, 300)),
    ("tree_3d", 3, "tree", 3, 4, 0, 1, (), (30000, 8000)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_solve(gpu, case):
    import torch
    from fmma_amd import _lib
    name, dim, solver, p, depth, prec, kid, params, sizes = case
    L = _lib.lib()
    h_ref = _plan(L, dim, solver, p, depth, prec, kid, params)
    h_na = _plan(L, dim, solver, p, depth, prec, kid, params)
    try:
        for k, n in enumerate(sizes):
            print("case %s call %d n=%d" % (name, k, n), flush=True)
            x = torch.from_numpy(synth.points(300 + k, n, dim)).cuda()
            w = torch.from_numpy(synth.normal(310 + k, n)).cuda()
            a_ref = torch.zeros(n, dtype=torch.float64, device="cuda")
            a_na = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
            _lib.check(_call(L.fmma_plan_solve, h_ref, x, w, a_ref))
            _lib.check(_call(L.fmma_plan_solve_noalloc, h_na, x, w, a_na))
            ref, got = a_ref.cpu().numpy(), a_na.cpu().numpy()
            assert np.all(np.isfinite(got))
            err = rel_l2(got, ref)
            print("  rel_l2 noalloc vs solve %.3e" % err, flush=True)
            assert err <= 1e-12, err
            d_ref, d_na = ctypes.c_int(), ctypes.c_int()
            L.fmma_plan_get_depth(h_ref, ctypes.byref(d_ref))
            L.fmma_plan_get_depth(h_na, ctypes.byref(d_na))
            assert d_ref.value == d_na.value  # the auto Depth is kept the same way
    finally:
        L.fmma_plan_destroy(h_ref)
        L.fmma_plan_destroy(h_na)


def test_noalloc_rejects_beyond_scratch(gpu):
    import torch
    from fmma_amd import _lib
    L = _lib.lib()
    h = _plan(L, 3, "fmm", 3, -1, 0, 1)
    try:
        n = 600000
        x = torch.from_numpy(synth.points(320, n, 3)).cuda()
        w = torch.from_numpy(synth.uniform(321, n)).cuda()
        a = torch.full((n,), 5.0, dtype=torch.float64, device="cuda")
        rc = _call(L.fmma_plan_solve_noalloc, h, x, w, a)
        assert rc == _lib.FMMA_ERR_SIZE
        assert b"no-alloc scratch" in L.fmma_last_error()
        assert torch.all(a == 5.0).item()
        # a size that fits still works on the same plan afterwards
        m = 20000
        y = torch.from_numpy(synth.points(322, m, 3)).cuda()
        v = torch.from_numpy(synth.uniform(323, m)).cuda()
        b = torch.zeros(m, dtype=torch.float64, device="cuda")
        ref = torch.zeros(m, dtype=torch.float64, device="cuda")
        h2 = _plan(L, 3, "fmm", 3, 4, 0, 1)
        L.fmma_plan_set_depth(h, 4)
        _lib.check(_call(L.fmma_plan_solve_noalloc, h, y, v, b))
        _lib.check(_call(L.fmma_plan_solve, h2, y, v, ref))
        L.fmma_plan_destroy(h2)
        assert rel_l2(b.cpu().numpy(), ref.cpu().numpy()) <= 1e-12
    finally:
        L.fmma_plan_destroy(h)
