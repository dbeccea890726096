"""GPU parity of the tree and nrnmm solvers against the oracle.

tree (src/tree.cpp:12-90): P2M + M2M, then M2P from the interaction lists of
levels 2..Depth-1 straight to the targets, then P2P.  nrnmm
(src/nrnmm.cpp:13-100): one grid of nrn_N^DIM boxes, M2P from every box at
Chebyshev distance >= 2, P2P from the rest.  Bars as for fmm: FP64 within
1e-12 relative (L2 over all targets) of the oracle at the same order and
Depth / nrn_N; FP32 within 3e-5 of the FP64 oracle.
"""
import numpy as np
import pytest

from fmma_amd import synth
from tests.conftest import rel_l2
from tests.test_gpu_parity import KID

pytestmark = pytest.mark.gpu


def _solve(gpu, solver, t, w, s, p, size=-1, kernel="inv_r", params=(), dtype=np.float64, shared=False,
           ans0=None, plan=None):
    import torch
    pl = plan or gpu.Plan(t.shape[1], dtype)
    pl.set_solver_type(solver)
    pl.set_poly_ord(p)
    if plan is None:
        if solver == "tree":
            pl.set_Depth(size)
        else:
            pl.set_nrn_N(size)
    pl.set_kernel(kernel, *params)
    tt = torch.from_numpy(np.ascontiguousarray(t, dtype=dtype)).cuda()
    st = tt if shared else torch.from_numpy(np.ascontiguousarray(s, dtype=dtype)).cuda()
    wt = None if w is None else torch.from_numpy(np.ascontiguousarray(w, dtype=dtype)).cuda()
    a = torch.zeros(t.shape[0], dtype=tt.dtype, device="cuda") if ans0 is None else \
        torch.from_numpy(np.ascontiguousarray(ans0, dtype=dtype)).cuda()
    pl.solve(tt, wt, st, a)
    torch.cuda.synchronize()
    return a.cpu().numpy().astype(np.float64), pl


def _oracle(oracle, solver, t, w, s, p, size, kernel, params):
    w = np.ones(s.shape[0]) if w is None else w
    if solver == "tree":
        return oracle.tree(t, w, s, p, size, KID[kernel], params)
    return oracle.nrnmm(t, w, s, p, size, KID[kernel], params)


CASES = [
    # solver, dim, nt, ns, shared, p, Depth or nrn_N (-1 auto), kernel, params
    ("tree", 1, 4096, 4096, False, 7, -1, "inv_diff", ()),
    ("tree", 2, 12000, 12000, True, 5, -1, "log_r", ()),
    ("tree", 2, 3000, 2500, False, 4, 5, "sum2_r2", ()),
    ("tree", 3, 20000, 20000, True, 3, -1, "inv_r_reg", ()),
    ("tree", 3, 9000, 7000, False, 4, 5, "yukawa", (2.0,)),
    ("tree", 3, 6000, 6000, False, 2, 4, "gauss", (0.1,)),
    ("tree", 3, 1001, 777, False, 3, 3, "inv_r", ()),
    ("tree", 3, 3000, 3000, True, 6, 3, "inv_r_reg", ()),
    ("nrnmm", 1, 4096, 4096, False, 7, -1, "inv_diff", ()),
    ("nrnmm", 2, 8000, 8000, True, 5, -1, "log_r", ()),
    ("nrnmm", 2, 3000, 2500, False, 4, 7, "sum2_r2", ()),
    ("nrnmm", 3, 12000, 12000, True, 3, -1, "inv_r_reg", ()),
    ("nrnmm", 3, 6000, 5000, False, 4, 5, "yukawa", (2.0,)),
    ("nrnmm", 3, 1001, 777, False, 2, 6, "gauss", (0.1,)),
    ("nrnmm", 3, 3000, 3000, False, 6, 3, "inv_r", ()),  # inv_r is inf at r = 0: distinct point sets
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s_%dd_%s_n%d_p%d_s%d%s" % (
    c[0], c[1], c[7], c[2], c[5], c[6], "_shared" if c[4] else ""))
def test_matches_oracle(gpu, oracle, case):
    solver, dim, nt, ns, shared, p, size, kernel, params = case
    t = synth.points(301, nt, dim)
    s = t if shared else synth.points(302, ns, dim)
    w = synth.normal(303, s.shape[0])
    print("case", case, flush=True)
    got, pl = _solve(gpu, solver, t, w, s, p, size, kernel, params, shared=shared)
    used = pl.depth if solver == "tree" else pl.nrn_N
    if size <= 0:
        auto = oracle.auto_depth(nt, dim) if solver == "tree" else oracle.nrn_auto(ns, dim)
        assert used == auto
    else:
        assert used == size
    ref = _oracle(oracle, solver, t, w, s, p, used, kernel, params)
    e = rel_l2(got, ref)
    print("  %s size %d: rel %.3e" % (solver, used, e))
    assert e <= 1e-12


@pytest.mark.parametrize("solver", ["tree", "nrnmm"])
def test_float32(gpu, oracle, solver):
    t = synth.points(311, 10000, 3)
    w = synth.uniform(312, 10000)
    got, pl = _solve(gpu, solver, t, w, t, 3, -1, "inv_r_reg", dtype=np.float32, shared=True)
    size = pl.depth if solver == "tree" else pl.nrn_N
    ref = _oracle(oracle, solver, t, w, t, 3, size, "inv_r_reg", ())
    assert rel_l2(got, ref) <= 3e-5


@pytest.mark.parametrize("solver,size", [("tree", 2), ("tree", 1), ("nrnmm", 2), ("nrnmm", 1)])
def test_no_far_field_is_exact(gpu, oracle, solver, size):
    t = synth.points(321, 900, 2)
    s = synth.points(322, 800, 2)
    w = synth.normal(323, 800)
    got, _ = _solve(gpu, solver, t, w, s, 3, size, "inv_r_reg", ans0=np.full(900, 7.0))
    assert rel_l2(got, oracle.exact(t, w, s, KID["inv_r_reg"])) <= 1e-13  # ans overwritten


def test_tree_depth3_equals_nrnmm_4(gpu):
    t = synth.points(331, 5000, 3)
    w = synth.normal(332, 5000)
    a, _ = _solve(gpu, "tree", t, w, t, 4, 3, "gauss", (0.2,), shared=True)
    b, _ = _solve(gpu, "nrnmm", t, w, t, 4, 4, "gauss", (0.2,), shared=True)
    assert rel_l2(a, b) <= 1e-13


def test_unweighted_sticky_sizes_and_time_log(gpu, oracle):
    """Unit weights (README.md:75-78), the auto sizes kept for later solves,
    and the stage keys of the reference's time log."""
    t = synth.points(341, 6000, 2)
    s = synth.points(342, 5000, 2)
    pl = gpu.Plan(2, np.float64)
    pl.set_solver_type("nrnmm")
    a, _ = _solve(gpu, "nrnmm", t, None, s, 4, -1, "log_r", plan=pl)
    n_auto = pl.nrn_N
    assert n_auto == oracle.nrn_auto(5000, 2)
    assert rel_l2(a, oracle.nrnmm(t, np.ones(5000), s, 4, n_auto, KID["log_r"])) <= 1e-12
    t2 = synth.points(343, 300, 2)
    b, _ = _solve(gpu, "nrnmm", t2, None, s, 4, -1, "log_r", plan=pl)
    assert pl.nrn_N == n_auto
    assert rel_l2(b, oracle.nrnmm(t2, np.ones(5000), s, 4, n_auto, KID["log_r"])) <= 1e-12
    log = pl.time_log()
    for k in ("P2M", "M2P", "P2P", "total time(nrnmm)"):
        assert k in log
    pl.set_solver_type("tree")
    pl.set_Depth(-1)
    c, _ = _solve(gpu, "tree", t, None, s, 4, -1, "log_r", plan=pl)
    assert pl.depth == oracle.auto_depth(6000, 2)
    assert rel_l2(c, oracle.tree(t, np.ones(5000), s, 4, pl.depth, KID["log_r"])) <= 1e-12
    log = pl.time_log()
    for k in ("M2M", "M2P", "total time(tree)"):
        assert k in log


def test_empty_and_unsupported(gpu):
    import torch
    from fmma_amd import FMMAError
    pl = gpu.Plan(3, np.float64)
    pl.set_solver_type("nrnmm")
    x = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    s = torch.rand((50, 3), dtype=torch.float64, device="cuda")
    a = torch.zeros(0, dtype=torch.float64, device="cuda")
    pl.solve(x, None, s, a)  # no targets: nothing to do, the auto nrn_N is still fixed
    assert pl.nrn_N == int(np.sqrt(3.0) * 50 ** (1 / 6)) + 1
    g = torch.zeros((50, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(FMMAError):
        pl.gradient(s, torch.ones(50, dtype=torch.float64, device="cuda"), s, g)
