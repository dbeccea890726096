"""Python kernels on the GPU through `import pyfmma`, as the reference's
Python tutorials use them (tutorials/tutorial0{1..5}/*.py): a callable over
the coordinates, set with set_fn2 (or set_fn1 for f(x - y)), then exact,
tree and fmm solves, weighted and unweighted, in 1-D and 2-D.  The
tutorials' kernel 1/|x-y|^2 (written as 1/(d*d) with d = |x0 - y0| in 1-D and
1/|dx^2 + dy^2| in 2-D) is the oracle's kernel 8, so every result is checked
against the oracle: FP64 at 1e-12 (exact 1e-13), FP32 at 3e-5.  Each
workload is the tutorials' (N = 10, sources i/N, targets (i + 0.5)/N, weights
i/N, Depth 3, order 3) and a larger one of the same form."""
import numpy as np
import pytest

from fmma_amd import synth
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

INV_R2 = 8


def k1d(x, y):
    d = np.fabs(x[0] - y[0])
    return 1.0 / (d * d)


def k2d(x, y):
    d = np.fabs((x[0] - y[0]) * (x[0] - y[0]) + (x[1] - y[1]) * (x[1] - y[1]))
    return 1.0 / d


def _grid(n, dim):
    src = np.zeros((n, dim))
    tgt = np.zeros((n, dim))
    for i in range(n):
        src[i, :] = (i + 0.0) / n
        tgt[i, :] = (i + 0.5) / n
    return src, tgt, np.arange(n) / n


@pytest.mark.parametrize("solver,weighted", [("exact", False), ("tree", False), ("fmm", False), ("fmm", True)])
def test_tutorial_1d_workloads(oracle, solver, weighted):
    import pyfmma
    print("1-D python kernel, %s, weighted %s" % (solver, weighted), flush=True)
    f = pyfmma.fmmad1()
    f.set_fn2(k1d)
    f.set_solver_type(solver)
    f.set_Depth(3)
    f.set_poly_ord(3)
    src, tgt, w = _grid(10, 1)
    ans = np.zeros(10)
    if weighted:
        f.solve(tgt, w, src, ans)
    else:
        f.solve(tgt, src, ans)
    ww = w if weighted else np.ones(10)
    if solver == "exact":
        ref = oracle.exact(tgt, ww, src, INV_R2)
        direct = np.array([sum(ww[j] * k1d(tgt[i], src[j]) for j in range(10)) for i in range(10)])
        assert rel_l2(ans, direct) <= 1e-13
    elif solver == "tree":
        ref = oracle.tree(tgt, ww, src, 3, 3, INV_R2)
    else:
        ref = oracle.fmm(tgt, ww, src, 3, 3, 0, INV_R2)
    assert rel_l2(ans, ref) <= 1e-12


def test_tutorial_2d_workload(oracle):
    import pyfmma
    f = pyfmma.fmmad2()
    f.set_fn2(k2d)
    f.set_solver_type("fmm")
    f.set_Depth(3)
    f.set_poly_ord(3)
    src, tgt, w = _grid(10, 2)
    ans = np.zeros(10)
    f.solve(tgt, w, src, ans)
    assert rel_l2(ans, oracle.fmm(tgt, w, src, 3, 3, 0, INV_R2)) <= 1e-12


@pytest.mark.parametrize("dim,cls,n,p,tol", [(1, "fmmad1", 20000, 5, 1e-12), (2, "fmmad2", 30000, 4, 1e-12),
                                             (2, "fmmaf2", 20000, 4, 3e-5)])
def test_python_kernel_larger_fmm(oracle, dim, cls, n, p, tol):
    """The tutorials' kernels on random points with a far field (auto
    Depth, trans_sym_flag 0: translation invariance is found in the traced
    expression), targets = sources excluded (the kernel is inf at r = 0)."""
    import pyfmma
    f = getattr(pyfmma, cls)()
    f.set_fn2(k1d if dim == 1 else k2d)
    f.set_solver_type("fmm")
    f.set_poly_ord(p)
    dt = np.float64 if cls.startswith("fmmad") else np.float32
    t = synth.points(501, n, dim).astype(dt)
    s = synth.points(502, n, dim).astype(dt)
    w = synth.normal(503, n).astype(dt)
    ans = np.zeros(n, dtype=dt)
    f.solve(t, w, s, ans)
    d = f.get_Depth()
    ref = oracle.fmm(t.astype(np.float64), w.astype(np.float64), s.astype(np.float64), p, d, 0, INV_R2)
    err = rel_l2(ans, ref)
    print("  %s N=%d Depth %d vs oracle %.3e" % (cls, n, d, err))
    assert err <= tol, err


def test_python_kernel_gradient_and_transpose(gpu, oracle):
    """Plan.set_fn with a Python kernel: its traced derivative drives the
    gradient; the transpose uses its symmetry; against the oracle's kernel 8."""
    import torch
    n = 8000
    t = torch.from_numpy(synth.points(511, n, 2)).cuda()
    s = torch.from_numpy(synth.points(512, n, 2)).cuda()
    w = torch.from_numpy(synth.normal(513, n)).cuda()
    cot = torch.from_numpy(synth.normal(514, n)).cuda()
    pl = gpu.Plan(2, "float64")
    pl.set_fn2(k2d)
    pl.set_solver_type("fmm")
    pl.set_poly_ord(4)
    pl.set_Depth(5)
    ans = torch.zeros(n, dtype=torch.float64, device="cuda")
    pl.solve(t, w, s, ans)
    g = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    pl.gradient(t, w, s, g)
    tr = torch.zeros(n, dtype=torch.float64, device="cuda")
    pl.solve_transpose(t, cot, s, tr)
    torch.cuda.synchronize()
    tn, sn, wn, cn = (a.cpu().numpy() for a in (t, s, w, cot))
    assert rel_l2(ans.cpu().numpy(), oracle.fmm(tn, wn, sn, 4, 5, 0, INV_R2)) <= 1e-12
    assert rel_l2(g.cpu().numpy(), oracle.fmm_grad(tn, wn, sn, 4, 5, 0, INV_R2)) <= 1e-12
    assert rel_l2(tr.cpu().numpy(), oracle.fmm_transpose(tn, cn, sn, 4, 5, 0, INV_R2)) <= 1e-12
