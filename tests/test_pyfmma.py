"""The pyfmma drop-in: `import pyfmma` as with the reference (setup.py:6),
the reference module's classes and methods (cpython/pyfmma.cpp:140-168), its
shape errors (cpython/pyfmma.cpp:22-53), Python kernels given to set_fn1 /
set_fn2, and -- on the GPU -- its results against the oracle, used the way
the reference's tutorials use it (tutorials/tutorial0*/tutorial0*.py)."""
import numpy as np
import pytest

from fmma_amd import synth
from tests.conftest import rel_l2


def _mod():
    import pyfmma
    return pyfmma


def test_top_level_module_is_the_extension():
    import pyfmma
    from fmma_amd import pyfmma as ext
    assert pyfmma.fmmad3 is ext.fmmad3


def test_module_surface():
    pyfmma = _mod()
    for name in ("fmmad1", "fmmaf1", "fmmad2", "fmmaf2", "fmmad3", "fmmaf3"):
        c = getattr(pyfmma, name)
        for meth in ("solve", "check_blas", "set_fn1", "set_fn2", "set_solver_type", "set_nrn_N", "set_poly_ord",
                     "set_Depth", "set_trans_sym_flag", "set_kernel", "set_m2l_precision"):
            assert hasattr(c, meth), (name, meth)
    assert pyfmma.fmmad2().check_blas() is False


@pytest.mark.parametrize("bad,msg", [
    (lambda: (np.zeros(10), np.zeros(10), np.zeros((10, 2)), np.zeros(10)), "target shape must be 2"),
    (lambda: (np.zeros((10, 2)), np.zeros((10, 1)), np.zeros((10, 2)), np.zeros(10)), "source_weight shape"),
    (lambda: (np.zeros((10, 3)), np.zeros(10), np.zeros((10, 2)), np.zeros(10)), "target dimension must be 2"),
    (lambda: (np.zeros((10, 2)), np.zeros(10), np.zeros((10, 3)), np.zeros(10)), "source dimension must be 2"),
    (lambda: (np.zeros((10, 2)), np.zeros(10), np.zeros((10, 2)), np.zeros(9)), "ans num 9"),
    (lambda: (np.zeros((10, 2)), np.zeros(8), np.zeros((10, 2)), np.zeros(10)), "source num 10"),
    (lambda: (np.zeros((10, 2)), np.zeros(10), np.zeros((10, 2)), np.zeros(10, dtype=np.float32)), "float64"),
])
def test_shape_errors_raise_runtime_error(bad, msg):
    """Checked before any device work (so this runs without a GPU)."""
    f = _mod().fmmad2()
    with pytest.raises(RuntimeError, match=msg):
        f.solve(*bad())


def test_untraceable_python_kernel_is_rejected():
    """A callable that branches on a coordinate cannot be traced for the GPU."""
    f = _mod().fmmad1()
    with pytest.raises(TypeError, match="branch"):
        f.set_fn2(lambda x, y: 1.0 if x[0] > y[0] else -1.0)


def test_unknown_kernel_name():
    with pytest.raises(RuntimeError, match="unknown kernel"):
        _mod().fmmad3().set_kernel("coulomb")


@pytest.mark.gpu
def test_builtin_kernel_tutorial05_layout(oracle):
    """The layout of the reference's 2-D tutorial (points on the diagonal,
    weights i/N, Depth 3, order 3) with a built-in kernel by name."""
    pyfmma = _mod()
    f = pyfmma.fmmad2()
    f.set_kernel("sum_r2")
    f.set_solver_type("fmm")
    f.set_Depth(3)
    f.set_poly_ord(3)
    f.set_trans_sym_flag(1)
    N = 10
    source = np.zeros((N, 2))
    target = np.zeros((N, 2))
    w = np.zeros(N)
    for i in range(N):
        source[i] = (i + 0.0) / N
        target[i] = (i + 0.5) / N
        w[i] = (i + 0.0) / N
    ans = np.zeros(N)
    f.solve(target, w, source, ans)
    ref = oracle.fmm(target, w, source, 3, 3, 1, 6)
    assert rel_l2(ans, ref) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("cls,dim,dtype,tol", [("fmmad1", 1, np.float64, 1e-12), ("fmmad3", 3, np.float64, 1e-12),
                                               ("fmmaf2", 2, np.float32, 3e-5), ("fmmaf3", 3, np.float32, 3e-5)])
def test_pyfmma_matches_oracle(oracle, cls, dim, dtype, tol):
    f = getattr(_mod(), cls)()
    n = 6000
    t = synth.points(40, n, dim).astype(dtype)
    s = synth.points(41, n, dim).astype(dtype)
    w = synth.uniform(42, n).astype(dtype)
    kern, kid = ("inv_diff", 2) if dim == 1 else ("inv_r_reg", 1)
    f.set_kernel(kern)
    f.set_solver_type("fmm")
    f.set_poly_ord(4)
    f.set_trans_sym_flag(1)
    ans = np.zeros(n, dtype=dtype)
    f.solve(t, w, s, ans)
    ref = oracle.fmm(t.astype(np.float64), w.astype(np.float64), s.astype(np.float64), 4, f.get_Depth(), 1, kid)
    assert rel_l2(ans, ref) <= tol
    # unweighted overload and the exact solver
    f.set_solver_type("exact")
    f.solve(t, s, ans)
    assert rel_l2(ans, oracle.exact(t.astype(np.float64), np.ones(n), s.astype(np.float64), kid)) <= max(tol, 1e-13)


@pytest.mark.gpu
def test_pyfmma_time_log():
    f = _mod().fmmad3()
    x = synth.points(43, 20000, 3)
    ans = np.zeros(20000)
    f.set_kernel("inv_r_reg")
    f.set_solver_type("fmm")
    f.set_poly_ord(3)
    f.solve(x, x, ans)
    log = f.time_log()
    for k in ("P2M", "M2M", "M2L", "L2L", "L2P", "P2P", "total time(fmm)"):
        assert k in log
