"""Python kernels (fmma_amd.jit): tracing a callable into an expression, its
symmetry and translation invariance, the derivative, and the compiled
functor against the callable -- on the host (hipcc builds the small library;
nothing here launches a kernel)."""
import math
import os

import numpy as np
import pytest

from fmma_amd import jit


@pytest.fixture(autouse=True, scope="module")
def _cache(tmp_path_factory):
    old = os.environ.get("FMMA_JIT_CACHE")
    os.environ["FMMA_JIT_CACHE"] = str(tmp_path_factory.mktemp("jit"))
    yield
    if old is None:
        os.environ.pop("FMMA_JIT_CACHE", None)
    else:
        os.environ["FMMA_JIT_CACHE"] = old


def tut1d(x, y):  # the form of the reference's 1-D tutorial kernels
    d = np.fabs(x[0] - y[0])
    return 1.0 / (d * d)


def tut2d(x, y):  # the form of its 2-D tutorial kernel
    d = np.fabs((x[0] - y[0]) * (x[0] - y[0]) + (x[1] - y[1]) * (x[1] - y[1]))
    return 1.0 / d


def yukawa1(d):
    r = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    return np.exp(-2.0 * r) / r


def odd(x, y):
    return (x[0] - y[0]) * np.exp(-((x[0] - y[0]) ** 2 + (x[1] - y[1]) ** 2))


def general(x, y):
    return (1.0 + x[0]) * np.exp(-(x[0] - y[0]) ** 2 - 2.0 * (x[1] - y[1]) ** 2)


@pytest.mark.parametrize("fn,dim,nargs,sym,ti", [(tut1d, 1, 2, 1, True), (tut2d, 2, 2, 1, True),
                                                 (yukawa1, 3, 1, 1, True), (odd, 2, 2, -1, True),
                                                 (general, 2, 2, 0, False)])
def test_trace_flags(fn, dim, nargs, sym, ti):
    spec = jit.KernelSpec(fn, dim, nargs, "float64")
    assert spec.symmetry == sym
    assert spec.translation_invariant == ti


@pytest.mark.parametrize("fn,dim,nargs", [(tut2d, 2, 2), (yukawa1, 3, 1), (general, 2, 2)])
def test_compiled_functor_and_gradient(fn, dim, nargs):
    """The compiled host functor equals the callable; its gradient equals
    central differences of the callable (w.r.t. x, or d for f(x - y))."""
    spec = jit.KernelSpec(fn, dim, nargs, "float64")
    L = jit.compile_kernel(spec)
    jit.check_against_python(L, fn, spec)
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(16):
        x, y = rng.random(dim), rng.random(dim)
        if np.linalg.norm(x - y) < 0.2:
            continue
        g = np.zeros(dim)
        L.fmma_jit_eval_grad(np.ascontiguousarray(x).ctypes.data, np.ascontiguousarray(y).ctypes.data, g.ctypes.data)
        for k in range(dim):
            e = np.zeros(dim)
            e[k] = h
            call = (lambda a: fn(list(a), list(y))) if nargs == 2 else (lambda a: fn(list(a - y)))
            fd = (call(x + e) - call(x - e)) / (2 * h)
            assert abs(g[k] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, g[k], fd)


def test_numpy_vector_forms_trace():
    """np.array over the point and np.linalg.norm trace through object arrays."""
    spec = jit.KernelSpec(lambda x, y: 1.0 / np.linalg.norm(np.array(x) - np.array(y)), 3, 2, "float64")
    assert spec.translation_invariant and spec.symmetry == 1


def test_untraceable_kernels_are_rejected():
    with pytest.raises(TypeError, match="numpy functions"):
        jit.KernelSpec(lambda x, y: math.sqrt(x[0] - y[0]), 1, 2, "float64")
    with pytest.raises(TypeError, match="branch"):
        jit.KernelSpec(lambda x, y: 1.0 if x[0] > y[0] else 0.0, 1, 2, "float64")
    with pytest.raises(TypeError):
        jit.KernelSpec(lambda x, y: "text", 1, 2, "float64")
