"""Full-size parity: the GPU fmm against the oracle (the CPU restatement of
the reference, OpenMP on the host) at the benchmark's own sizes -- config 3
(3-D 1/r, N = 4,194,304, 4 nodes per dimension, Depth 6: 128- and 256-box M2L
tiles, 128-point leaves, 15-bit keys) and config 2 (2-D log|x-y|, N = 2^20, 6
nodes per dimension, Depth 8).  Bars: FP64 1e-12; FP32 M2L 2e-6; BF16x3 2e-6
(3-D) and 1e-5 (2-D log, see test_m2l_forms_2d_log)."""
import numpy as np
import pytest

from fmma_amd import synth
from tests.conftest import rel_l2
from tests.test_gpu_parity import KID, gpu_solve

pytestmark = pytest.mark.gpu

_REF = {}


def _oracle(oracle, key, x, w, p, depth, kid):
    if key not in _REF:
        print("oracle %s ..." % key, flush=True)
        _REF[key] = oracle.fmm(x, w, x, p, depth, 1, kid)
    return _REF[key]


@pytest.mark.parametrize("prec,tol", [("native", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-6)])
def test_fullsize_config3(gpu, oracle, prec, tol):
    n = 1 << 22
    x = synth.points(1, n, 3)
    w = synth.uniform(2, n)
    print("config 3 full size, %s" % prec, flush=True)
    got, pl = gpu_solve(gpu, "fmm", x, w, x, 3, 6, 1, "inv_r_reg", shared=True, precision=prec)
    ref = _oracle(oracle, "c3", x, w, 3, 6, KID["inv_r_reg"])
    err = rel_l2(got, ref)
    print("  N=%d Depth 6 %s vs oracle %.3e, tiles %s" % (n, prec, err, pl.m2l_tiles(batch=False)), flush=True)
    assert err <= tol, err


@pytest.mark.parametrize("prec,tol", [("native", 1e-12), ("fp32", 2e-6), ("bf16x3", 1e-5)])
def test_fullsize_config2(gpu, oracle, prec, tol):
    n = 1 << 20
    x = synth.points(1, n, 2)
    w = synth.normal(2, n)
    print("config 2 full size, %s" % prec, flush=True)
    got, pl = gpu_solve(gpu, "fmm", x, w, x, 5, 8, 1, "log_r", shared=True, precision=prec)
    ref = _oracle(oracle, "c2", x, w, 5, 8, KID["log_r"])
    err = rel_l2(got, ref)
    print("  N=%d Depth 8 %s vs oracle %.3e" % (n, prec, err), flush=True)
    assert err <= tol, err
