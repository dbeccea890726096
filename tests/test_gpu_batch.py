"""The batching form: many independent problems in one pass of batched
kernels (Plan.solve_batch), replayed from one captured hipGraph, with M2L
tile widths picked by the autotuner.

Each problem of a batch must come out as a single solve of it would (its own
bounding cube and Depth): checked against a loop of Plan.solve calls on the
same slices -- bit for bit in FP64, since the batched kernels do the same
arithmetic in the same order -- and against the oracle on a few problems
(1e-12).  Graph replays must equal the eager pass, also after the inputs
change in place.
"""
import numpy as np
import pytest

from fmma_amd import synth
from tests.conftest import rel_l2
from tests.test_gpu_parity import KID

pytestmark = pytest.mark.gpu


def _plan(gpu, kernel="yukawa", params=(2.0,), p=4, precision="native", dtype="float64", depth=-1):
    pl = gpu.Plan(3, dtype)
    pl.set_solver_type("fmm")
    pl.set_poly_ord(p)
    pl.set_Depth(depth)
    pl.set_trans_sym_flag(1)
    pl.set_kernel(kernel, *params)
    pl.set_m2l_precision(precision)
    return pl


def _loop(gpu, counts, x, w, s=None, s_counts=None, **kw):
    """Per-problem solves on slices of the same tensors (the reference way)."""
    import torch
    ans = torch.zeros(x.shape[0], dtype=x.dtype, device="cuda")
    s = x if s is None else s
    s_counts = counts if s_counts is None else s_counts
    t0 = s0 = 0
    for nt, ns in zip(counts, s_counts):
        pl = _plan(gpu, **kw)  # a fresh plan: its own auto Depth, as a fresh reference object
        tt = x[t0:t0 + nt]
        ss = tt if s is x and s_counts is counts else s[s0:s0 + ns]
        pl.solve(tt, w[s0:s0 + ns], ss, ans[t0:t0 + nt])
        t0 += nt
        s0 += ns
    torch.cuda.synchronize()
    return ans


def _cuda(*arrs, dtype=np.float64):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).cuda() for a in arrs]


@pytest.mark.parametrize("precision", ["native", "fp32", "bf16x3"])
def test_batch_equals_loop_of_single_solves(gpu, precision):
    import torch
    counts = synth.batch_counts(201, 24, 500, 9000)
    x, w = _cuda(synth.batch_points(202, counts, 3), synth.normal(203, int(counts.sum())))
    print("counts", counts.tolist(), flush=True)
    pl = _plan(gpu, precision=precision)
    ans = torch.zeros(x.shape[0], dtype=torch.float64, device="cuda")
    pl.solve_batch(counts, x, w, counts, x, ans)
    torch.cuda.synchronize()
    ref = _loop(gpu, counts, x, w, precision=precision)
    a, r = ans.cpu().numpy(), ref.cpu().numpy()
    print("  %s: max |batch - loop| %.3e, rel %.3e, tiles %s" % (precision, np.abs(a - r).max(), rel_l2(a, r),
                                                             pl.m2l_tiles()))
    assert rel_l2(a, r) <= 1e-14


def test_batch_matches_oracle_and_replays(gpu, oracle):
    import torch
    counts = synth.batch_counts(211, 16, 600, 6000)
    xn = synth.batch_points(212, counts, 3)
    wn = synth.normal(213, int(counts.sum()))
    x, w = _cuda(xn, wn)
    pl = _plan(gpu)
    ans = torch.zeros(x.shape[0], dtype=torch.float64, device="cuda")
    pl.solve_batch(counts, x, w, counts, x, ans)  # eager, runs the autotuner
    torch.cuda.synchronize()
    first = ans.clone()
    off = np.concatenate([[0], np.cumsum(counts)])
    for p in (0, 7, 15):
        sl = slice(off[p], off[p + 1])
        ref = oracle.fmm(xn[sl], wn[sl], xn[sl], 4, -1, 1, KID["yukawa"], (2.0,))
        e = rel_l2(first.cpu().numpy()[sl], ref)
        print("  problem %d (n=%d) vs oracle %.3e" % (p, counts[p], e))
        assert e <= 1e-12
    # graph: capture + launch, then replays
    for _ in range(3):
        ans.zero_()
        pl.solve_batch(counts, x, w, counts, x, ans)
        torch.cuda.synchronize()
        assert torch.equal(ans, first)
    # new values in the same buffers: the replay computes with them
    w.mul_(-0.5)
    x.add_(0.25)
    pl.solve_batch(counts, x, w, counts, x, ans)
    torch.cuda.synchronize()
    ref = _loop(gpu, counts, x, w)
    assert rel_l2(ans.cpu().numpy(), ref.cpu().numpy()) <= 1e-14
    log = pl.time_log()
    assert log["total time(fmm batch)"] > 0


def test_batch_separate_targets_and_graph_off(gpu):
    import torch
    tc = synth.batch_counts(221, 10, 800, 5000)
    sc = synth.batch_counts(222, 10, 500, 7000)
    t, s, w = _cuda(synth.batch_points(223, tc, 3), synth.batch_points(224, sc, 3), synth.normal(225, int(sc.sum())))
    pl = _plan(gpu, kernel="inv_r_reg", params=(), p=3)
    pl.set_batch_graph(False)
    pl.set_m2l_autotune(False)
    ans = torch.zeros(t.shape[0], dtype=torch.float64, device="cuda")
    pl.solve_batch(tc, t, w, sc, s, ans)
    torch.cuda.synchronize()
    ref = _loop(gpu, tc, t, w, s=s, s_counts=sc, kernel="inv_r_reg", params=(), p=3)
    assert rel_l2(ans.cpu().numpy(), ref.cpu().numpy()) <= 1e-14


def test_batch_fixed_depth_and_float32(gpu):
    import torch
    counts = synth.batch_counts(231, 12, 2000, 20000)
    xn = synth.batch_points(232, counts, 3)
    wn = synth.normal(233, int(counts.sum()))
    x, w = _cuda(xn, wn, dtype=np.float32)
    pl = _plan(gpu, dtype="float32", depth=4)
    ans = torch.zeros(x.shape[0], dtype=torch.float32, device="cuda")
    pl.solve_batch(counts, x, w, counts, x, ans)
    pl.solve_batch(counts, x, w, counts, x, ans)  # graph
    torch.cuda.synchronize()
    ref = _loop(gpu, counts, x, w, dtype="float32", depth=4)
    assert rel_l2(ans.cpu().numpy(), ref.cpu().numpy()) <= 1e-6


def test_batch_errors(gpu):
    import torch
    from fmma_amd import FMMAError
    x, w = _cuda(synth.points(241, 300, 3), synth.normal(242, 300))
    pl = _plan(gpu)
    ans = torch.zeros(300, dtype=torch.float64, device="cuda")
    with pytest.raises(FMMAError):  # auto Depth 2 for 150 points: no far field to batch
        pl.solve_batch([150, 150], x, w, [150, 150], x, ans)
    with pytest.raises(FMMAError):
        pl.solve_batch([300, 0], x, w, [300, 0], x, ans)
