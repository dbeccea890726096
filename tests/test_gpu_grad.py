"""Differentiation on the GPU: the gradient with respect to the target
coordinates (Plan.gradient), the transpose in the weights
(Plan.solve_transpose) and the torch.autograd.Function built on them
(fmma_amd.autograd).

Bars: FP64 fmm gradient and transpose within 1e-12 of the oracle's (which
central differences and the adjoint identity pin, tests/test_oracle_grad.py);
FP64 exact gradient within 1e-13; reuse of the forward's expansions bitwise
equal to recomputing them; autograd gradients equal to the entry points'
results, checked against central differences on 64 coordinates and against
the exact solver's gradients (within the fmm approximation error, stated in
the test); the FP32 form within 1e-4 of FP64.
"""
import numpy as np
import pytest

from fmma_amd import synth
from tests.conftest import rel_l2
from tests.test_gpu_parity import KID

pytestmark = pytest.mark.gpu


def _plan(gpu, dim, kernel, params=(), p=3, depth=-1, ts=1, dtype="float64", solver="fmm"):
    pl = gpu.Plan(dim, dtype)
    pl.set_solver_type(solver)
    pl.set_poly_ord(p)
    pl.set_Depth(depth)
    pl.set_trans_sym_flag(ts)
    pl.set_kernel(kernel, *params)
    return pl


def _cuda(*arrs, dtype=np.float64):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).cuda() for a in arrs]


GRAD_CASES = [
    dict(id="3d_invr_sep", dim=3, nt=6000, ns=20000, k="inv_r_reg", params=(), p=3, depth=4, ts=1, src="uniform"),
    dict(id="3d_invr_ts0", dim=3, nt=3000, ns=5000, k="inv_r_reg", params=(), p=2, depth=4, ts=0, src="uniform"),
    dict(id="3d_gauss_mix", dim=3, nt=20000, ns=30000, k="gauss", params=(0.05,), p=3, depth=5, ts=1,
         src="mixture"),
    dict(id="2d_log_shared", dim=2, nt=16384, ns=16384, k="log_r", params=(), p=5, depth=6, ts=1, src="shared"),
    dict(id="2d_yukawa", dim=2, nt=5000, ns=8000, k="yukawa", params=(3.0,), p=4, depth=5, ts=1, src="uniform"),
    dict(id="1d_invdiff", dim=1, nt=4096, ns=4096, k="inv_diff", params=(), p=7, depth=9, ts=1, src="uniform"),
    dict(id="2d_sum2_general", dim=2, nt=3000, ns=2500, k="sum2_r2", params=(), p=4, depth=4, ts=0, src="uniform"),
    dict(id="3d_p6_d3", dim=3, nt=4000, ns=4000, k="inv_r_reg", params=(), p=6, depth=3, ts=1, src="uniform"),
    dict(id="3d_depth2", dim=3, nt=700, ns=900, k="gauss", params=(0.2,), p=3, depth=2, ts=1, src="uniform"),
]


def _case_inputs(c):
    t = synth.points(91, c["nt"], c["dim"])
    if c["src"] == "shared":
        s = t
    elif c["src"] == "mixture":
        s = synth.gaussian_mixture(92, c["ns"], c["dim"])
    else:
        s = synth.points(92, c["ns"], c["dim"])
    w = synth.normal(93, s.shape[0])
    return t, w, s


@pytest.mark.parametrize("case", GRAD_CASES, ids=[c["id"] for c in GRAD_CASES])
def test_fmm_gradient_and_transpose_match_oracle(gpu, oracle, case):
    import torch
    print("case", case["id"], flush=True)
    t, w, s = _case_inputs(case)
    g = synth.normal(94, t.shape[0])
    kid, p, depth, ts, prm = KID[case["k"]], case["p"], case["depth"], case["ts"], case["params"]
    pl = _plan(gpu, case["dim"], case["k"], prm, p, depth, ts)
    tt, ww, gg = _cuda(t, w, g)
    ss = tt if case["src"] == "shared" else _cuda(s)[0]
    grad = torch.empty_like(tt)
    pl.gradient(tt, ww, ss, grad, reuse=False)
    out = torch.empty_like(ww)
    pl.solve_transpose(tt, gg, ss, out)
    torch.cuda.synchronize()
    ref_g = oracle.fmm_grad(t, w, s, p, depth, ts, kid, prm)
    ref_t = oracle.fmm_transpose(t, g, s, p, depth, ts, kid, prm)
    eg = rel_l2(grad.cpu().numpy(), ref_g)
    et = rel_l2(out.cpu().numpy(), ref_t)
    print("  gradient vs oracle %.3e, transpose vs oracle %.3e" % (eg, et))
    assert eg <= 1e-12, eg
    assert et <= 1e-12, et


@pytest.mark.parametrize("name,dim,params", [("inv_r_reg", 3, ()), ("gauss", 3, (0.1,)), ("log_r", 2, ()),
                                             ("inv_diff", 1, ()), ("sum_r2", 2, ())])
def test_exact_gradient_and_transpose_match_oracle(gpu, oracle, name, dim, params):
    import torch
    t = synth.points(101, 1500, dim)
    s = synth.points(102, 2500, dim)
    w = synth.normal(103, 2500)
    g = synth.normal(104, 1500)
    pl = _plan(gpu, dim, name, params, solver="exact")
    tt, ww, ss, gg = _cuda(t, w, s, g)
    grad = torch.empty_like(tt)
    pl.gradient(tt, ww, ss, grad)
    out = torch.empty_like(ww)
    pl.solve_transpose(tt, gg, ss, out)
    torch.cuda.synchronize()
    kid = KID[name]
    assert rel_l2(grad.cpu().numpy(), oracle.exact_grad(t, w, s, kid, params)) <= 1e-13
    assert rel_l2(out.cpu().numpy(), oracle.SYMMETRY[kid] * oracle.exact(s, g, t, kid, params)) <= 1e-13


def test_gradient_reuse_equals_recompute_and_survives_other_solves(gpu):
    import torch
    t = synth.points(111, 20000, 3)
    s = synth.gaussian_mixture(112, 30000, 3)
    w = synth.normal(113, 30000)
    pl = _plan(gpu, 3, "gauss", (0.05,), depth=5)
    tt, ss, ww = _cuda(t, s, w)
    ans = torch.zeros(20000, dtype=torch.float64, device="cuda")
    pl.solve(tt, ww, ss, ans)
    g_reuse = torch.empty_like(tt)
    pl.gradient(tt, ww, ss, g_reuse, reuse=True)
    g_new = torch.empty_like(tt)
    pl.gradient(tt, ww, ss, g_new, reuse=False)
    assert torch.equal(g_reuse, g_new)
    # another solve on other arrays in between: reuse must notice and recompute
    t2, = _cuda(synth.points(114, 5000, 3))
    pl.solve(t2, ww, ss, torch.zeros(5000, dtype=torch.float64, device="cuda"))
    g_after = torch.empty_like(tt)
    pl.gradient(tt, ww, ss, g_after, reuse=True)
    assert torch.equal(g_after, g_new)
    # unweighted and empty forms
    g_ones = torch.empty_like(tt)
    pl.gradient(tt, None, ss, g_ones, reuse=False)
    assert torch.isfinite(g_ones).all()
    empty = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    g_empty = torch.full_like(tt, 7.0)
    pl.gradient(tt, torch.zeros(0, dtype=torch.float64, device="cuda"), empty, g_empty)
    assert torch.count_nonzero(g_empty).item() == 0


def _config4(n_t, n_s):
    c = synth.config("c4_gauss", n_s)
    return synth.points(5, n_t, 3), c["sources"], c["weights"], c["params"]


def _interior_targets(oracle, t, s, depth, count, h):
    """Indices of `count` targets that are not extreme in any coordinate and
    stay inside their leaf box when moved by +-h."""
    _, rel = oracle.leaf_boxes(t, s, t, depth, rel=True)
    o, Len = oracle.bbox(t, s)
    margin = 4 * h / (Len / 2 ** (depth - 1)) * 2
    ok = np.all(np.abs(rel) < 1 - margin, axis=1)
    ok &= np.all(t > t.min(0) + 1e-9, axis=1) & np.all(t < t.max(0) - 1e-9, axis=1)
    idx = np.flatnonzero(ok)
    return idx[synth.sample_indices(121, idx.size, count)]


@pytest.mark.parametrize("n", [1 << 17, 1 << 20], ids=["n131k", "n1m"])
def test_autograd_config4_fd_and_exact(gpu, oracle, n):
    """The differentiation workload: 3-D Gaussian h = 0.05, sources from a
    16-component mixture, targets uniform, loss = sum_i g_i u_i with g ~ N(0,1);
    backward with respect to weights and target coordinates."""
    import torch
    from fmma_amd import autograd as fa
    t, s, w, prm = _config4(n, n)
    g = synth.normal(122, n)
    print("n", n, flush=True)
    pl = fa.make_plan(3, torch.float64, "gauss", prm, poly_ord=3)
    tt, ss, ww, gg = _cuda(t, s, w, g)
    tt.requires_grad_(True)
    ww.requires_grad_(True)
    u = fa.fmm(tt, ww, ss, pl)
    loss = (gg * u).sum()
    loss.backward()
    depth = pl.depth
    gx, gw = tt.grad.detach().clone(), ww.grad.detach().clone()

    # the autograd results are the entry points' results
    grad = torch.empty_like(tt)
    pl.gradient(tt.detach(), ww.detach(), ss, grad, reuse=False)
    assert torch.equal(gx, gg[:, None] * grad)
    tr = torch.empty_like(ww)
    pl.solve_transpose(tt.detach(), gg, ss, tr)
    assert torch.equal(gw, tr)

    # central differences on 64 coordinates (two extra solves: u_i depends on x_i only)
    h = 1e-6
    idx = _interior_targets(oracle, t, s, depth, 64, h)
    dims = np.arange(64) % 3
    tp, tm = t.copy(), t.copy()
    tp[idx, dims] += h
    tm[idx, dims] -= h
    up = torch.zeros(n, dtype=torch.float64, device="cuda")
    um = torch.zeros_like(up)
    pl.solve(*_cuda(tp), ww.detach(), ss, up)
    pl.solve(*_cuda(tm), ww.detach(), ss, um)
    fd = ((up - um).cpu().numpy()[idx] / (2 * h)) * g[idx]
    an = gx.cpu().numpy()[idx, dims]
    e_fd = rel_l2(an, fd)
    print("  x-grad vs central differences on 64 coordinates %.3e" % e_fd)
    assert e_fd < 1e-6, e_fd

    # against the exact solver's gradients (the fmm's own approximation error)
    ex = fa.make_plan(3, torch.float64, "gauss", prm, solver="exact")
    sample = synth.sample_indices(123, n, 2048)
    ts_, = _cuda(t[sample])
    ge = torch.empty_like(ts_)
    ex.gradient(ts_, ww.detach(), ss, ge)
    e_ex = rel_l2(gx.cpu().numpy()[sample], g[sample][:, None] * ge.cpu().numpy())
    wt = torch.empty_like(ww)
    ex.solve_transpose(tt.detach(), gg, ss, wt)
    e_wt = rel_l2(gw.cpu().numpy(), wt.cpu().numpy())
    print("  x-grad vs exact-solver gradient %.3e, w-grad vs exact transpose %.3e" % (e_ex, e_wt))
    assert e_ex < 1e-2 and e_wt < 1e-2

    # the FP32 form
    pl32 = fa.make_plan(3, torch.float32, "gauss", prm, poly_ord=3, depth=depth)
    t32, s32, w32, g32 = _cuda(t, s, w, g, dtype=np.float32)
    t32.requires_grad_(True)
    w32.requires_grad_(True)
    (g32 * fa.fmm(t32, w32, s32, pl32)).sum().backward()
    e32x = rel_l2(t32.grad.double().cpu().numpy(), gx.cpu().numpy())
    e32w = rel_l2(w32.grad.double().cpu().numpy(), gw.cpu().numpy())
    print("  fp32 vs fp64: x-grad %.3e, w-grad %.3e" % (e32x, e32w))
    assert e32x < 1e-4 and e32w < 1e-4


def test_autograd_recompute_and_shared_plan(gpu):
    import torch
    from fmma_amd import autograd as fa
    t, s, w, prm = _config4(30000, 40000)
    pl = fa.make_plan(3, torch.float64, "gauss", prm, depth=5)
    tt, ss, ww = _cuda(t, s, w)
    grads = []
    for recompute in (False, True):
        x = tt.clone().requires_grad_(True)
        v = ww.clone().requires_grad_(True)
        u = fa.fmm(x, v, ss, pl, recompute=recompute)
        # a second call on the same plan before the backward
        u2 = fa.fmm(x * 1.0, v, ss, pl, recompute=recompute)
        (u.sum() + 2 * u2.sum()).backward()
        grads.append((x.grad.clone(), v.grad.clone()))
    assert torch.equal(grads[0][0], grads[1][0])
    assert torch.equal(grads[0][1], grads[1][1])


def test_autograd_rejects_source_gradients(gpu):
    import torch
    from fmma_amd import autograd as fa
    x, = _cuda(synth.points(131, 2000, 2))
    w, = _cuda(synth.normal(132, 2000))
    pl = fa.make_plan(2, torch.float64, "log_r", poly_ord=3, depth=4)
    x.requires_grad_(True)
    u = fa.fmm(x, w, x, pl)
    with pytest.raises(NotImplementedError):
        u.sum().backward()
