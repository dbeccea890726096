"""Inputs the reference cannot bin: every point at one position (the cube has
size 0, src/fmm_core.cpp:261 divides by it) and non-finite coordinates.  The
bounding cube is checked before any kernel uses it, so these solves fail with
FMMA_ERR_INVALID and a message instead of binning garbage; the plan stays
usable afterwards."""
import numpy as np
import pytest

from fmma_amd import synth

pytestmark = pytest.mark.gpu


def _plan(fm, solver, dim=3):
    pl = fm.Plan(dim, np.float64)
    pl.set_solver_type(solver)
    pl.set_poly_ord(2)
    pl.set_Depth(4)
    pl.set_kernel("inv_r_reg")
    return pl


def _run(pl, x, w):
    import torch
    xt = torch.from_numpy(x).cuda()
    wt = torch.from_numpy(w).cuda()
    a = torch.full((x.shape[0],), 7.0, dtype=torch.float64, device="cuda")
    pl.solve(xt, wt, xt, a)
    torch.cuda.synchronize()
    return a.cpu().numpy()


@pytest.mark.parametrize("solver", ["fmm", "tree", "nrnmm"])
@pytest.mark.parametrize("case", ["same_point", "nan", "inf", "neg_inf"])
def test_degenerate_inputs_rejected(gpu, solver, case):
    n = 2000
    x = synth.points(201, n, 3)
    w = synth.uniform(202, n)
    if case == "same_point":
        x[:] = [0.25, 0.5, 0.75]
    elif case == "nan":
        x[17, 1] = np.nan
    elif case == "inf":
        x[5, 0] = np.inf
    else:
        x[1999, 2] = -np.inf
    pl = _plan(gpu, solver)
    with pytest.raises(gpu.FMMAError) as e:
        _run(pl, x, w)
    assert e.value.code == 1
    msg = str(e.value)
    assert ("same position" in msg) if case == "same_point" else ("NaN or infinite" in msg), msg
    # the plan still solves valid input
    y = synth.points(203, n, 3)
    out = _run(pl, y, w)
    assert np.all(np.isfinite(out))


def test_degenerate_ans_untouched(gpu):
    """A rejected single solve leaves ans as it was."""
    x = np.full((500, 2), 0.5)
    w = np.ones(500)
    pl = _plan(gpu, "fmm", dim=2)
    import torch
    xt = torch.from_numpy(x).cuda()
    wt = torch.from_numpy(w).cuda()
    a = torch.full((500,), 3.0, dtype=torch.float64, device="cuda")
    with pytest.raises(gpu.FMMAError):
        pl.solve(xt, wt, xt, a)
    assert torch.all(a == 3.0).item()


def test_degenerate_batch_problem_rejected(gpu):
    """A batch fails when one problem cannot be binned; it names the problem."""
    import torch
    counts = [3000, 4000, 3500]
    x = np.concatenate([synth.points(210 + i, c, 3) for i, c in enumerate(counts)])
    x[3000:7000] = 0.125  # problem 1: every point at one position
    w = synth.normal(220, x.shape[0])
    pl = _plan(gpu, "fmm")
    pl.set_Depth(-1)
    pl.set_trans_sym_flag(1)
    xt = torch.from_numpy(x).cuda()
    wt = torch.from_numpy(w).cuda()
    a = torch.zeros(x.shape[0], dtype=torch.float64, device="cuda")
    with pytest.raises(gpu.FMMAError) as e:
        pl.solve_batch(counts, xt, wt, counts, xt, a)
    assert "batch problem 1" in str(e.value)
