"""The mutual P2P form: with targets = sources and a kernel that declares its
symmetry, each pair of distinct neighbour boxes is evaluated once and the
reverse sums are added by a combine pass in a fixed order.

Checked against the direct form (every target box against its whole 3^D
neighbourhood, as src/fmm.cpp:122-141 does) and against the oracle: the two
forms add the same terms in a different order, so they agree to rounding
(1e-13 relative), and both stay within the 1e-12 parity bar.  Cases cover
boxes of more than 256 points (several passes per box), empty boxes, leaf
level 1 (P2P adds onto ans) and every dimension.
"""
import numpy as np
import pytest

from fmma_amd import synth
from tests.conftest import rel_l2
from tests.test_gpu_parity import KID, gpu_solve

pytestmark = pytest.mark.gpu

CASES = [
    dict(id="3d_invr_d4", dim=3, n=32768, p=3, depth=4, k="inv_r_reg", w="u01"),
    dict(id="3d_invr_d5_mixture", dim=3, n=40000, p=3, depth=5, k="inv_r_reg", w="normal", src="mixture"),
    dict(id="3d_invr_d3_big_boxes", dim=3, n=30000, p=2, depth=3, k="inv_r_reg", w="u01"),
    dict(id="3d_yukawa", dim=3, n=16384, p=3, depth=4, k="yukawa", w="normal", params=(2.0,)),
    dict(id="3d_gauss_mixture", dim=3, n=20000, p=3, depth=5, k="gauss", w="normal", params=(0.05,),
         src="mixture"),
    dict(id="2d_log", dim=2, n=16384, p=5, depth=6, k="log_r", w="normal"),
    dict(id="2d_log_mixture", dim=2, n=20000, p=4, depth=5, k="log_r", w="normal", src="mixture"),
    dict(id="1d_log", dim=1, n=8000, p=7, depth=7, k="log_r", w="pm1"),
    dict(id="2d_gauss", dim=2, n=12000, p=5, depth=5, k="gauss", w="normal", params=(0.1,)),
]


def _points(case):
    dim, n = case["dim"], case["n"]
    if case.get("src") == "mixture":
        x = synth.gaussian_mixture(41, n, dim)
    else:
        x = synth.points(41, n, dim)
    w = {"u01": synth.uniform(42, n), "pm1": 2 * synth.uniform(42, n) - 1, "normal": synth.normal(42, n)}[case["w"]]
    return x, w


@pytest.mark.parametrize("case", CASES, ids=[c["id"] for c in CASES])
def test_mutual_matches_direct_and_oracle(gpu, oracle, case):
    x, w = _points(case)
    params = case.get("params", ())
    ts = case.get("ts", 1)
    print("case", case["id"], flush=True)
    direct_plan = gpu.Plan(case["dim"], "float64")
    direct_plan.set_p2p_mutual(False)
    direct_plan.set_Depth(case["depth"])
    direct, _ = gpu_solve(gpu, "fmm", x, w, x, case["p"], case["depth"], ts, case["k"], params, shared=True,
                          plan=direct_plan)
    mutual, pl = gpu_solve(gpu, "fmm", x, w, x, case["p"], case["depth"], ts, case["k"], params, shared=True)
    ref = oracle.fmm(x, w, x, case["p"], pl.depth, ts, KID[case["k"]], params)
    e_dir = rel_l2(mutual, direct)
    e_ref = rel_l2(mutual, ref)
    print("  mutual vs direct %.3e, vs oracle %.3e (direct vs oracle %.3e)" % (e_dir, e_ref, rel_l2(direct, ref)))
    assert e_dir <= 1e-13, e_dir
    assert e_ref <= 1e-12, e_ref
    log = pl.time_log()
    assert log["P2P sum"] >= 0.0


def test_mutual_leaf_level_one_adds_onto_ans(gpu, oracle):
    """Depth 2 (leaf level 1): no far field; P2P adds onto the caller's ans
    (src/fmm.cpp:114-139), in the mutual form too."""
    x = synth.points(51, 3000, 3)
    w = synth.uniform(52, 3000)
    ans0 = synth.normal(53, 3000)
    got, _ = gpu_solve(gpu, "fmm", x, w, x, 2, 2, 1, "inv_r_reg", shared=True, ans0=ans0)
    ref = oracle.fmm(x, w, x, 2, 2, 1, KID["inv_r_reg"], (), ans=ans0.copy())
    assert rel_l2(got, ref) <= 1e-12


def test_mutual_is_deterministic(gpu):
    x, w = _points(dict(dim=3, n=50000, w="normal", src="mixture"))
    a, pl = gpu_solve(gpu, "fmm", x, w, x, 3, 5, 1, "inv_r_reg", shared=True)
    b, _ = gpu_solve(gpu, "fmm", x, w, x, 3, 5, 1, "inv_r_reg", shared=True, plan=pl)
    assert np.array_equal(a, b)


def test_mutual_float32(gpu, oracle):
    x, w = _points(dict(dim=3, n=20000, w="u01"))
    got, pl = gpu_solve(gpu, "fmm", x, w, x, 3, 4, 1, "inv_r_reg", shared=True, dtype=np.float32)
    ref = oracle.fmm(x, w, x, 3, pl.depth, 1, KID["inv_r_reg"], ())
    assert rel_l2(got, ref) <= 3e-5


def test_separate_copies_use_direct_form(gpu, oracle):
    """The same points passed as two arrays take the direct form (no shared
    tree), which must agree with the mutual form on one array."""
    x, w = _points(dict(dim=3, n=20000, w="u01"))
    a, _ = gpu_solve(gpu, "fmm", x, w, x, 3, 4, 1, "inv_r_reg", shared=True)
    b, pl = gpu_solve(gpu, "fmm", x, w, x.copy(), 3, 4, 1, "inv_r_reg", shared=False)
    assert rel_l2(a, b) <= 1e-13
