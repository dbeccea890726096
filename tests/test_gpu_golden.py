"""GPU results against the golden vectors the compiled reference wrote
(tests/golden/*.npz): fmm at 1e-12, exact at 1e-13, leaf boxes bit-exact
(the reference's own P2M lists, with points on box faces and the far
boundary), per-level coefficients at 1e-12; FP32 fixtures at 3e-5 with
boxes bit-exact; domains away from the origin with both M2L branches."""
import glob
import json
import os

import numpy as np
import pytest

from tests.conftest import ROOT, rel_l2
from tests.test_gpu_parity import gpu_solve

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(ROOT, "tests", "golden")
FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "[cbot]_*.npz")) + glob.glob(os.path.join(GOLDEN, "c[0-9]_*.npz")))
FIXTURES = sorted(set(FIXTURES))


def _load(path):
    from tests.golden.make_golden import digest, spec_inputs
    z = np.load(path)
    spec = json.loads(str(z["spec"]))
    t, w, s = spec_inputs(spec)
    assert digest(t, w, s) == spec["input_sha256"]
    return z, spec, t, w, s


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_gpu_fmm_matches_reference_golden(gpu, path):
    z, spec, t, w, s = _load(path)
    print("golden", spec["name"], flush=True)
    f32 = spec.get("dtype") == "float32"
    dt = np.float32 if f32 else np.float64
    got, pl = gpu_solve(gpu, "fmm", t, w, s, spec["poly_ord"], spec["depth"], spec["trans_sym"], spec["kernel"],
                        tuple(spec["params"]), shared=spec["shared"], dtype=dt)
    err = rel_l2(got, z["fmm"])
    print("  fmm vs reference %.3e" % err, flush=True)
    assert err <= (3e-5 if f32 else 1e-12)
    assert np.array_equal(pl.point_boxes("s", s.shape[0]), z["source_boxes"])
    if not spec["shared"]:
        assert np.array_equal(pl.point_boxes("t", t.shape[0]), z["target_boxes"])
    ex, _ = gpu_solve(gpu, "exact", t, w, s, kernel=spec["kernel"], params=tuple(spec["params"]), dtype=dt)
    assert rel_l2(ex, z["exact"]) <= (3e-5 if f32 else 1e-13)


def test_gpu_levels_match_reference_golden(gpu):
    z, spec, t, w, s = _load(os.path.join(GOLDEN, "levels_2d_d5.npz"))
    _, pl = gpu_solve(gpu, "fmm", t, w, s, spec["poly_ord"], spec["depth"], spec["trans_sym"], spec["kernel"])
    for l in range(2, spec["depth"]):
        assert rel_l2(pl.level_coeffs("Wm", l), z["Wm%d" % l]) <= 1e-12
        assert rel_l2(pl.level_coeffs("Wl", l), z["Wl%d" % l]) <= 1e-12
