"""Write golden vectors from the compiled reference (oracle/_ref, built by
`make -C oracle` from fockl/FMMA's own src/*.cpp where those sources exist).

    python tests/golden/make_golden.py

Each fixture stores the workload spec (inputs are regenerated from SplitMix64
seeds by fmma_amd.synth; a SHA-256 of the inputs is stored to detect any
generator change) and the reference's outputs.  The reference itself never
travels; these .npz files do.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fmma_amd import synth  # noqa: E402
from oracle import pyoracle  # noqa: E402

KID = {"inv_r": 0, "inv_r_reg": 1, "inv_diff": 2, "log_r": 3, "gauss": 4, "yukawa": 5, "sum_r2": 6, "sum2_r2": 7}

# (name, dim, n_t, n_s, shared, weights, poly_ord, depth, trans_sym, kernel, params, extra input options)
CASES = [
    ("c1_1d_invdiff", 1, 4096, 4096, False, "pm1", 7, -1, 0, "inv_diff", (), {}),
    ("c2_2d_log", 2, 8192, 8192, True, "normal", 5, 5, 1, "log_r", (), {}),
    ("c3_3d_invr_d4", 3, 8192, 8192, True, "u01", 3, 4, 1, "inv_r_reg", (), {}),
    ("c3_3d_invr_d5_ts0", 3, 6000, 6000, True, "u01", 3, 5, 0, "inv_r_reg", (), {}),
    ("c5_3d_yukawa", 3, 4096, 4096, True, "normal", 4, 4, 1, "yukawa", (2.0,), {}),
    ("t_2d_sum2_r2", 2, 2000, 1500, False, "u01", 5, 4, 0, "sum2_r2", (), {}),
    # points on leaf-box faces and on the far boundary of the cube
    ("b_3d_faces", 3, 6000, 5000, False, "u01", 3, 5, 1, "inv_r_reg", (), dict(faces=400)),
    ("b_1d_faces", 1, 3000, 3000, False, "pm1", 5, 8, 0, "log_r", (), dict(faces=300)),
    ("b_2d_faces_f32", 2, 5000, 4000, False, "normal", 4, 6, 1, "log_r", (), dict(faces=300, dtype="float32")),
    # domains away from the origin and of small scale, both M2L branches
    ("o_3d_offset1e4_ts0", 3, 4000, 4000, True, "u01", 3, 4, 0, "inv_r_reg", (), dict(offset=1e4)),
    ("o_3d_offset1e4_ts1", 3, 4000, 4000, True, "u01", 3, 4, 1, "inv_r_reg", (), dict(offset=1e4)),
    ("o_3d_offset1e5_ts0", 3, 3000, 3000, True, "u01", 3, 4, 0, "inv_r_reg", (), dict(offset=1e5)),
    ("o_3d_offsetm5e3_ts0", 3, 3000, 2500, False, "normal", 2, 4, 0, "yukawa", (2.0,),
     dict(offset=-5e3, scale=2.0)),
    ("o_2d_scaled1e-3_ts0", 2, 4000, 4000, True, "normal", 5, 5, 0, "log_r", (), dict(scale=1e-3)),
    ("o_2d_offset1e3_ts0", 2, 4000, 4000, True, "normal", 5, 5, 0, "log_r", (), dict(offset=1e3)),
]


def inputs(dim, nt, ns, shared, wkind, offset=0.0, scale=1.0, faces=0, depth=None, dtype="float64"):
    """The fixture's inputs from SplitMix64 seeds 11 (targets), 12 (sources)
    and 13 (weights).  faces > 0: that many points of each set moved onto
    the faces of the leaf boxes of `depth` (multiples of 2^-(depth-1)), one
    source at the far corner 1.0 and one target at 0.0, so that the cube is
    exactly [0, 1]^dim.  Then x -> x * scale + offset, and the dtype."""
    t = synth.points(11, nt, dim)
    s = t if shared else synth.points(12, ns, dim)
    if faces:
        NL = 1 << (depth - 1)
        t = t.copy()
        s = t if shared else s.copy()
        t[:faces] = np.floor(t[:faces] * NL) / NL
        t[0] = 0.0
        if not shared:
            s[:faces] = np.floor(s[:faces] * NL) / NL
        s[-1] = 1.0
    if scale != 1.0 or offset != 0.0:
        t = t * scale + offset
        s = t if shared else s * scale + offset
    n = s.shape[0]
    w = {"u01": synth.uniform(13, n), "pm1": 2 * synth.uniform(13, n) - 1, "normal": synth.normal(13, n)}[wkind]
    if dtype == "float32":
        t, w, s = (a.astype(np.float32) for a in (t, w, s))
    return t, w, s


def spec_inputs(spec):
    """inputs() for a stored fixture spec."""
    return inputs(spec["dim"], spec["nt"], spec["ns"], spec["shared"], spec["weights"], spec.get("offset", 0.0),
                  spec.get("scale", 1.0), spec.get("faces", 0), spec.get("depth"), spec.get("dtype", "float64"))


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def main():
    if not pyoracle.have_ref():
        raise SystemExit("oracle/_ref/libfmma_ref.so is not built (needs the reference sources)")
    manifest = {}
    for name, dim, nt, ns, shared, wk, p, depth, ts, k, prm, opt in CASES:
        d = depth if depth > 0 else pyoracle.auto_depth(nt, dim)
        spec = dict(name=name, dim=dim, nt=nt, ns=ns, shared=shared, weights=wk, poly_ord=p, depth=d,
                    trans_sym=ts, kernel=k, params=list(prm), **opt)
        t, w, s = spec_inputs(spec)
        spec["input_sha256"] = digest(t, w, s)
        fmm = pyoracle.ref_solve("fmm", t, w, s, p, depth, ts, KID[k], prm)
        exact = pyoracle.ref_solve("exact", t, w, s, p, depth, ts, KID[k], prm)
        # leaf boxes from the reference's own binning (P2M's source_ind_in_box)
        sboxes = pyoracle.ref_leaf_boxes(t, s, s, d)
        tboxes = pyoracle.ref_leaf_boxes(t, s, t, d)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, spec=json.dumps(spec), fmm=fmm, exact=exact, source_boxes=sboxes.astype(np.int32),
                            target_boxes=tboxes.astype(np.int32))
        manifest[name + ".npz"] = hashlib.sha256(open(path, "rb").read()).hexdigest()
        print(name, "depth", d, "rel(fmm,exact) = %.3e" % (np.linalg.norm(fmm - exact) / np.linalg.norm(exact)))
    # per-level coefficients of a tiny case (M2M and L2L run: depth 5)
    t, w, s = inputs(2, 700, 900, False, "u01")
    Wm, Wl = pyoracle.ref_levels(t, w, s, 3, 5, 1, KID["inv_r_reg"])
    np.savez_compressed(os.path.join(HERE, "levels_2d_d5.npz"),
                        spec=json.dumps(dict(dim=2, nt=700, ns=900, shared=False, weights="u01", poly_ord=3,
                                             depth=5, trans_sym=1, kernel="inv_r_reg",
                                             input_sha256=digest(t, w, s))),
                        **{"Wm%d" % l: Wm[l] for l in range(2, 5)}, **{"Wl%d" % l: Wl[l] for l in range(2, 5)})
    manifest["levels_2d_d5.npz"] = hashlib.sha256(open(os.path.join(HERE, "levels_2d_d5.npz"), "rb").read()).hexdigest()
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
