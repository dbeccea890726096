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

# (name, dim, n_t, n_s, shared, weights, poly_ord, depth, trans_sym, kernel, params)
CASES = [
    ("c1_1d_invdiff", 1, 4096, 4096, False, "pm1", 7, -1, 0, "inv_diff", ()),
    ("c2_2d_log", 2, 8192, 8192, True, "normal", 5, 5, 1, "log_r", ()),
    ("c3_3d_invr_d4", 3, 8192, 8192, True, "u01", 3, 4, 1, "inv_r_reg", ()),
    ("c3_3d_invr_d5_ts0", 3, 6000, 6000, True, "u01", 3, 5, 0, "inv_r_reg", ()),
    ("c5_3d_yukawa", 3, 4096, 4096, True, "normal", 4, 4, 1, "yukawa", (2.0,)),
    ("t_2d_sum2_r2", 2, 2000, 1500, False, "u01", 5, 4, 0, "sum2_r2", ()),
]


def inputs(dim, nt, ns, shared, wkind):
    t = synth.points(11, nt, dim)
    s = t if shared else synth.points(12, ns, dim)
    n = s.shape[0]
    w = {"u01": synth.uniform(13, n), "pm1": 2 * synth.uniform(13, n) - 1, "normal": synth.normal(13, n)}[wkind]
    return t, w, s


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def main():
    if not pyoracle.have_ref():
        raise SystemExit("oracle/_ref/libfmma_ref.so is not built (needs the reference sources)")
    manifest = {}
    for name, dim, nt, ns, shared, wk, p, depth, ts, k, prm in CASES:
        t, w, s = inputs(dim, nt, ns, shared, wk)
        fmm = pyoracle.ref_solve("fmm", t, w, s, p, depth, ts, KID[k], prm)
        exact = pyoracle.ref_solve("exact", t, w, s, p, depth, ts, KID[k], prm)
        d = depth if depth > 0 else pyoracle.auto_depth(nt, dim)
        boxes = pyoracle.leaf_boxes(t, s, s, d)
        spec = dict(name=name, dim=dim, nt=nt, ns=ns, shared=shared, weights=wk, poly_ord=p, depth=d,
                    trans_sym=ts, kernel=k, params=list(prm), input_sha256=digest(t, w, s))
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, spec=json.dumps(spec), fmm=fmm, exact=exact, source_boxes=boxes.astype(np.int32))
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
