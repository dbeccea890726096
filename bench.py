"""Benchmark: 3-D 1/r FMM solve, N = 4,194,304 points, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-exact]

A step is one full FMM solve (tree build, P2M, M2M, M2L, L2L, L2P, P2P) of
the same synthetic workload: N points uniform in the unit cube (SplitMix64
seed 1), targets = sources, weights U[0,1) (seed 2), f = 1/|x-y| with f = 0 at
r = 0, 4 Chebyshev nodes per dimension (poly_ord 3), Depth 6 (the auto rule),
trans_sym_flag 1.  Inputs are resident in HBM before the timed region.

The headline `value` is ms per solve of the FP32-MFMA M2L form (P2P, geometry
and operator set-up in FP64).  The same line reports the plain FP64 GPU form,
the split-BF16x3 form, both error relations (precision error against the
FP64 GPU fmm over all N targets; total error against the FP64 exact sum on
8,192 sampled targets, computed on the GPU), per-stage times, the roofline of
the dominant kernel (live HIP-event stage times), and the CPU baseline: the
compiled reference (oracle/_ref) or its C restatement (oracle/) timed on this
host on a bounded sample and scaled to N.

Beside the headline: the same workload at the Chebyshev orders whose error
against the exact sum is <= 1e-6 (5 and 6 nodes per dimension, Depth 6) in
every form (`at_rel_err_1e-6`), the batching form (512 Yukawa problems), and
the tree and nrnmm solvers at N = 1,048,576.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D N=4M FMM solve ms @ rel-err≤1e-6: FP32-MFMA M2L vs FP64 GPU vs CPU"
PEAK_HBM = 8.0e12          # B/s
PEAK_FP64 = 78.6e12        # FLOP/s, FP64 vector and matrix
PEAK_FP32 = 157.3e12       # FLOP/s, FP32 matrix
PEAK_BF16 = 2.5e15         # FLOP/s, dense bf16 matrix
P2P_FLOPS_PER_PAIR = 12    # 3 sub, 3 mul, 2 add (r^2), sqrt, div (1/r), mul (w f), add (accumulate)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 22)
    ap.add_argument("--depth", type=int, default=-1)
    ap.add_argument("--poly-ord", type=int, default=3, help="Chebyshev order (nodes per dimension - 1)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-exact", action="store_true")
    ap.add_argument("--forms", default="fp32,native,bf16x3")
    ap.add_argument("--no-batch", action="store_true", help="skip the batching-form measurement")
    ap.add_argument("--batch-only", action="store_true", help="only the batching-form measurement")
    ap.add_argument("--batch-problems", type=int, default=512)
    ap.add_argument("--no-solvers", action="store_true", help="skip the tree and nrnmm measurement")
    ap.add_argument("--accuracy-configs", default="4:6,5:6",
                    help="poly_ord:Depth pairs also timed in every form with their errors vs the exact sum "
                         "(the orders that reach rel-err <= 1e-6); empty to skip")
    ap.add_argument("--solvers-only", action="store_true", help="only the tree and nrnmm measurement")
    ap.add_argument("--solvers-n", type=int, default=1 << 20)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (3 times): the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rms_rel(a, b):
    return float(np.sqrt(np.mean(((a - b) / b) ** 2)))


def pair_count(x_np, depth):
    """Exact P2P pair interactions of the uniform tree (targets = sources):
    sum over leaf boxes of n_b * (sources in the clipped 3^3 neighbourhood),
    and the pairs within one box, sum of n_b^2."""
    NL = 1 << (depth - 1)
    lo, hi = x_np.min(0), x_np.max(0)
    Len = float(np.max(hi - lo))
    org = (hi + lo) / 2 - Len / 2
    b = np.minimum(((x_np - org) / (Len / NL)).astype(np.int64), NL - 1)
    cnt = np.zeros((NL, NL, NL), dtype=np.int64)
    np.add.at(cnt, (b[:, 0], b[:, 1], b[:, 2]), 1)
    pad = np.pad(cnt, 1)
    nb = np.zeros_like(cnt)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                nb += pad[1 + dx:1 + dx + NL, 1 + dy:1 + dy + NL, 1 + dz:1 + dz + NL]
    return int((cnt * nb).sum()), int((cnt * cnt).sum())


def m2l_pairs(depth, dim=3):
    return sum((6 * 2 ** l - 8) ** dim - (3 * 2 ** l - 2) ** dim for l in range(2, depth))


def measured_traffic(kernel, n, depth):
    """HBM-side bytes per dispatch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled
    by the plain-copy calibration) from the committed rocprofv3 --pmc summary
    of this workload, or None when the summary does not cover this run."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_3d_n4m_depth6.md")
    if n != 1 << 22 or depth != 6 or not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip().strip("`") for c in line.split("|")]
        if len(cells) > 5 and kernel in cells[1]:
            return int(cells[4]) + int(cells[5])
    return None


def cpu_baseline():
    """Time the reference fmm (compiled oracle/_ref if present, else the C
    restatement) on N = 524,288 points (Depth 5, 128 points per leaf as at
    N = 4M, Depth 6), trans_sym_flag 1, and scale by 8 to N = 4,194,304."""
    from oracle import pyoracle
    from fmma_amd import synth
    n = 1 << 19
    x = synth.points(1, n, 3)
    w = synth.uniform(2, n)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 16))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    times = []
    if pyoracle.have_ref():
        kind = "reference"
        pyoracle.ref_set_threads(cores)
        for _ in range(2):
            t0 = time.perf_counter()
            pyoracle.ref_solve("fmm", x, w, x, 3, -1, 1, 1)
            times.append(time.perf_counter() - t0)
    else:
        kind = "port"
        for _ in range(2):
            t0 = time.perf_counter()
            pyoracle.fmm(x, w, x, 3, -1, 1, 1)
            times.append(time.perf_counter() - t0)
    ms = min(times) * 1e3
    cpu_model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                cpu_model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return {"value": round(ms * 8, 1), "unit": "ms (extrapolated x8 to N=4,194,304)", "cores": cores, "kind": kind,
            "sample": "fmm solve, N=524,288 uniform points (Depth 5, 128 points/leaf), 4 nodes/dim, 1/r, "
                      "trans_sym_flag 1; best of 2 runs = %.1f ms; scaled x8 (O(N))" % ms,
            "cpu": cpu_model, "omp_threads": cores}


def batch_bench(fmma_amd, synth, torch, nprob, steps):
    """The batching form: nprob independent 3-D Yukawa (kappa = 2) problems,
    N log-uniform in [1024, 65536], points uniform in each problem's unit
    cube, weights N(0,1), 5 nodes per dimension, FP32-MFMA M2L.  One batched
    solve (replayed from its captured hipGraph, M2L tiles from the
    autotuner) against a loop of per-problem solves on the same data (one
    plan per Depth, buffers warmed)."""
    c = synth.config("c5_batch", nprob)
    counts = c["counts"]
    x = torch.from_numpy(c["targets"]).cuda()
    w = torch.from_numpy(c["weights"]).cuda()
    n = x.shape[0]

    def mkplan(depth=-1):
        pl = fmma_amd.Plan(3, "float64")
        pl.set_solver_type("fmm")
        pl.set_poly_ord(c["poly_ord"])
        pl.set_Depth(depth)
        pl.set_trans_sym_flag(1)
        pl.set_kernel(c["kernel"], *c["params"])
        pl.set_m2l_precision("fp32")
        return pl

    pl = mkplan()
    ans = torch.zeros(n, dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    pl.solve_batch(counts, x, w, counts, x, ans)  # eager + autotuner
    torch.cuda.synchronize()
    first_ms = (time.perf_counter() - t0) * 1e3
    pl.solve_batch(counts, x, w, counts, x, ans)  # capture + launch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pl.solve_batch(counts, x, w, counts, x, ans)
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / steps
    print("# batch of %d problems (%d points): first call %.1f ms, graph replay %.3f ms, tiles %s"
          % (nprob, n, first_ms, graph_ms, pl.m2l_tiles()), file=sys.stderr, flush=True)

    depths = [int(np.log(k) / 3) + 1 for k in counts]
    plans = {d: mkplan(d) for d in sorted(set(depths))}
    ref = torch.zeros_like(ans)
    off = np.concatenate([[0], np.cumsum(counts)])

    def loop():
        for p in range(nprob):
            sl = slice(int(off[p]), int(off[p + 1]))
            plans[depths[p]].solve(x[sl], w[sl], x[sl], ref[sl])
        torch.cuda.synchronize()

    loop()  # warm the buffers
    t0 = time.perf_counter()
    loop()
    loop_ms = (time.perf_counter() - t0) * 1e3
    diff = float((ans - ref).abs().max().item())
    print("# per-problem loop %.1f ms; max |batch - loop| %.3e" % (loop_ms, diff), file=sys.stderr, flush=True)
    return {"workload": "%d Yukawa problems, N log-uniform in [1024, 65536], %d points, 5 nodes/dim, M2L fp32, "
                        "targets = sources" % (nprob, n),
            "problems": nprob, "points": int(n), "depths": {str(d): depths.count(d) for d in sorted(set(depths))},
            "graph_replay_ms": round(graph_ms, 3), "first_call_ms": round(first_ms, 1),
            "per_problem_loop_ms": round(loop_ms, 1), "speedup_vs_loop": round(loop_ms / graph_ms, 1),
            "max_abs_diff_vs_loop": diff, "m2l_tiles": {str(k): v for k, v in pl.m2l_tiles().items()}}


def _clip_window(b, N, w):
    """Boxes of a window [2(b//2) - 2, 2(b//2) + 3] (w = 6) or [b - 1, b + 1]
    (w = 3) that lie in [0, N), per coordinate."""
    if w == 6:
        lo = 2 * (b // 2) - 2
        return np.minimum(lo + 6, N) - np.maximum(lo, 0)
    return np.minimum(b + 2, N) - np.maximum(b - 1, 0)


def m2p_evaluations(x_np, P, depth=None, grid=None):
    """Kernel evaluations of the M2P stage (targets x far-box nodes): tree --
    every level 2..Depth-1, the interaction list of the target's box; nrnmm --
    every box of the grid at Chebyshev distance >= 2."""
    lo, hi = x_np.min(0), x_np.max(0)
    Len = float(np.max(hi - lo))
    org = (hi + lo) / 2 - Len / 2
    total = 0
    if grid is not None:
        b = np.minimum(((x_np - org) / (Len / grid)).astype(np.int64), grid - 1)
        near = np.prod(_clip_window(b, grid, 3), axis=1)
        return int(((grid ** 3 - near) * P).sum())
    L = depth - 1
    bl = np.minimum(((x_np - org) / (Len / (1 << L))).astype(np.int64), (1 << L) - 1)
    for l in range(2, L + 1):
        b = bl >> (L - l)
        N = 1 << l
        cnt = np.prod(_clip_window(b, N, 6), axis=1) - np.prod(_clip_window(b, N, 3), axis=1)
        total += int(cnt.sum()) * P
    return total


def solver_bench(fmma_amd, synth, torch, n, steps, with_cpu):
    """The tree and nrnmm solvers (src/tree.cpp, src/nrnmm.cpp) on N points
    uniform in the unit cube, targets = sources, f = 1/r (0 at r = 0), 4 nodes
    per dimension, auto Depth / nrn_N.  GPU: best of `steps` solves with the
    stage times; roofline of the M2P stage against FP64 peak.  CPU: the
    compiled reference (or the C restatement) with two target samples of the
    points (tree: 8192 and 65536, nrnmm: 2048 and 4096) and all N sources,
    Depth / nrn_N fixed to the GPU's: the per-target cost (from the
    difference of the two) times N plus the rest."""
    x_np = synth.points(1, n, 3)
    w_np = synth.uniform(2, n)
    x = torch.from_numpy(x_np).cuda()
    w = torch.from_numpy(w_np).cuda()
    out = {}
    for solver in ("tree", "nrnmm"):
        pl = fmma_amd.Plan(3, "float64")
        pl.set_solver_type(solver)
        pl.set_poly_ord(3)
        pl.set_kernel("inv_r_reg")
        a = torch.zeros(n, dtype=torch.float64, device="cuda")
        pl.solve(x, w, x, a)  # warm-up; fixes the auto Depth / nrn_N
        torch.cuda.synchronize()
        best, log = None, None
        for _ in range(steps):
            pl.clear_time_log()
            t0 = time.perf_counter()
            pl.solve(x, w, x, a)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            if best is None or ms < best:
                best, log = ms, {k: round(v, 3) for k, v in pl.time_log().items() if not k.startswith("total")}
        size = pl.depth if solver == "tree" else pl.nrn_N
        ev = m2p_evaluations(x_np, 64, depth=size) if solver == "tree" else m2p_evaluations(x_np, 64, grid=size)
        m2p_s = log["M2P"] * 1e-3
        flops = ev * P2P_FLOPS_PER_PAIR
        rec = {"ms": round(best, 3), "stages_ms": log, ("depth" if solver == "tree" else "nrn_N"): size,
               "m2p_evaluations": ev,
               "m2p_roofline": {"kernel": "m2p_kernel", "bound": "fp64-valu",
                                "achieved": float("%.4g" % (flops / m2p_s / 1e12)), "peak": PEAK_FP64 / 1e12,
                                "unit": "TFLOP/s", "frac": float("%.4g" % (flops / m2p_s / PEAK_FP64)),
                                "work": "%d evaluations x %d flops" % (ev, P2P_FLOPS_PER_PAIR)}}
        print("# %s (%s %d): %.2f ms, stages %s, M2P %.3g evaluations/s" % (
            solver, "Depth" if solver == "tree" else "nrn_N", size, best, json.dumps(log), ev / m2p_s),
            file=sys.stderr, flush=True)
        if with_cpu:
            rec["cpu_baseline"] = solver_cpu(solver, x_np, w_np, size)
            rec["speedup_vs_cpu"] = round(rec["cpu_baseline"]["value"] / best, 1)
            print("# %s CPU %s" % (solver, json.dumps(rec["cpu_baseline"])), file=sys.stderr, flush=True)
        out[solver] = rec
    out["workload"] = ("3-D 1/r (0 at r=0), N=%d uniform points in the unit cube, targets = sources, weights U[0,1), "
                       "4 Chebyshev nodes/dim, auto Depth / nrn_N" % n)
    return out


def solver_cpu(solver, x_np, w_np, size):
    from oracle import pyoracle
    n = x_np.shape[0]
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 16))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    ref = pyoracle.have_ref()
    if ref:
        pyoracle.ref_set_threads(cores)
    rng = np.random.default_rng(5)
    times = {}
    sizes = (8192, 65536) if solver == "tree" else (2048, 4096)
    for s in sizes:
        t = np.ascontiguousarray(x_np[rng.choice(n, s, replace=False)])
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            if ref:
                if solver == "tree":
                    pyoracle.ref_solve("tree", t, w_np, x_np, 3, size, 0, 1)
                else:
                    pyoracle.ref_solve("nrnmm", t, w_np, x_np, 3, -1, 0, 1, nrn_N=size)
            elif solver == "tree":
                pyoracle.tree(t, w_np, x_np, 3, size, 1)
            else:
                pyoracle.nrnmm(t, w_np, x_np, 3, size, 1)
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        times[s] = best
    s0, s1 = sizes
    per_target = max(times[s1] - times[s0], 0.0) / (s1 - s0)
    rest = max(times[s0] - per_target * s0, 0.0)
    ms = (rest + per_target * n) * 1e3
    return {"value": round(ms, 1), "unit": "ms (extrapolated to N=%d targets)" % n, "cores": cores,
            "kind": "reference" if ref else "port",
            "sample": "%s solve with %d and %d of the points as targets, all %d as sources: %.3f s and %.3f s; "
                      "per-target cost x N + the rest" % (solver, s0, s1, n, times[s0], times[s1])}


def run_forms(fmma_amd, torch, x, w, poly_ord, depth_in, forms, steps, warmup, exact):
    """The 3-D 1/r fmm solve (trans_sym 1, 0 at r = 0) in each M2L form: ms per
    solve over `steps` after `warmup` (the first runs the tile autotuner),
    HIP-event stage times, the precision error of each form against the FP64
    form over all targets and, with exact = (indices, sums), the total error
    against the exact sum on those targets."""
    n = x.shape[0]
    plans, answers, results = {}, {}, {}
    for form in forms:
        pl = fmma_amd.Plan(3, "float64")
        pl.set_solver_type("fmm")
        pl.set_poly_ord(poly_ord)
        pl.set_Depth(depth_in)
        pl.set_trans_sym_flag(1)
        pl.set_kernel("inv_r_reg")
        pl.set_m2l_precision(form)
        plans[form] = pl
        answers[form] = torch.zeros(n, dtype=torch.float64, device=x.device)
    for form in forms:
        pl, a = plans[form], answers[form]
        for _ in range(warmup):
            pl.solve(x, w, x, a)
        torch.cuda.synchronize()
        pl.clear_time_log()
        t0 = time.perf_counter()
        for _ in range(steps):
            pl.solve(x, w, x, a)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        log = {k: v / steps for k, v in pl.time_log().items() if not k.startswith("total")}
        results[form] = {"ms": round(ms, 3), "stages_ms": {k: round(v, 4) for k, v in log.items()},
                         "m2l_tiles": {str(k): v for k, v in pl.m2l_tiles(batch=False).items()}}
        print("# p=%d depth %d form %-7s %.3f ms/solve  stages %s" % (
            poly_ord, pl.depth, form, ms, json.dumps(results[form]["stages_ms"])), file=sys.stderr, flush=True)
    depth = plans[forms[0]].depth
    ref_form = "native" if "native" in forms else forms[0]
    a_ref = answers[ref_form].cpu().numpy()
    for form in forms:
        if form != ref_form:
            results[form]["precision_err_vs_fp64_fmm"] = rel_l2(answers[form].cpu().numpy(), a_ref)
    if exact is not None:
        idx, u = exact
        for form in forms:
            got = answers[form].cpu().numpy()[idx]
            results[form]["total_err_vs_exact"] = rel_l2(got, u)
            results[form]["rms_rel_err_vs_exact"] = rms_rel(got, u)
    return results, depth


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this FMM build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import fmma_amd
    from fmma_amd import synth

    dev = torch.device("cuda:0")
    if args.copy_mib:
        a = torch.ones(args.copy_mib << 18, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4}))
        return
    if args.batch_only:
        print(json.dumps({"batching": batch_bench(fmma_amd, synth, torch, args.batch_problems, args.steps)}))
        return
    if args.solvers_only:
        print(json.dumps({"solvers": solver_bench(fmma_amd, synth, torch, args.solvers_n, max(3, args.steps // 2),
                                                  not args.no_cpu)}))
        return
    n = args.n
    x_np = synth.points(1, n, 3)
    w_np = synth.uniform(2, n)
    x = torch.from_numpy(x_np).to(dev)
    w = torch.from_numpy(w_np).to(dev)
    torch.cuda.synchronize()

    forms = [f for f in args.forms.split(",") if f]
    exact = None
    if not args.no_exact:
        # the FP64 exact sum on 8192 sampled targets (GPU exact solver), for the total error
        idx = synth.sample_indices(3, n, 8192)
        ts = x[torch.from_numpy(idx).to(dev)].contiguous()
        ex = fmma_amd.Plan(3, "float64")
        ex.set_solver_type("exact")
        ex.set_kernel("inv_r_reg")
        u = torch.zeros(len(idx), dtype=torch.float64, device=dev)
        ex.solve(ts, w, x, u)
        exact = (idx, u.cpu().numpy())
    results, depth = run_forms(fmma_amd, torch, x, w, args.poly_ord, args.depth, forms, args.steps, args.warmup,
                               exact)

    accuracy = None
    if exact is not None and args.accuracy_configs:
        accuracy = {}
        for cfg in args.accuracy_configs.split(","):
            p, d = (int(v) for v in cfg.split(":"))
            res, _ = run_forms(fmma_amd, torch, x, w, p, d, forms, 5, 2, exact)
            entry = {"nodes_per_dim": p + 1, "depth": d, "forms": res}
            if "native" in res:
                for f in res:
                    if f != "native":
                        entry["speedup_%s_vs_fp64_gpu" % f] = round(res["native"]["ms"] / res[f]["ms"], 3)
            entry["all_forms_within_1e-6"] = all(r["total_err_vs_exact"] <= 1e-6 for r in res.values())
            accuracy["p%d_depth%d" % (p, d)] = entry

    head = "fp32" if "fp32" in forms else forms[0]
    ms_head = results[head]["ms"]
    stages = results[head]["stages_ms"]
    # roofline of the dominant kernel
    pairs, self_pairs = pair_count(x_np, depth)
    # the mutual form evaluates each pair of distinct neighbour boxes once
    evaluated = self_pairs + (pairs - self_pairs) // 2
    mpairs = m2l_pairs(depth)
    dom = max(stages, key=stages.get)
    P = (args.poly_ord + 1) ** 3
    if dom == "P2P":
        flops = pairs * P2P_FLOPS_PER_PAIR
        ach = flops / (stages["P2P"] * 1e-3)
        roof = {"kernel": "p2p_mutual_kernel", "bound": "fp64-valu", "achieved": ach / 1e12,
                "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": ach / PEAK_FP64,
                "traffic": measured_traffic("p2p_mutual_kernel", n, depth),
                "work": "%d pair interactions x %d flops" % (pairs, P2P_FLOPS_PER_PAIR),
                "pairs_per_s": pairs / (stages["P2P"] * 1e-3),
                "evaluated_pairs": evaluated,
                "evaluations_per_s": evaluated / (stages["P2P"] * 1e-3)}
    elif dom == "M2L":
        flops = 2 * P * P * mpairs
        ach = flops / (stages["M2L"] * 1e-3)
        peak = {"fp32": PEAK_FP32, "native": PEAK_FP64, "bf16x3": PEAK_BF16 / 3}[head]
        roof = {"kernel": "m2l_pipe", "bound": "mfma", "achieved": ach / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s", "frac": ach / peak, "traffic": measured_traffic("m2l_pipe", n, depth),
                "work": "2*P^2 x %d M2L pairs" % mpairs}
    else:
        nbytes = n * 4 * 8 + (8 ** (depth - 1)) * P * 8
        ach = nbytes / (stages[dom] * 1e-3)
        roof = {"kernel": dom, "bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": ach / PEAK_HBM, "traffic": None}
    for k in ("achieved", "peak", "frac"):
        roof[k] = float("%.4g" % roof[k])

    batching = None
    if not args.no_batch:
        batching = batch_bench(fmma_amd, synth, torch, args.batch_problems, max(3, args.steps // 2))

    solvers = None
    if not args.no_solvers:
        solvers = solver_bench(fmma_amd, synth, torch, args.solvers_n, 3, not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()

    line = {
        "metric": METRIC,
        "value": ms_head,
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_head,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (M2L %s)" % head,
        "data": "synthetic",
        "config": {"workload": "3-D 1/r FMM solve, N=%d uniform points in the unit cube, targets=sources, "
                               "%d Chebyshev nodes/dim, Depth %d, trans_sym 1, M2L form %s"
                               % (n, args.poly_ord + 1, depth, head),
                   "n": n, "poly_ord": args.poly_ord, "depth": depth, "kernel": "1/|x-y| (0 at r=0)", "m2l": head,
                   "parallelism": "single GPU"},
        "forms": results,
        "speedup_fp32_vs_fp64_gpu": round(results["native"]["ms"] / ms_head, 3) if "native" in results else None,
        "p2p_pairs": pairs,
        "m2l_pairs": mpairs,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batching": batching,
        "solvers": solvers,
        "at_rel_err_1e-6": accuracy,
    }
    if cpu and "native" in results:
        line["speedup_fp64_gpu_vs_cpu"] = round(cpu["value"] / results["native"]["ms"], 1)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
