#!/usr/bin/env python3
"""Benchmark of the treecode Gauss operator on one MI355X.

Workload (the 1M-point configuration the headline metric is quoted on):
1,000,000 points on a union of 6 seeded random ellipsoids (semi-axes
U[0.2,0.6], centres U[-0.5,0.5]^3, noise sigma = 5e-4), normalised as
CUDA-based/main_CAGR.py:42-45 does, float32; leaf size 32, theta = 0.5,
smoothing width 0.002 for every point (the l0 preset's minimum).

One step = one matvec pair: A.mu (A21, A22, A23 on mu [N,3], one fused
upward pass + traversal) followed by A^T.r (A2T1, A2T2, A2T3 on the three
columns of r = -A.mu, summed; one fused pass with three sets of centres), with
the inputs resident in HBM.  Upward passes are inside the step; the tree
build and the balancing of the walks (cut into pieces costed from exact
interaction counts) are outside (build_ms, balance_ms).  The timed steps run
without instrumentation; the per-launch HIP events of the roofline come from
the same steps run once more.  The main figure is the
FP32-compensated form; the plain FP64 operator on the same workload is timed
beside it (fp64_value), and the oracle (the FP64 CPU restatement) is timed on
a bounded sample of queries on the host cores (cpu_baseline).

Solve comparison (solve): CGNR on the normal equations of the linear operator
A mu = -[A21; A22; A23] mu (unit point weights, the same width), b = 0.5,
to ||A^T(b - A mu)|| / ||A^T b|| <= --solve-tol (default 1e-6, the metric's
tolerance: the operator's condition number, near 1e8, puts it about 10^4
CG iterations away, some 2.5 minutes of FP64; both solves restart CG from
the FP64 residual when it stalls), once with the plain FP64
operator and exact adjoint, once with FP32c iterations and FP64-verified
residual replacement (its adjoint through an adjoint plan, whose build is
timed inside the solve; also reported without the plan); all start from
mu = 0 on the tree built above.

Seeds (seed_sweep): the same workload on seeds 1..--seeds-1 (new clouds and
attributes; same setup and steps, solves to --sweep-solve-tol), with the
medians over all seeds.
``value`` and the roofline stay those of the seed-0 run (--seed).

Prints one JSON line.

The other configurations of the metric's table run on request, each printing
one JSON line of its own (they are parity cases of the test suite, measured
here at full size): --config c1 (10k sphere: the reference's 40-iteration
schedule on the FP64 oracle on the host cores and on the GPU, 128^3 grid
query), c2 (100k torus with 5 % outliers: FP64 drop-in matvec pairs, 256^3
grid query), c4 (500k displaced sphere: autograd forward + backward, FP64
finite-difference check on 64 entries), c5 (64 clouds, ~3M points: autotuned
batched persistent launch against one launch per cloud), pgr (GaussRecon on
the 1M-point ellipsoid cloud: the PGR octree query set built on the GPU
against the reference's own Octree::setTree on a host core, the A3 winding
numbers at its corners and samples, and marching cubes on the octree against
the reference's own GetMCIsoTriangles).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Gauss-operator matvecs/s (A+Aᵀ) at 1M pts; solve time vs plain FP64 at 1e-6"
PEAK_FP32_VECTOR_TFLOPS = 157.3     # MI355X published FP32 vector peak
F_OP = {"A2k": 25, "A2Tk": 27}      # algorithmic flops per interaction (sqrt, div, exp = 1)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--leaf", type=int, default=32)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--width", type=float, default=0.002)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU time of the oracle sample")
    ap.add_argument("--solve-tol", type=float, default=1e-6,
                    help="FP64-verified relative residual of the solve comparison")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--solve-fp64-plain", action="store_true",
                    help="also time the FP64 solve as one CG run without restarts")
    ap.add_argument("--solve-max-iter", type=int, default=40000,
                    help="iteration cap of the solve comparison (both solves)")
    ap.add_argument("--sweep-solve-tol", type=float, default=1e-4,
                    help="tolerance of the solves of the other seeds (1e-5 takes ~20 s a seed)")
    ap.add_argument("--seeds", type=int, default=5,
                    help="1M config: also time seeds 1..SEEDS-1 (pairs/s, FP64 ratio, solves) "
                         "and report the medians over seeds 0..SEEDS-1")
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5", "pgr"],
                    help="c3 (default): the headline workload; others: see the module docstring")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times and exit (calibrates "
                         "the HBM byte counters FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


TRAFFIC_FILE = "profiles/pmc_traffic_1M_ellipsoids.md"


def measured_traffic(kernel, args):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated), for the default workload only."""
    if (args.points, args.leaf, args.theta, args.width, args.seed) != (1_000_000, 32, 0.5, 0.002, 0):
        return None
    path = os.path.join(ROOT, TRAFFIC_FILE)
    try:
        text = open(path).read()
        blob = text[text.index("```json") + 7:]
        data = json.loads(blob[:blob.index("```")])
        return data["traffic_bytes"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    args = parse()
    # the oracle's OpenMP threads stay on their cores (set before any
    # OpenMP runtime starts)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no ROCm GPU visible"}))
        sys.exit(1)
    if args.copy_probe:
        src = torch.empty(args.copy_probe << 18, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe << 20}))
        return
    from cagr_amd import WindingNumberTreecode, synth

    dev = torch.device("cuda:0")
    if args.config != "c3":
        line = {"c1": config_c1, "c2": config_c2, "c4": config_c4, "c5": config_c5,
                "pgr": config_pgr}[args.config](args)
        print(json.dumps(line), flush=True)
        return
    n = args.points
    log(f"generating {n} points")
    pts = synth.normalize(synth.ellipsoids(n, args.seed)).astype(np.float32)
    mu_np = synth.random_normals(n, args.seed + 1).astype(np.float32)
    pts_t = torch.from_numpy(pts).to(dev)
    mu = torch.from_numpy(mu_np).to(dev)
    w = torch.full((n,), args.width, dtype=torch.float32, device=dev)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    wt = WindingNumberTreecode(pts_t, leaf_size=args.leaf, theta=args.theta, precision="fp32c")
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    info = wt.info()
    log(f"tree built in {build_ms:.1f} ms: {info}")
    # the walks cut into pieces of about a quarter of the mean wave cost
    # (WindingNumberTreecode.balance, costed on A2 with mu; setup, like the
    # build; outside the timed steps)
    t0 = time.perf_counter()
    wt.balance("A2", mu)
    torch.cuda.synchronize()
    balance_ms = (time.perf_counter() - t0) * 1e3
    log(f"walks balanced in {balance_ms:.1f} ms")

    def step(prec):
        a = wt.forward_A2(mu, w, precision=prec)
        return wt.forward_A2T_sum(-a, w, precision=prec), a

    def timed(prec, steps, warmup, profile=False):
        for _ in range(warmup):
            step(prec)
        torch.cuda.synchronize()
        if profile:
            wt.profile(True)
        t = time.perf_counter()
        for _ in range(steps):
            out = step(prec)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        prof = None
        if profile:
            wt.profile(False)
            prof = wt.profile_read()
        return el, out, prof

    # ---------------------------------------------------------------- FP32c
    log(f"fp32c: {args.warmup} warmup + {args.steps} timed steps")
    # the timed steps carry no instrumentation; the per-launch HIP events of
    # the roofline come from the same number of steps run once more
    el, (r32, a32), _ = timed("fp32c", args.steps, args.warmup)
    _, _, prof = timed("fp32c", args.steps, 0, profile=True)
    value = args.steps / el
    ms_per_step = el / args.steps * 1e3
    trav = np.array(prof["traverse"]).reshape(args.steps, 2)   # [A2, A2T_sum] per step
    upw = np.array(prof["upward"]).reshape(args.steps, 2)
    t_a2, t_a2t = trav[:, 0].mean(), trav[:, 1].mean()
    log(f"fp32c {value:.2f} pairs/s, {ms_per_step:.3f} ms/step; traversal A2 {t_a2:.3f} ms, "
        f"A2T {t_a2t:.3f} ms; upward {upw[:, 0].mean():.3f} / {upw[:, 1].mean():.3f} ms")

    # exact interaction counts of both traversals (same decisions in every precision)
    far_a, near_a, _ = wt.counts("A2", mu)
    s3 = (-a32).contiguous()
    far_t, near_t, _ = wt.counts("A2T_sum", s3)
    inter_a = int(far_a.sum().item() + near_a.sum().item())
    inter_t = int(far_t.sum().item() + near_t.sum().item())
    flop_a = inter_a * 3 * F_OP["A2k"]          # three directions per interaction
    flop_t = inter_t * F_OP["A2Tk"]             # interactions already summed over the 3 sets
    dom = "A2" if t_a2 >= t_a2t else "A2T_sum"
    dom_kernel = "k_trav_v32<2, 0, false>" if dom == "A2" else "k_trav_s32<3, 0>"
    dom_flop, dom_ms = (flop_a, t_a2) if dom == "A2" else (flop_t, t_a2t)
    achieved = dom_flop / (dom_ms * 1e-3) / 1e12
    log(f"interactions per launch: A2 {inter_a:.3e}, A2T_sum {inter_t:.3e}; dominant {dom}: "
        f"{achieved:.1f} TFLOP/s")

    # ---------------------------------------------------------------- FP64
    fp64 = None
    if not args.no_fp64:
        ks = args.fp64_steps or max(2, args.steps // 4)
        log(f"fp64: 1 warmup + {ks} timed steps")
        el64, (r64, a64), _ = timed("fp64", ks, 1)
        _, _, prof64 = timed("fp64", ks, 0, profile=True)
        tr64 = np.array(prof64["traverse"]).reshape(ks, 2)
        # same inputs for both precisions (A^T.r is applied to the FP32c A.mu:
        # its node centres follow |r|, so different inputs give different trees)
        r64s = wt.forward_A2T_sum(-a32, w, precision="fp64")
        a_err = ((a32.double() - a64.double()).abs().max() / a64.double().abs().max()).item()
        r_err = ((r32.double() - r64s.double()).abs().max() / r64s.double().abs().max()).item()
        fp64 = dict(value=ks / el64, ms_per_step=el64 / ks * 1e3,
                    traverse_ms=[float(tr64[:, 0].mean()), float(tr64[:, 1].mean())],
                    fp32c_err_vs_fp64={"A_mu": a_err, "AT_r": r_err})
        log(f"fp64 {fp64['value']:.2f} pairs/s; fp32c normwise err vs fp64: A {a_err:.2e}, AT {r_err:.2e}")

    # ---------------------------------------------------------------- solve
    solve = None
    if not args.no_solve:
        from cagr_amd.solve import NormalOperator, cgnr_solve, cgnr_solve_mixed
        # warm both precisions' kernels once so compilation and first-touch stay out
        op = NormalOperator(wt, args.width)
        torch.cuda.synchronize()
        t = time.perf_counter()
        opp = NormalOperator(wt, args.width, plan=True)
        torch.cuda.synchronize()
        plan_ms = (time.perf_counter() - t) * 1e3
        for prec in ("fp64", "fp32c"):
            op.AT(op.A(torch.zeros((n, 3), dtype=torch.float64, device=dev), prec), prec)
        opp.AT(op.A(torch.zeros((n, 3), dtype=torch.float64, device=dev), "fp32c"), "fp32c")
        torch.cuda.synchronize()
        # per-application times of the solve's operator and its exact adjoint
        op_ms = {}
        x3 = mu.double()
        for prec in ("fp32c", "fp64"):
            reps = 5
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                y = op.A(x3, prec)
            torch.cuda.synchronize()
            ta = (time.perf_counter() - t) / reps * 1e3
            t = time.perf_counter()
            for _ in range(reps):
                op.AT(y, prec)
            torch.cuda.synchronize()
            tat = (time.perf_counter() - t) / reps * 1e3
            op_ms[prec] = {"A": round(ta, 4), "AT": round(tat, 4)}
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(5):
            opp.AT(y, "fp32c")
        torch.cuda.synchronize()
        op_ms["fp32c"]["AT_planned"] = round((time.perf_counter() - t) / 5 * 1e3, 4)
        op_ms["plan_build"] = round(plan_ms, 3)
        log(f"solve operator ms: {op_ms}")
        log(f"solve: fp64 CGNR to {args.solve_tol:g}")
        mu64, i64 = cgnr_solve(wt, args.width, tol=args.solve_tol, max_iter=args.solve_max_iter,
                               log=log)
        log(f"solve: fp64 {i64['iters']} iterations ({i64['restarts']} restarts), {i64['seconds']:.3f} s, "
            f"residual {i64['rel_residual']:.3e}")
        mu32, i32 = cgnr_solve_mixed(wt, args.width, tol=args.solve_tol,
                                     max_iter=args.solve_max_iter, log=log)
        plain = None
        if args.solve_fp64_plain:
            log(f"solve: fp64 CGNR to {args.solve_tol:g}, one CG run (no restarts)")
            _, ipl = cgnr_solve(wt, args.width, tol=args.solve_tol, max_iter=args.solve_max_iter,
                                log=log, stall=0)
            plain = {"iters": ipl["iters"], "seconds": round(ipl["seconds"], 4),
                     "rel_residual": ipl["rel_residual"],
                     "speedup_of_mixed": round(ipl["seconds"] / i32["seconds"], 3)}
            log(f"solve: fp64 without restarts {ipl['iters']} iterations, {ipl['seconds']:.3f} s")
        log(f"solve: mixed {i32['iters']} iterations ({i32['restarts']} restarts), "
            f"{i32['seconds']:.3f} s (adjoint plan built inside), residual {i32['rel_residual']:.3e}")
        # the adjoint plan's worth, at the sweep tolerance (a shorter solve)
        _, ip = cgnr_solve_mixed(wt, args.width, tol=args.sweep_solve_tol)
        _, ia = cgnr_solve_mixed(wt, args.width, tol=args.sweep_solve_tol, plan=False)
        log(f"solve to {args.sweep_solve_tol:g}: mixed {ip['iters']} iterations, "
            f"{ip['seconds']:.3f} s; without plan {ia['iters']} iterations, {ia['seconds']:.3f} s")
        dmu = float(torch.linalg.norm(mu32 - mu64) / torch.linalg.norm(mu64))
        solve = {
            "tol": args.solve_tol,
            "fp64": {"iters": i64["iters"], "restarts": i64["restarts"],
                     "seconds": round(i64["seconds"], 4), "rel_residual": i64["rel_residual"]},
            "fp64_one_cg_run": plain,
            "mixed": {"iters": i32["iters"], "restarts": i32["restarts"],
                      "seconds": round(i32["seconds"], 4), "rel_residual": i32["rel_residual"]},
            "speedup": round(i64["seconds"] / i32["seconds"], 3),
            "build_s": round((build_ms + balance_ms) / 1e3, 5),
            "speedup_with_build": round((i64["seconds"] + (build_ms + balance_ms) / 1e3)
                                        / (i32["seconds"] + (build_ms + balance_ms) / 1e3), 3),
            "mixed_plan_vs_atomic_adjoint": {
                "tol": args.sweep_solve_tol,
                "planned": {"iters": ip["iters"], "seconds": round(ip["seconds"], 4)},
                "atomic": {"iters": ia["iters"], "seconds": round(ia["seconds"], 4)}},
            "operator_ms": op_ms,
            "mu_rel_diff": dmu,
        }
        log(f"solve: speedup {solve['speedup']:.2f}x, ||mu32 - mu64|| / ||mu64|| = {dmu:.2e}")
        del mu64, mu32, op, opp

    # ---------------------------------------------------------------- other seeds
    sweep = None
    if args.seeds > 1:
        sweep = {"seeds": [args.seed], "fp32c_pairs_per_s": [round(value, 4)],
                 "fp64_pairs_per_s": [round(fp64["value"], 4)] if fp64 else None,
                 "solve_speedup": [solve["speedup"]] if solve else None,
                 "solve_tol": [args.solve_tol, args.sweep_solve_tol] if solve else None}
        del wt
        torch.cuda.empty_cache()
        for sd in range(args.seed + 1, args.seed + args.seeds):
            r = seed_run(args, sd, fp64 is not None, solve is not None)
            sweep["seeds"].append(sd)
            sweep["fp32c_pairs_per_s"].append(r["fp32c"])
            if fp64:
                sweep["fp64_pairs_per_s"].append(r["fp64"])
            if solve:
                sweep["solve_speedup"].append(r["solve_speedup"])
        sweep["median_fp32c_pairs_per_s"] = float(np.median(sweep["fp32c_pairs_per_s"]))
        if fp64:
            sweep["median_speedup_vs_fp64"] = round(float(np.median(
                np.array(sweep["fp32c_pairs_per_s"]) / np.array(sweep["fp64_pairs_per_s"]))), 3)
        if solve:
            sweep["median_solve_speedup"] = float(np.median(sweep["solve_speedup"]))
        log(f"seeds {sweep['seeds']}: fp32c {sweep['fp32c_pairs_per_s']} pairs/s, median "
            f"{sweep['median_fp32c_pairs_per_s']:.2f}")

    # ---------------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pts, mu_np, s3.cpu().numpy().astype(np.float64), args, r32, a32)

    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "matvec pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "1M points on 6 seeded ellipsoids (sigma 5e-4), A.mu + A^T.r matvec pair, "
                        "FP32 interactions with FP64 totals, the FP64 operator's far/near and "
                        "cutoff decisions (FP32 tests, FP64 retests in bounds)",
            "points": n, "leaf_size": args.leaf, "theta": args.theta, "width": args.width,
            "nodes": info["num_nodes"], "depth": info["depth"],
        },
        "build_ms": round(build_ms, 3),
        "balance_ms": round(balance_ms, 3),
        "roofline": {
            "bound": "valu",
            "kernel": f"{dom_kernel} ({dom}, fp32c)",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_VECTOR_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_VECTOR_TFLOPS, 4),
            "traffic": measured_traffic(dom_kernel, args),
            "traffic_source": TRAFFIC_FILE,
            "flop_per_launch": dom_flop,
            "launch_ms": round(float(dom_ms), 4),
        },
        "kernels_ms": {"traverse_A2": round(float(t_a2), 4), "traverse_A2T_sum": round(float(t_a2t), 4),
                       "upward_A2": round(float(upw[:, 0].mean()), 4),
                       "upward_A2T_sum": round(float(upw[:, 1].mean()), 4)},
        "interactions": {"A2": inter_a, "A2T_sum": inter_t},
    }
    if fp64:
        line["fp64_value"] = round(fp64["value"], 4)
        line["speedup_vs_fp64"] = round(value / fp64["value"], 3)
        line["fp64"] = fp64
    if solve:
        line["solve"] = solve
    if sweep:
        line["seed_sweep"] = sweep
    if cpu:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


def seed_run(args, seed, with_fp64, with_solve):
    """The 1M workload on another seed: same setup as the main run (tree
    build, walks balanced), then timed FP32c and FP64 pairs and the two
    tolerance solves."""
    import torch
    from cagr_amd import WindingNumberTreecode, synth
    from cagr_amd.solve import cgnr_solve, cgnr_solve_mixed
    dev = torch.device("cuda:0")
    n = args.points
    pts = synth.normalize(synth.ellipsoids(n, seed)).astype(np.float32)
    mu = torch.from_numpy(synth.random_normals(n, seed + 1).astype(np.float32)).to(dev)
    w = torch.full((n,), args.width, dtype=torch.float32, device=dev)
    wt = WindingNumberTreecode(torch.from_numpy(pts).to(dev), leaf_size=args.leaf,
                               theta=args.theta, precision="fp32c")
    wt.balance("A2", mu)
    out = {}
    for prec, steps, warm in (("fp32c", args.steps, args.warmup),
                              ("fp64", args.fp64_steps or max(2, args.steps // 4), 1)):
        if prec == "fp64" and not with_fp64:
            continue
        for _ in range(warm):
            wt.forward_A2T_sum(-wt.forward_A2(mu, w, precision=prec), w, precision=prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            wt.forward_A2T_sum(-wt.forward_A2(mu, w, precision=prec), w, precision=prec)
        torch.cuda.synchronize()
        out[prec] = round(steps / (time.perf_counter() - t), 4)
    if with_solve:
        _, i64 = cgnr_solve(wt, args.width, tol=args.sweep_solve_tol)
        _, i32 = cgnr_solve_mixed(wt, args.width, tol=args.sweep_solve_tol)
        out["solve_speedup"] = round(i64["seconds"] / i32["seconds"], 3)
    log(f"seed {seed}: {out}")
    del wt
    torch.cuda.empty_cache()
    return out


def cpu_baseline(pts, mu_np, s3, args, r32, a32):
    """The FP64 oracle (CPU restatement, OpenMP) on a bounded random sample of
    queries of the same workload, extrapolated to all N queries."""
    import oracle
    n = pts.shape[0]
    p64 = pts.astype(np.float64)
    mu64 = mu_np.astype(np.float64)
    L = oracle.lib()
    cores = int(L.orc_num_threads())
    t0 = time.perf_counter()
    tree = oracle.build_tree(p64, 15, args.leaf)
    t_build = time.perf_counter() - t0
    log(f"cpu: oracle tree built in {t_build:.1f} s ({tree.num_nodes} nodes), {cores} threads")
    # upward passes of one matvec pair (1 for A2k, 3 for A2Tk), full size
    t0 = time.perf_counter()
    mom_mu = oracle.upward(tree, p64, mu64)
    mom_s = [oracle.upward(tree, p64, s3[:, k:k + 1]) for k in range(3)]
    t_up = time.perf_counter() - t0

    rng = np.random.default_rng(args.seed + 7)

    def run(q_idx):
        q = p64[q_idx]
        wq = np.full(q.shape[0], args.width)
        t = time.perf_counter()
        outs = [oracle.apply(f"A2{k + 1}", tree, p64, mu64, wq, queries=q, node_moments=mom_mu,
                             theta=args.theta) for k in range(3)]
        outs_t = [oracle.apply(f"A2T{k + 1}", tree, p64, s3[:, k:k + 1], wq, queries=q,
                               node_moments=mom_s[k], theta=args.theta) for k in range(3)]
        return time.perf_counter() - t, np.concatenate(outs, 1), sum(outs_t)

    # size the sample to about --cpu-seconds of host work: a pilot, then grow
    # the sample until one run takes at least half the target (fixed per-call
    # costs make small pilots overestimate the per-query time)
    m = 4096
    while True:
        q_idx = np.sort(rng.choice(n, m, replace=False))
        dt, ref_a, ref_r = run(q_idx)
        if dt >= 0.5 * args.cpu_seconds or m >= n:
            break
        m = int(min(n, max(2 * m, m * args.cpu_seconds / max(dt, 1e-3))))
    per_pair = t_up + dt * n / m
    # the reference's as-shipped team size (omp_set_num_threads(20),
    # cpu_kernels.cpp:244) on the same sample, however many cores the host has
    L.orc_set_num_threads(20)
    t0 = time.perf_counter()
    oracle.upward(tree, p64, mu64)
    for k in range(3):
        oracle.upward(tree, p64, s3[:, k:k + 1])
    t_up20 = time.perf_counter() - t0
    dt20, _, _ = run(q_idx)
    L.orc_set_num_threads(cores)
    per_pair20 = t_up20 + dt20 * n / m
    log(f"cpu: the same sample with 20 threads (the reference's setting): {1 / per_pair20:.5f} pairs/s")
    # accuracy of the GPU FP32c result on the sampled queries against the FP64 oracle
    ga = a32.double().cpu().numpy()[q_idx]
    gr = r32.double().cpu().numpy()[q_idx]
    ea = float(np.abs(ga - ref_a).max() / np.abs(ref_a).max())
    er = float(np.abs(gr - ref_r).max() / np.abs(ref_r).max())
    log(f"cpu: {m} sampled queries in {dt:.1f} s + upward {t_up:.2f} s -> {1 / per_pair:.5f} pairs/s; "
        f"gpu fp32c err vs oracle on sample: A {ea:.2e}, AT {er:.2e}")
    return {
        "value": round(1.0 / per_pair, 6),
        "unit": "matvec pairs/s",
        "cores": cores,
        "host_cpu": host_cpu(),
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "unset"),
        "as_shipped_20_threads": {"value": round(1.0 / per_pair20, 6),
                                  "note": f"omp_set_num_threads(20) as cpu_kernels.cpp:244 "
                                          f"hard-codes it, on {_USABLE_CORES} usable host "
                                          f"cores"},
        "kind": "port",
        "sample": f"{m} random queries of the same 1M-point tree (all 6 operators), full-size "
                  f"upward passes ({t_up:.2f} s), extrapolated linearly to {n} queries; "
                  f"oracle tree build {t_build:.1f} s not included",
        "gpu_fp32c_err_vs_oracle_on_sample": {"A_mu": ea, "AT_r": er},
    }


_USABLE_CORES = len(os.sched_getaffinity(0))   # before any OpenMP team pins this thread


def host_cpu() -> str:
    """The host CPU model (first model name of /proc/cpuinfo) and the cores
    this process could run on when it started."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}; {_USABLE_CORES} usable cores of {os.cpu_count()}"


def _sync_time(fn, reps=1):
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps, out


def _pairs_per_s(wt, mu, w, precision, steps):
    def step():
        a = wt.forward_A2(mu, w, precision=precision)
        return wt.forward_A2T_sum(-a, w, precision=precision)
    step()
    dt, _ = _sync_time(step, steps)
    return 1.0 / dt


def config_c1(args):
    """10k sphere (sigma 0.002), leaf 32, theta 0.5: the reference's schedule
    (main_CAGR.py:90-174, preset l0, An = 1) on the host cores with the FP64
    oracle and on the GPU in both precisions; 128^3 grid query."""
    import torch
    import oracle
    from oracle import solve_ref
    from cagr_amd import WindingNumberTreecode, synth
    from cagr_amd.query import grid_values, sample_attributes
    from cagr_amd.solve import PRESETS, cagr_schedule
    dev = torch.device("cuda:0")
    n = 10_000
    pts = synth.normalize(synth.sphere(n, args.seed, 0.002)).astype(np.float32)
    wsmin, wsmax = PRESETS["l0"]
    out = {"metric": "schedule seconds (40 iterations + 8 G steps) and 128^3 grid query seconds",
           "config": {"workload": "10k points on the unit sphere, sigma 0.002, leaf 32, theta 0.5, "
                                  "preset l0, An 1", "points": n}, "data": "synthetic"}
    pt = torch.from_numpy(pts).to(dev)
    res = {}
    for prec, dt in (("fp64", torch.float64), ("fp32c", torch.float32)):
        wt = WindingNumberTreecode(pt.to(dt), leaf_size=32, theta=0.5, precision=prec)
        cagr_schedule(wt, wsmin, wsmax, precision=prec)   # warm-up
        sec, (normals, rn) = _sync_time(lambda: cagr_schedule(wt, wsmin, wsmax, precision=prec))
        mu = sample_attributes(normals)
        grid_values(wt, mu, [-1.2] * 3, [1.2] * 3, [128] * 3, 0.01, precision=prec)
        gsec, vals = _sync_time(lambda: grid_values(wt, mu, [-1.2] * 3, [1.2] * 3, [128] * 3, 0.01,
                                                    precision=prec))
        res[prec] = {"schedule_s": round(sec, 4), "grid128_s": round(gsec, 4),
                     "normals": normals.double().cpu().numpy()}
        log(f"c1 {prec}: schedule {sec:.3f} s, 128^3 grid {gsec * 1e3:.1f} ms")
    L = oracle.lib()
    cores = int(L.orc_num_threads())
    t = time.perf_counter()
    ref_normals, _ = solve_ref.cagr_schedule(pts.astype(np.float64), wsmin, wsmax, leaf_size=32,
                                             theta=0.5)
    cpu_s = time.perf_counter() - t
    log(f"c1 cpu oracle schedule {cpu_s:.2f} s on {cores} threads")
    def diff(a):
        return float(np.abs(a - ref_normals).max() / np.abs(ref_normals).max())
    out["gpu"] = {k: {kk: vv for kk, vv in v.items() if kk != "normals"} for k, v in res.items()}
    out["normals_err_vs_oracle"] = {k: diff(v["normals"]) for k, v in res.items()}
    out["cpu_baseline"] = {"value": round(cpu_s, 3), "unit": "s (schedule)", "cores": cores,
                           "kind": "port", "sample": "the whole 40-iteration schedule + 8 G steps"}
    out["speedup_fp32c_vs_cpu"] = round(cpu_s / res["fp32c"]["schedule_s"], 1)
    out["speedup_fp32c_vs_fp64"] = round(res["fp64"]["schedule_s"] / res["fp32c"]["schedule_s"], 3)
    return out


def config_c2(args):
    """100k torus (R 1, r 0.3, sigma 0.001, 5 % uniform outliers), leaf 32,
    theta 0.5: FP64 drop-in matvec pairs/s (and FP32c), 256^3 grid query."""
    import torch
    from cagr_amd import WindingNumberTreecode, synth
    from cagr_amd.query import grid_values, sample_attributes
    dev = torch.device("cuda:0")
    n = 100_000
    pts = synth.normalize(synth.torus(n, args.seed)).astype(np.float32)
    mu_np = synth.random_normals(n, args.seed + 1)
    out = {"metric": "matvec pairs/s and 256^3 grid query seconds",
           "config": {"workload": "100k torus points, sigma 0.001, 5% outliers, leaf 32, theta 0.5, "
                                  "w 0.002 (pairs), w 0.01 (grid)", "points": n},
           "data": "synthetic"}
    for prec, dt in (("fp64", torch.float64), ("fp32c", torch.float32)):
        pt = torch.from_numpy(pts).to(dev).to(dt)
        wt = WindingNumberTreecode(pt, leaf_size=32, theta=0.5, precision=prec)
        mu = torch.from_numpy(mu_np).to(dev).to(dt)
        w = torch.full((n,), 0.002, dtype=dt, device=dev)
        pps = _pairs_per_s(wt, mu, w, prec, 10)
        att = sample_attributes(mu)
        grid_values(wt, att, [-1.2] * 3, [1.2] * 3, [64] * 3, 0.01, precision=prec)
        gsec, vals = _sync_time(lambda: grid_values(wt, att, [-1.2] * 3, [1.2] * 3, [256] * 3, 0.01,
                                                    precision=prec))
        out[prec] = {"pairs_per_s": round(pps, 3), "grid256_s": round(gsec, 4),
                     "grid256_queries_per_s": round(256 ** 3 / gsec, 1)}
        if prec == "fp64":
            g64 = vals.double()
        else:
            out[prec]["grid_err_vs_fp64"] = float((vals.double() - g64).abs().max() / g64.abs().max())
        log(f"c2 {prec}: {pps:.1f} pairs/s, 256^3 grid {gsec:.3f} s")
    out["speedup_fp32c_vs_fp64"] = {"pairs": round(out["fp32c"]["pairs_per_s"] / out["fp64"]["pairs_per_s"], 3),
                                    "grid": round(out["fp64"]["grid256_s"] / out["fp32c"]["grid256_s"], 3)}
    return out


def config_c4(args):
    """500k displaced sphere: loss = ||A2 mu - target||^2 through the autograd
    form; forward + backward time in FP32c and FP64; FP64 gradient against
    central differences (h = 1e-4) on 64 random entries; FP32c gradient
    against the FP64 one."""
    import torch
    from cagr_amd import WindingNumberTreecode, synth
    from cagr_amd.autograd import treecode_op
    dev = torch.device("cuda:0")
    n = 500_000
    pts = synth.normalize(synth.displaced_sphere(n, args.seed)).astype(np.float32)
    gen = torch.Generator(device="cpu").manual_seed(args.seed + 3)
    mu0 = torch.randn(n, 3, generator=gen, dtype=torch.float64).to(dev)
    target = torch.randn(n, 3, generator=gen, dtype=torch.float64).to(dev)
    out = {"metric": "autograd forward+backward seconds; gradient error",
           "config": {"workload": "500k displaced-sphere points, leaf 32, theta 0.5, w 0.002, A2",
                      "points": n}, "data": "synthetic"}
    grads = {}
    for prec, dt in (("fp64", torch.float64), ("fp32c", torch.float32)):
        wt = WindingNumberTreecode(torch.from_numpy(pts).to(dev).to(dt), leaf_size=32, theta=0.5,
                                   precision=prec)
        widths = torch.full((n,), 0.002, dtype=dt, device=dev)

        def fb():
            mu = mu0.to(dt).clone().requires_grad_(True)
            y = treecode_op(wt, mu, widths, op="A2", precision=prec)
            loss = ((y - target.to(dt)) ** 2).sum()
            loss.backward()
            return mu.grad
        fb()
        sec, g = _sync_time(fb, 3)
        grads[prec] = g.double()
        out[prec] = {"fwd_bwd_s": round(sec, 5)}
        log(f"c4 {prec}: forward+backward {sec * 1e3:.2f} ms")
        if prec == "fp64":
            def loss_fn(m):
                y = treecode_op(wt, m, widths, op="A2", precision="fp64")
                return float(((y - target) ** 2).sum())
            h = 1e-4
            idx = np.random.default_rng(args.seed + 4).choice(n * 3, 64, replace=False)
            fd = []
            with torch.no_grad():
                for k in idx:
                    e = torch.zeros(n * 3, dtype=torch.float64, device=dev)
                    e[k] = h
                    e = e.view(n, 3)
                    fd.append((loss_fn(mu0 + e) - loss_fn(mu0 - e)) / (2 * h))
            fd = np.array(fd)
            got = g.double().cpu().numpy().reshape(-1)[idx]
            out["fp64"]["fd_rel_err"] = float(np.abs(got - fd).max() / np.abs(fd).max())
    out["fp32c"]["grad_err_vs_fp64"] = float((grads["fp32c"] - grads["fp64"]).abs().max()
                                             / grads["fp64"].abs().max())
    out["speedup_fp32c_vs_fp64"] = round(out["fp64"]["fwd_bwd_s"] / out["fp32c"]["fwd_bwd_s"], 3)
    return out


def config_c5(args):
    """64 clouds (sizes log-uniform in [2e3, 2e5], sphere / torus / ellipsoid,
    sigma 0.001): autotune leaf {16, 32, 64} x theta {0.4, 0.5, 0.6} x waves
    per block {1, 2, 4, 8} under the error budget against direct sums (5e-2,
    the tuner's default), once for the whole batch and once per size class
    (< 10k, 10k..60k, >= 60k points); then the tuned batched persistent form,
    the per-class forests, the batched form at leaf 32 / theta 0.5 and one
    single-cloud launch pair per cloud at leaf 32 / theta 0.5."""
    import torch
    from cagr_amd import WindingNumberTreecode, synth
    from cagr_amd.autotune import ClassBatches, autotune, autotune_classes
    from cagr_amd.batched import BatchedTreecode
    dev = torch.device("cuda:0")
    clouds_np = synth.batch_clouds(64, args.seed)
    clouds = [torch.from_numpy(c.astype(np.float32)).to(dev) for c in clouds_np]
    mus = [torch.from_numpy(synth.random_normals(c.shape[0], args.seed + i).astype(np.float32)).to(dev)
           for i, c in enumerate(clouds_np)]
    total = int(sum(c.shape[0] for c in clouds_np))
    rep = autotune(clouds, mus, 0.002, steps=3, rounds=3, log=log)
    b = rep.best
    log(f"c5 tuned: leaf {b.leaf_size}, theta {b.theta}, waves/block {b.waves_per_block}: "
        f"{b.ms_per_pair:.3f} ms/pair (error {b.err:.2e}, budget {rep.budget:.2e}, "
        f"stable {rep.stable}, lead {rep.lead:.1%})")
    classes = autotune_classes(clouds, mus, 0.002, steps=3, rounds=3, log=log)
    mu = torch.cat(mus).contiguous()
    w = torch.full((total,), 0.002, dtype=torch.float32, device=dev)
    bt = BatchedTreecode(clouds, leaf_size=b.leaf_size, theta=b.theta, precision="fp32c")
    bt.balance("A2", mu)
    bt.set_launch(b.waves_per_block)
    tuned = _pairs_per_s(bt, mu, w, "fp32c", 10)
    del bt
    cb = ClassBatches(clouds, classes)
    cb.balance(mu)
    per_class = _pairs_per_s(cb, mu, w, "fp32c", 10)
    del cb
    bd = BatchedTreecode(clouds, leaf_size=32, theta=0.5, precision="fp32c")
    bd.balance("A2", mu)
    default = _pairs_per_s(bd, mu, w, "fp32c", 10)
    del bd
    singles = [WindingNumberTreecode(c, leaf_size=32, theta=0.5, precision="fp32c") for c in clouds]
    ws = [torch.full((c.shape[0],), 0.002, dtype=torch.float32, device=dev) for c in clouds]

    def all_singles():
        for wt, m, wi in zip(singles, mus, ws):
            a = wt.forward_A2(m, wi, precision="fp32c")
            wt.forward_A2T_sum(-a, wi, precision="fp32c")
    all_singles()
    sec, _ = _sync_time(all_singles, 5)
    per_cloud = 1.0 / sec
    log(f"c5: tuned batched {tuned:.1f} pairs/s, per-class forests {per_class:.1f}, batched "
        f"leaf 32 / theta 0.5 {default:.1f}, one launch pair per cloud {per_cloud:.1f}")
    return {"metric": "batched matvec pairs/s (all 64 clouds per pair)",
            "config": {"workload": f"64 clouds, {total} points, w 0.002, A2 + A2T_sum, FP32c",
                       "points": total},
            "data": "synthetic",
            "tuned": {"leaf_size": b.leaf_size, "theta": b.theta,
                      "waves_per_block": b.waves_per_block, "pairs_per_s": round(tuned, 3),
                      "sampled_err": b.err, "err_budget": rep.budget, "stable": rep.stable,
                      "lead": round(rep.lead, 4)},
            "per_class": {"pairs_per_s": round(per_class, 3),
                          "classes": [{"clouds": int(len(sel)), "leaf_size": r.best.leaf_size,
                                       "theta": r.best.theta,
                                       "waves_per_block": r.best.waves_per_block,
                                       "sampled_err": r.best.err, "stable": r.stable}
                                      for sel, r in classes]},
            "batched_leaf32_theta05": round(default, 3),
            "per_cloud_launches": round(per_cloud, 3),
            "speedup_tuned_vs_leaf32_theta05": round(tuned / default, 3),
            "speedup_tuned_vs_per_cloud": round(tuned / per_cloud, 3),
            "table": [[r.leaf_size, r.theta, r.waves_per_block, round(r.ms_per_pair, 4), r.err]
                      for r in rep.table],
            "errors": {f"{k[0]}/{k[1]}": v for k, v in rep.errors.items()}}


def config_pgr(args):
    """GaussRecon (main_GaussRecon_cuda.cu:255-486 before marching cubes) on
    the 1M-point ellipsoid cloud, -d 10 -m 1 -w 0.01: the PGR octree query set
    on the GPU (cagr_pgr_build) against the reference's own Octree::setTree
    (oracle/_ref, one host core, as the reference runs it; the port when the
    reference library is absent), then -wn at every corner and at the samples
    with the leaf-1 treecode (FP32c and FP64)."""
    import torch
    import oracle
    from cagr_amd import WindingNumberTreecode, synth
    from cagr_amd.pgr import PGROctree
    from cagr_amd.query import sample_attributes
    dev = torch.device("cuda:0")
    n, D, M, width = args.points, 10, 1, 0.01
    pts = synth.normalize(synth.ellipsoids(n, args.seed)).astype(np.float32)
    nml = synth.random_normals(n, args.seed + 1).astype(np.float32)
    pt = torch.from_numpy(pts).to(dev)
    PGROctree(pt, D, M)
    builds = []
    for _ in range(3):
        sec, oct_ = _sync_time(lambda: PGROctree(pt, D, M))
        builds.append(sec)
    build_s = float(np.median(builds))
    nc, nl = oct_.corners.shape[0], oct_.leaves.shape[0]
    log(f"pgr: GPU build {build_s * 1e3:.1f} ms, {nl} leaves, {nc} corners, depth {oct_.depth}")
    out = {"metric": "PGR query-set build seconds; A3 values at its corners (queries/s); mesh seconds",
           "config": {"workload": f"GaussRecon on {n} ellipsoid points, -d {D} -m {M} -w {width}, "
                                  "leaf-1 treecode, theta 0.5", "points": n},
           "data": "synthetic",
           "pgr": {"build_s": round(build_s, 4), "leaves": nl, "corners": nc,
                   "depth": oct_.depth}}
    p11 = oct_.sample_points()
    mu = sample_attributes(torch.from_numpy(nml).to(dev))
    q = oct_.query_points()
    for prec, dt in (("fp32c", torch.float32), ("fp64", torch.float64)):
        tb, wt = _sync_time(lambda: WindingNumberTreecode(p11.to(dt), leaf_size=1, theta=0.5,
                                                          precision=prec))
        m = mu.to(dt)
        wt.query(q.to(dt), m, width, precision=prec)
        qs, vals = _sync_time(lambda: wt.query(q.to(dt), m, width, precision=prec))
        w = torch.full((n,), width, dtype=dt, device=dev)
        ss, wn_s = _sync_time(lambda: wt.apply("A3", m, w, continuous=True, precision=prec))
        iso = -float(torch.kthvalue(wn_s[:, 0].double().cpu(), n // 2 + 1).values)
        out[prec] = {"tree_build_s": round(tb, 4), "corner_query_s": round(qs, 4),
                     "corner_queries_per_s": round(nc / qs, 1), "sample_query_s": round(ss, 4),
                     "isovalue": iso}
        if prec == "fp32c":
            v32 = vals.double()
        else:
            out["fp32c"]["corner_err_vs_fp64"] = float((v32 - vals.double()).abs().max()
                                                       / vals.double().abs().max())
        log(f"pgr {prec}: corners {qs * 1e3:.1f} ms, samples {ss * 1e3:.1f} ms, iso {iso:.4g}")
        if prec == "fp32c":
            vals32, iso32 = (-vals[:, 0]).float().contiguous(), iso
            oct_.mesh(vals32, iso32, width / 2)
            ms, (mv, mt) = _sync_time(lambda: oct_.mesh(vals32, iso32, width / 2))
            out["mesh"] = {"seconds": round(ms, 4), "vertices": int(mv.shape[0]),
                           "triangles": int(mt.shape[0]),
                           "field": "FP32c -wn at the corners, its isovalue, smoothWidth w/2"}
            log(f"pgr mesh: {ms * 1e3:.1f} ms, {mv.shape[0]} vertices, {mt.shape[0]} triangles")
    # the whole of main_GaussRecon_cuda.cu as one call (octree, treecode,
    # -wn at corners and samples, isovalue, mesh, vertices back to the host)
    from cagr_amd.pgr import gauss_recon
    nml_t = torch.from_numpy(nml).to(dev)
    gauss_recon(pt, nml_t, None, width=width, max_depth=D, min_depth=M, precision="fp32c")
    gs, (gv, gt, _) = _sync_time(lambda: gauss_recon(pt, nml_t, None, width=width, max_depth=D,
                                                     min_depth=M, precision="fp32c"))
    out["gauss_recon_s"] = round(gs, 4)
    log(f"pgr: gauss_recon end to end {gs * 1e3:.1f} ms ({len(gv)} vertices, {len(gt)} triangles)")
    if not args.no_cpu:
        kind = "reference" if oracle.pgr_ref_available() else "port"
        fn = oracle.pgr_ref_build if kind == "reference" else oracle.pgr_build
        t = time.perf_counter()
        ref = fn(pts, D, M)
        cpu_s = time.perf_counter() - t
        same = bool(np.array_equal(ref.corners, oct_.corners.cpu().numpy())
                    and np.array_equal(ref.leaves, oct_.leaves.cpu().numpy()))
        log(f"pgr: CPU {kind} build {cpu_s:.2f} s, identical query set: {same}")
        if kind == "reference" and "mesh" in out:
            t = time.perf_counter()
            rv, roff, ridx = oracle.pgr_ref_mesh(vals32.cpu().numpy(), iso32, width / 2)
            mesh_cpu = time.perf_counter() - t
            out["mesh"]["cpu_reference_s"] = round(mesh_cpu, 3)
            out["mesh"]["identical_to_cpu"] = bool(
                np.array_equal(rv, mv.cpu().numpy())
                and np.array_equal(ridx.reshape(-1, 3), mt.cpu().numpy()))
            out["mesh"]["speedup_vs_cpu"] = round(mesh_cpu / out["mesh"]["seconds"], 1)
            log(f"pgr: CPU reference mesh {mesh_cpu:.2f} s, identical: "
                f"{out['mesh']['identical_to_cpu']}")
        out["cpu_baseline"] = {"value": round(cpu_s, 3), "unit": "s (PGR build)", "cores": 1,
                               "kind": kind,
                               "sample": "the whole build on the same 1M points"
                                         + (" (incl. writing and parsing its .xyz file)"
                                            if kind == "reference" else "")}
        out["pgr"]["identical_to_cpu"] = same
        out["pgr"]["speedup_vs_cpu"] = round(cpu_s / build_s, 1)
    return out


if __name__ == "__main__":
    main()
