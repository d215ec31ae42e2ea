"""Benchmark: semi-discrete OT dual solves to 1e-5 max relative cell-mass error.

Default workload (--workload shapes65k): 256x256 jittered sites (N = 65,536), 1024x1024 target
of 40 random disks and rectangles at 100:1 contrast, uniform source masses 1/N, mixed precision
(FP32 cell geometry and integration, FP64 accumulation and L-BFGS, m = 10), until
max_i |mass_i - nu_i| / nu_i < 1e-5.

--workload curves262k: N = 262,144 sites on a 30-degree tilted source plane, 2048x2048 target
of 30 cubic Bezier strokes.

--workload batch64: 64 uneven problems (N_k from 2,048 to 32,768, 512x512 Gaussian-mixture
targets) solved together in one batched context, each to max-rel 1e-5.

One step = complete solves from nothing but the problems, the inputs resident in host memory
as the caller hands them over: the context is created inside the timed region (target prefix
tables, site bucketing, uploads) and the coarse-to-fine solve runs there too (cot_solve_ladder:
each problem on coarse jittered grids -- 16, 32, 64 and 128 sites a side for shapes65k, up to
256 for curves262k -- with box-downsampled targets, solved to max-rel 0.1, each solution carried
up through the weight field; then the full-resolution problem; every stage L-BFGS until max-rel
2 with no empty cell and damped Newton steps after: the library's defaults).  --start learned starts the full-resolution problem from the
network's field instead; --start zero from w = 0.  Launch shapes per level come from the
autotuner (caustic_amd.autotune), run before the timed region.

Outside the timed region, once: the plain GPU solve (FP64, L-BFGS from w = 0 with the
reference's rules and arithmetic, ot_op.cuh:289-537, to the same stopping rule) for the speedup;
the zero start with mixed precision and the Newton tail; a CPU baseline (the FP64 oracle on the
host cores: a bounded sample of the workload's evaluations, and measured solves of the
N = 4,096 configuration); the batch64 workload; the learned-vs-zero comparison.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes moved per evaluation (construction: pre-pass, lane pass, deferred wave pass;
# integration with the conformity check in the same launch) of the N=65,536 workload, measured
# with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on MI355X, each scaled by the same counter's
# reading for a copy of known size (profiles/shapes65k_eval_traffic_pmc.md)
TRAFFIC_BYTES = {"shapes65k": 85.17e6}
# VALU issue utilisation of the evaluation kernels, rocprofv3 SQ counters
# (profiles/shapes65k_eval_valu_pmc.md)
VALU_BUSY = {"shapes65k": {"cell_reappear_kernel": 0.102, "cell_kernel": 0.109,
                           "cell_defer_kernel": 0.016, "integrate_kernel": 0.286}}
WORKLOADS = {
    "shapes65k": {
        "desc": "single large solve: N=65536 jittered sites, 1024x1024 disks/rects target "
                "(100:1), L-BFGS m=10 to max-rel 1e-5, mixed precision",
        "metric": "solves/s to 1e-5 cell-mass err at N=65k",
    },
    "curves262k": {
        "desc": "single large solve: N=262144 jittered sites on a 30-degree tilted plane, "
                "2048x2048 target of 30 Bezier strokes (100:1), L-BFGS m=10 to max-rel 1e-5, "
                "mixed precision",
        "metric": "solves/s to 1e-5 cell-mass err at N=262k",
    },
    "batch64": {
        "desc": "64 uneven problems in one batch: N_k in [2048, 32768] jittered sites, 512x512 "
                "Gaussian-mixture targets, L-BFGS m=10 to max-rel 1e-5 each, mixed precision",
        "metric": "solves/s to 1e-5 cell-mass err, 64 uneven problems batched",
    },
}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(problem, evals_per_solve, budget_s=20.0):
    """Oracle (FP64 C++/OpenMP + qhull) evaluation time on the host, scaled to solves/s with the
    plain GPU solve's evaluation count (the same algorithm: FP64 L-BFGS from w = 0).  Sample:
    repeated full evaluations of the same problem at zero weights and at small random weights,
    until budget_s seconds have elapsed."""
    from oracle import oracle as O

    rng = np.random.default_rng(0)
    t0 = time.perf_counter()
    n_eval = 0
    while True:
        w = np.zeros(problem.n) if n_eval == 0 else rng.normal(0.0, 1e-6, problem.n)
        O.evaluate(problem.sites, w, problem.nu, problem.rho, problem.proj_domain)
        n_eval += 1
        if time.perf_counter() - t0 > budget_s:
            break
    per_eval = (time.perf_counter() - t0) / n_eval
    return {
        "value": 1.0 / (per_eval * evals_per_solve),
        "unit": "solves/s",
        "cores": _threads(),
        "kind": "port",
        "cpu": _cpu_model(),
        "sample": (f"{n_eval} FP64 oracle evaluations of the N=65536, 1024^2 workload "
                   f"({per_eval:.3f} s each), scaled by the plain GPU solve's "
                   f"{evals_per_solve} evaluations (FP64 L-BFGS from w = 0)"),
    }


def cpu_c1_solves(n_solves=3, log=print):
    """Measured CPU solves: the FP64 oracle's L-BFGS (the reference's rules) of the N = 4,096
    configuration (64x64 jittered grid, 256^2 FP64 target of 3 Gaussians, m = 10, from w = 0)
    to max-rel 1e-4, median of n_solves."""
    from caustic_amd import workloads as W
    from oracle import oracle as O

    p = W.plain_fp64()
    times, rows = [], []
    for _ in range(n_solves):
        t0 = time.perf_counter()
        r = O.solve(p, m=10, stop="maxrel", tol=1e-4, total_iter=20000, h0_threshold=1e-5)
        times.append(time.perf_counter() - t0)
        rows.append((r.iterations, r.evaluations, bool(r.converged)))
        log(f"[cpu c1 solve] {times[-1]:.1f} s, {r.iterations} iterations")
    return {"median_seconds": float(np.median(times)), "seconds": times,
            "iterations": [x[0] for x in rows], "evaluations": [x[1] for x in rows],
            "converged": all(x[2] for x in rows), "cores": _threads(), "cpu": _cpu_model(),
            "workload": "N=4096 (64x64 jittered), 256x256 FP64 3-Gaussian target, L-BFGS m=10 "
                        "from w=0 to max-rel 1e-4, FP64 oracle (OpenMP)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="shapes65k", choices=sorted(WORKLOADS))
    ap.add_argument("--start", default="multiscale", choices=["multiscale", "learned", "zero"])
    ap.add_argument("--no-cold", action="store_true",
                    help="skip the zero-start solves outside the timed region")
    ap.add_argument("--plain-iters", type=int, default=20000,
                    help="iteration limit of the plain FP64 L-BFGS solve from w = 0")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-c1-solves", type=int, default=3,
                    help="measured CPU oracle solves of the N = 4,096 configuration (median)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batch64 line")
    ap.add_argument("--learned-seeds", type=int, default=64,
                    help="held-out problems for the learned-vs-zero warm-start comparison "
                         "(batches of 64)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--device-loop", default="graph", choices=["graph", "eager"],
                    help="eager: the same device rounds launched without graph capture")
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch shapes instead of the autotuner's")
    ap.add_argument("--levels", default=None,
                    help="comma-separated coarse grid sides of the warm start (default: the "
                         "library's, cot_default_ladder_opts)")
    ap.add_argument("--coarse-tol", type=float, default=None,
                    help="max-rel the coarse levels are solved to (default: the library's, "
                         "cot_default_ladder_opts)")
    ap.add_argument("--handoff", type=float, default=None,
                    help="max-rel at which every stage hands L-BFGS over to Newton (default: the "
                         "library's)")
    ap.add_argument("--newton-sweeps", type=int, default=2,
                    help="block-local Jacobi sweeps of the Newton steps' CG preconditioner")
    ap.add_argument("--newton-coarse", type=int, default=256,
                    help="patches per problem in the coarse term of the Newton steps' "
                         "preconditioner (0: Jacobi only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build",
              file=sys.stderr)
        return 2

    import torch

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    from caustic_amd import warmstart as WS
    from caustic_amd import workloads as W
    from caustic_amd.ot import OTBatch

    wl = WORKLOADS[args.workload]
    loop = {"graph": 1, "eager": 2}[args.device_loop]
    make_probs = {"shapes65k": lambda: [W.large_shapes()],
                  "curves262k": lambda: [W.tilted_curves()],
                  "batch64": W.batched_uneven}[args.workload]
    probs = make_probs()
    tol = 1e-5
    log = lambda s: print(s, flush=True)
    net = WS.WarmStartNet.load(WS.DEFAULT_MODEL) if args.start == "learned" else None
    # verbose: the solver prints a progress line every 2 s of a solve
    kw = dict(history=10, tol=tol, stop="maxrel", h0_threshold=0.0, max_iter=20000,
              device_loop=loop, verbose=1)
    from caustic_amd.ot import ladder_defaults

    schedule = ladder_defaults(probs)
    coarse_tol = args.coarse_tol if args.coarse_tol is not None else schedule["coarse_tol"]
    handoff = args.handoff if args.handoff is not None else schedule["handoff"]
    newton = dict(newton_from=handoff, newton_coarse=args.newton_coarse,
                  newton_sweeps=args.newton_sweeps)
    if args.levels:
        schedule["sides"] = [int(v) for v in args.levels.split(",")]
        schedule["floors"] = [0.0] * len(schedule["sides"])
    levels = tuple(schedule["sides"])
    tuning = {"final": None, "levels": [None] * len(levels)}

    tune_report = None
    if not args.no_tune:
        # launch shapes per level, outside the timed region, level by level: each level is
        # tuned from the weights the already tuned levels below hand it (the coarse problems
        # built as caustic_amd.warmstart builds them, of the sizes the ladder solves)
        from caustic_amd import autotune as AT

        tune_report = {}
        stretch = dict(iters=400, stop_tol=handoff)
        coarse = WS.coarse_problems(probs, levels)
        fields = None
        for li, (n_side, cps) in enumerate(zip(levels, coarse)):
            init = None if fields is None else np.concatenate(
                [WS.weights_from_field(U, cp.sites, cp.frame) for U, cp in zip(fields, cps)])
            tuning["levels"][li], r = AT.tune(cps, "mixed", w0=init, log=log, **stretch)
            tune_report[f"level{n_side}"] = r
            with OTBatch(cps, precision="mixed", tuning=tuning["levels"][li]) as cb:
                wc, _ = cb.solve(init, **dict(kw, tol=coarse_tol, verbose=0), **newton)
                cen, _ = cb.centroids(wc)
                fields = [WS.field_from_solution(cp.sites, wk, ck, n_side, cp.frame)
                          for cp, wk, ck in zip(cps, cb.split(wc), cb.split(cen.reshape(-1, 2)))]
        w_probe = None if fields is None else np.concatenate(
            [WS.weights_from_field(U, p.sites, p.frame) for U, p in zip(fields, probs)])
        tuning["final"], r = AT.tune(probs, "mixed", w0=w_probe, log=log, **stretch)
        tune_report["final"] = r
        log(f"[autotune] {json.dumps(tune_report)}")

    stats_total = {}

    def add_stats(st):
        for k, v in st.items():
            stats_total[k] = stats_total.get(k, 0) + v

    def one_solve():
        """One timed step: context creation and the whole solve of fresh problems.  Returns
        the full-resolution reports."""
        with OTBatch(probs, precision="mixed", tuning=tuning["final"]) as b:
            if args.start == "multiscale":
                extra = {k: v for k, v in (("coarse_tol", args.coarse_tol),
                                           ("handoff", args.handoff)) if v is not None}
                lad = b.ladder_options(level_tunings=tuning["levels"], sides=levels,
                                       floors=schedule["floors"],
                                       final_floors=schedule["final_floors"], **extra)
                w, rep = b.solve_ladder(ladder=lad, **kw)
            else:
                w0 = None
                if args.start == "learned":
                    w0 = np.concatenate([WS.LEARNED_ALPHA * WS.weights_from_field(
                        WS.smooth_field(U, WS.LEARNED_SIGMA), p.sites, p.frame)
                        for U, p in zip(WS.predict_fields(net, probs), probs)])
                w, rep = b.solve(w0, **kw, **newton)
            add_stats(b.stats())
        return rep

    cold = plain = learned = None
    if not args.no_cold:
        # the plain GPU solve: FP64, L-BFGS from w = 0 with the reference's rules (H0 threshold
        # 1e-5, halving line search, restarts), to the same stopping rule.  The one change: a
        # trial is also accepted when its directional derivative is <= 0 -- with the
        # reference's decrease-only test this solve stalls at max-rel 3.4e-4 (the decrease left,
        # ~|g|^2, falls below the rounding of f) and never reaches 1e-5
        t0 = time.perf_counter()
        with OTBatch(probs, precision="fp64") as b:
            _, rc = b.solve(None, **dict(kw, h0_threshold=1e-5, max_iter=args.plain_iters),
                            accept_rule="derivative", raise_on_fail=False)
        plain = {"seconds": time.perf_counter() - t0,
                 "iterations": max(r["iterations"] for r in rc),
                 "evaluations": sum(r["evaluations"] for r in rc),
                 "converged": all(bool(r["converged"]) for r in rc),
                 "max_rel_err": max(r["max_rel_err"] for r in rc),
                 "algorithm": "FP64 L-BFGS m=10 from w=0, the reference's rules "
                              "(ot_op.cuh:289-537) with derivative-aware acceptance (the "
                              "decrease-only test stalls at max-rel 3.4e-4), max-rel stopping "
                              "rule",
                 "iteration_limit": args.plain_iters}
        log(f"[plain FP64 L-BFGS solve from w=0] {plain}")
        t0 = time.perf_counter()
        with OTBatch(probs, precision="mixed", tuning=tuning["final"]) as b:
            _, rc = b.solve(None, **kw, **newton)
        cold = {"seconds": time.perf_counter() - t0,
                "iterations": max(r["iterations"] for r in rc),
                "evaluations": sum(r["evaluations"] for r in rc),
                "converged": all(bool(r["converged"]) for r in rc),
                "algorithm": f"mixed precision, L-BFGS from w=0 to max-rel {handoff:g} with no "
                             "empty cell, then damped Newton"}
        log(f"[zero start, mixed + Newton] {cold}")
        if args.workload == "shapes65k" and os.path.exists(WS.DEFAULT_MODEL):
            # the learned start at this size (the network's 128 x 128 field resampled to the
            # sites), the same algorithm as the zero start
            lnet = WS.WarmStartNet.load(WS.DEFAULT_MODEL)
            t0 = time.perf_counter()
            with OTBatch(probs, precision="mixed", tuning=tuning["final"]) as b:
                w0, rp, rz = WS.choose_start(probs, WS.learned_start(lnet, probs), "mixed", b)
                _, rc = b.solve(np.concatenate(w0), **kw, **newton)
            learned = {"seconds": time.perf_counter() - t0,
                       "iterations": max(r["iterations"] for r in rc),
                       "converged": all(bool(r["converged"]) for r in rc),
                       "learned_used": bool(rp[0] < rz[0]),
                       "algorithm": "as the zero start, from the network's field (low-passed, "
                                    "halved) where its |mass - nu|_2 beats zeros"}
            lnet.close()
            log(f"[learned start, mixed + Newton] {learned}")
    for _ in range(args.warmup):
        r = one_solve()
        log(f"[warmup] iters {max(x['iterations'] for x in r)} "
            f"maxrel {max(x['max_rel_err'] for x in r):.3e}")
    torch.cuda.synchronize()
    stats_total.clear()
    reps = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        r = one_solve()
        reps.append((r, time.perf_counter() - ts))
        log(f"[step {k}] iters {max(x['iterations'] for x in r)} "
            f"maxrel {max(x['max_rel_err'] for x in r):.3e} {reps[-1][1]:.3f}s")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = stats_total
    conv = all(x["converged"] for r, _ in reps for x in r)
    # dominant work of an evaluation: construction + integration passes of the full batch
    cell_avg = st["cell_seconds"] / max(st["cell_launches"], 1)
    integ_avg = st["integ_seconds"] / max(st["integ_launches"], 1)
    P = sum(p.rho.size for p in probs)
    N = sum(p.n for p in probs)
    alg_bytes = P * 4 + 48 * N  # FP32 density read once; sites, w, nu, mass, f
    achieved = alg_bytes / (cell_avg + integ_avg) / 1e9
    n_solves = len(probs) * args.steps
    step_s = dt / args.steps
    out = {
        "metric": wl["metric"],
        "value": n_solves / dt if conv else 0.0,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * step_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded jittered grids + disks/rectangles or Gaussian-mixture targets)",
        "config": {
            "workload": wl["desc"],
            "start": args.start, "problems": len(probs), "n_sites": N,
            "target_pixels": P,
            "step": "context creation (prefix tables, bucketing, uploads) + coarse-to-fine "
                    "solve (cot_solve_ladder) of fresh problems",
            "final_iterations_max": int(np.mean([max(x["iterations"] for x in r)
                                                 for r, _ in reps])),
            "converged": conv,
            "max_rel_err": max(x["max_rel_err"] for r, _ in reps for x in r),
            "device_loop": args.device_loop,
            "levels": list(levels) if args.start == "multiscale" else [],
            "newton_coarse": args.newton_coarse,
            "newton_sweeps": args.newton_sweeps,
            "final_level_newton_per_solve": {
                "steps": st["newton_steps"] / args.steps,
                "cg_iterations": st["cg_iterations"] / args.steps,
                "coarse_inverse_host_ms": 1e3 * st["coarse_seconds"] / args.steps},
            "launch_shapes": tuning,
        },
        "plain_gpu_solve": plain,
        "speedup_vs_plain": (plain["seconds"] / step_s) if plain else None,
        "speedup_vs_plain_is_lower_bound": bool(plain and not plain["converged"]),
        "zero_start_mixed_newton": cold,
        "learned_start_mixed_newton": learned,
        "roofline": {
            "kernel": "cell evaluation: construction (cell_reappear_kernel, the pre-pass of the "
                      "cells the last evaluation left empty + cell_kernel + its deferred wave "
                      "pass cell_defer_kernel, one in-kernel timer span) + integrate_kernel "
                      "(exact mass integral; the conformity check runs in its last "
                      "workgroups), full-resolution passes of the timed solves",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_BYTES.get(args.workload),
            "limiter": ("latency: data-dependent walks, one site per lane; VALU issue "
                        "utilisation " + json.dumps(VALU_BUSY.get(args.workload)) +
                        " (rocprofv3 SQ counters) -- the HBM fraction is not the binding roof"),
            "avg_launch_us": {"cell_kernel": cell_avg * 1e6, "integrate_kernel": integ_avg * 1e6},
            "launches": st["cell_launches"], "alg_bytes_per_launch": alg_bytes,
        },
        # in-kernel timers of the timed solves' full-resolution passes: seconds and launches
        "kernels": {
            "cell": [st["cell_seconds"], st["cell_launches"]],
            "integrate": [st["integ_seconds"], st["integ_launches"]],
            "repairs": [st["repair_seconds"], st["repair_launches"]],
        },
        "autotune": tune_report,
    }
    if args.workload == "shapes65k" and not args.no_batch:
        out["batch64"] = batch64_line(OTBatch, W, kw, log)
    if not args.no_cpu_baseline and args.workload == "shapes65k":
        n_eval = plain["evaluations"] if plain else 10000
        out["cpu_baseline"] = cpu_baseline(probs[0], n_eval, args.cpu_budget)
        if args.cpu_c1_solves > 0:
            out["cpu_c1_solve"] = cpu_c1_solves(args.cpu_c1_solves, log)
    if args.learned_seeds > 0 and os.path.exists(WS.DEFAULT_MODEL):
        # the metric's second part: the learned warm start against the zero start on held-out
        # N = 16,384 problems (128 x 128 jittered sites, 512 x 512 shapes targets), solved in
        # batches of 64; the library's coarse-to-fine ladder on the same batches for reference
        rows, bts = WS.evaluate_warm_start_batched(
            WS.WarmStartNet.load(WS.DEFAULT_MODEL),
            range(900000, 900000 + args.learned_seeds), newton_from=handoff, log=log)
        ratio = np.array([r["zero_iters"] / max(r["learned_iters"], 1) for r in rows])
        zs, ls_ = sum(b_["zero_s"] for b_ in bts), sum(b_["learned_s"] for b_ in bts)
        out["learned_vs_zero"] = {
            "workload": "N=16,384 (128x128 jittered grid), 512x512 shapes target, mixed "
                        f"precision, L-BFGS to max-rel {handoff:g} then Newton to 1e-5, batches "
                        "of 64 problems",
            "held_out_seeds": [rows[0]["seed"], rows[-1]["seed"]], "problems": len(rows),
            "zero_iterations": int(sum(r["zero_iters"] for r in rows)),
            "learned_iterations": int(sum(r["learned_iters"] for r in rows)),
            "iteration_ratio_median": float(np.median(ratio)),
            "iteration_ratio_p10_p90": [float(np.percentile(ratio, 10)),
                                        float(np.percentile(ratio, 90))],
            "learned_slower_problems": int(np.sum(ratio < 1.0)),
            "zero_seconds": zs, "learned_seconds": ls_, "speedup": zs / ls_,
            "ladder_seconds": sum(b_["ladder_s"] for b_ in bts),
            "learned_used": int(sum(r["learned_used"] for r in rows)),
            "converged": all(r["converged"] for r in rows)}
    print(json.dumps(out), flush=True)
    return 0


def batch64_line(OTBatch, W, kw, log, steps=2):
    """The batched configuration: 64 uneven problems (N_k 2,048-32,768 log-uniform, 512^2
    Gaussian-mixture targets) solved together to max-rel 1e-5 each by cot_solve_ladder, context
    creation included; default launch shapes.  One untimed solve, then `steps` timed ones."""
    probs = W.batched_uneven()
    times, reps = [], None
    for k in range(steps + 1):
        t0 = time.perf_counter()
        with OTBatch(probs, precision="mixed") as b:
            _, reps = b.solve_ladder(**dict(kw, verbose=0), raise_on_fail=False)
        if k:
            times.append(time.perf_counter() - t0)
    its = [r["iterations"] for r in reps]
    line = {"problems": len(probs), "n_sites": int(sum(p.n for p in probs)),
            "solves_per_s": len(probs) * len(times) / sum(times),
            "batch_seconds": float(np.mean(times)),
            "per_problem_latency_s": float(np.mean(times)),
            "converged": int(sum(bool(r["converged"]) for r in reps)),
            "max_rel_err": max(r["max_rel_err"] for r in reps),
            "final_iterations_mean": float(np.mean(its)), "final_iterations_max": int(max(its)),
            "note": "every problem of the batch finishes with the batch: its latency is the "
                    "batch's wall time"}
    log(f"[batch64] {line}")
    return line


if __name__ == "__main__":
    sys.exit(main())
