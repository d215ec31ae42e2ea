"""Benchmark: semi-discrete OT dual solves to 1e-5 max relative cell-mass error.

Default workload (--workload shapes65k): 256x256 jittered sites (N = 65,536), 1024x1024 target
of 40 random disks and rectangles at 100:1 contrast, uniform source masses 1/N, mixed precision
(FP32 cell geometry and integration, FP64 accumulation and L-BFGS, m = 10), until
max_i |mass_i - nu_i| / nu_i < 1e-5.

--workload curves262k: N = 262,144 sites on a 30-degree tilted source plane, 2048x2048 target
of 30 cubic Bezier strokes.

--workload batch64: 64 uneven problems (N_k from 2,048 to 32,768, 512x512 Gaussian-mixture
targets) solved together in one batched context, each to max-rel 1e-5.

One step = complete solves from nothing but the problems, inputs resident in HBM: the warm start
(coarse-to-fine: each problem on coarse jittered grids -- 16, 32, 64 and 128 sites a side for
shapes65k, up to 256 for curves262k, 16 and 32 for batch64 -- with box-downsampled targets,
solved to 1e-3 (--coarse-tol), each solution carried up through the weight field; or the
network's field,
--start learned), then the full-resolution solve.  Each level runs L-BFGS until max-rel 1 and
damped Newton steps after (--newton-from).  Every
level's work is inside the timed region.  The plain solve from zero weights (the reference's
start) is timed once, outside the timed region, for the speedup.  Launch shapes come from the
autotuner (caustic_amd.autotune), run before the timed region.

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
# HBM bytes moved per evaluation (construction + integration launch) of the N=65,536 workload,
# measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on MI355X, each scaled by the same
# counter's reading for a copy of known size (profiles/shapes65k_eval_traffic_pmc.md)
TRAFFIC_BYTES = {"shapes65k": 88.98e6}
WORKLOADS = {
    "shapes65k": {
        "levels": (16, 32, 64, 128),
        "desc": "single large solve: N=65536 jittered sites, 1024x1024 disks/rects target "
                "(100:1), L-BFGS m=10 to max-rel 1e-5, mixed precision",
        "metric": "solves/s to 1e-5 cell-mass err at N=65k",
    },
    "curves262k": {
        "levels": (16, 32, 64, 128, 256),
        "desc": "single large solve: N=262144 jittered sites on a 30-degree tilted plane, "
                "2048x2048 target of 30 Bezier strokes (100:1), L-BFGS m=10 to max-rel 1e-5, "
                "mixed precision",
        "metric": "solves/s to 1e-5 cell-mass err at N=262k",
    },
    "batch64": {
        "levels": (16, 32),
        "desc": "64 uneven problems in one batch: N_k in [2048, 32768] jittered sites, 512x512 "
                "Gaussian-mixture targets, L-BFGS m=10 to max-rel 1e-5 each, mixed precision",
        "metric": "solves/s to 1e-5 cell-mass err, 64 uneven problems batched",
    },
}


def cpu_baseline(problem, evals_per_solve, budget_s=20.0):
    """Oracle (FP64 C++/OpenMP + qhull) evaluation time on the host, scaled to solves/s with the
    plain GPU solve's evaluation count.  Sample: repeated full evaluations of the same problem
    at zero weights and at small random weights, until budget_s seconds have elapsed."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
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
        "cores": threads,
        "kind": "port",
        "sample": (f"{n_eval} FP64 oracle evaluations of the N=65536, 1024^2 workload "
                   f"({per_eval:.3f} s each), scaled by the plain (zero-start) GPU solve's "
                   f"{evals_per_solve} evaluations"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="shapes65k", choices=sorted(WORKLOADS))
    ap.add_argument("--start", default="multiscale", choices=["multiscale", "learned", "zero"])
    ap.add_argument("--no-cold", action="store_true", help="skip the plain zero-start solve")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--learned-seeds", type=int, default=4,
                    help="held-out problems for the learned-vs-zero warm-start comparison")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--device-loop", default="graph", choices=["graph", "eager"],
                    help="eager: the same device rounds launched without graph capture")
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch shapes instead of the autotuner's")
    ap.add_argument("--levels", default=None,
                    help="comma-separated coarse grid sides of the warm start (default: the "
                         "workload's)")
    ap.add_argument("--coarse-tol", type=float, default=1e-2,
                    help="max-rel to which each coarse level of the warm start is solved")
    ap.add_argument("--newton-from", type=float, default=1.0,
                    help="max-rel at which each level hands over from L-BFGS to damped Newton "
                         "steps (0: L-BFGS to the end)")
    ap.add_argument("--coarse-precision", default="mixed", choices=["mixed", "fp64"],
                    help="geometry precision of the warm start's coarse levels")
    ap.add_argument("--newton-sweeps", type=int, default=2,
                    help="block-local Jacobi sweeps of the Newton steps' preconditioner")
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
    probs = {"shapes65k": lambda: [W.large_shapes()], "curves262k": lambda: [W.tilted_curves()],
             "batch64": W.batched_uneven}[args.workload]()
    tol = 1e-5
    log = lambda s: print(s, flush=True)
    net = WS.WarmStartNet.load(WS.DEFAULT_MODEL) if args.start == "learned" else None
    # verbose: the solver prints a progress line every 2 s of a solve
    kw = dict(history=10, tol=tol, stop="maxrel", h0_threshold=0.0, max_iter=20000,
              device_loop=loop, verbose=1, newton_from=args.newton_from or None,
              newton_coarse=args.newton_coarse, newton_sweeps=args.newton_sweeps)

    ms_levels = tuple(int(v) for v in args.levels.split(",")) if args.levels else wl["levels"]
    levels = {"multiscale": ms_levels, "learned": (ms_levels[-1],), "zero": ()}[args.start]
    coarse = WS.coarse_problems(probs, levels)
    tuning = {"final": None, "levels": [None] * len(levels)}

    def start_weights(before_level=None):
        """The warm start's weights for the full-resolution problems and its evaluations;
        before_level(li, problems, init): called with each coarse level's starting weights
        (None: zeros) before that level is solved."""
        if args.start == "zero":
            return None, 0
        fields = None if args.start == "multiscale" else [
            WS.LEARNED_ALPHA * WS.smooth_field(U, WS.LEARNED_SIGMA)
            for U in WS.predict_fields(net, probs)]
        evals = 0
        for li, (n_side, cps) in enumerate(zip(levels, coarse)):
            init = None if fields is None else np.concatenate(
                [WS.weights_from_field(U, cp.sites, cp.frame) for U, cp in zip(fields, cps)])
            if before_level is not None:
                before_level(li, cps, init)
            tl = time.perf_counter()
            with OTBatch(cps, precision=args.coarse_precision,
                         tuning=tuning["levels"][li]) as cb:
                wc, rc = cb.solve(init, **dict(kw, tol=args.coarse_tol))
                cen, _ = cb.centroids(wc)
                wcs, ccs = cb.split(wc), cb.split(cen.reshape(-1, 2))
            evals += sum(r["evaluations"] for r in rc)
            log(f"[level {n_side}] {1e3 * (time.perf_counter() - tl):.1f} ms, iterations "
                f"{max(r['iterations'] for r in rc)} (Newton "
                f"{max(r.get('newton_iterations', 0) for r in rc)}, L-BFGS "
                f"{1e3 * max(r.get('lbfgs_seconds', 0.0) for r in rc):.1f} ms, Newton "
                f"{1e3 * max(r.get('newton_seconds', 0.0) for r in rc):.1f} ms)")
            fields = [WS.field_from_solution(cp.sites, wk, ck, n_side, cp.frame)
                      for cp, wk, ck in zip(cps, wcs, ccs)]
        return np.concatenate([WS.weights_from_field(U, p.sites, p.frame)
                               for U, p in zip(fields, probs)]), evals

    def one_solve(batch):
        """One timed step.  Returns (reports of the full-resolution solves, evaluations)."""
        w0, evals = start_weights()
        tl = time.perf_counter()
        w, rep = batch.solve(w0, **kw)
        log(f"[level final] {1e3 * (time.perf_counter() - tl):.1f} ms, iterations "
            f"{max(r['iterations'] for r in rep)} (Newton "
            f"{max(r.get('newton_iterations', 0) for r in rep)}, L-BFGS "
            f"{1e3 * max(r.get('lbfgs_seconds', 0.0) for r in rep):.1f} ms, Newton "
            f"{1e3 * max(r.get('newton_seconds', 0.0) for r in rep):.1f} ms)")
        return rep, evals + sum(r["evaluations"] for r in rep)

    tune_report = None
    if not args.no_tune:
        # launch shapes per level, outside the timed region, level by level: each level is
        # tuned from the weights the already tuned levels below hand it, then solved with its
        # shapes -- the solves are deterministic, so the tuner measures exactly the trajectory
        # the timed solves will follow
        from caustic_amd import autotune as AT

        tune_report = {}
        # the measured stretch: each level's L-BFGS phase up to the Newton hand-off (or 40
        # iterations of a pure L-BFGS solve)
        stretch = (dict(iters=400, stop_tol=args.newton_from) if args.newton_from
                   else dict(iters=40))

        def tune_level(li, cps, init):
            tuning["levels"][li], r = AT.tune(cps, args.coarse_precision, w0=init, log=log,
                                              **stretch)
            tune_report[f"level{levels[li]}"] = r

        w_probe, _ = start_weights(before_level=tune_level)
        tuning["final"], r = AT.tune(probs, "mixed", w0=w_probe, log=log, **stretch)
        tune_report["final"] = r
        log(f"[autotune] {json.dumps(tune_report)}")

    batch = OTBatch(probs, precision="mixed", tuning=tuning["final"])
    cold = None
    if not args.no_cold:
        t0 = time.perf_counter()
        _, rc = batch.solve(None, **kw)
        cold = {"seconds": time.perf_counter() - t0,
                "iterations": max(r["iterations"] for r in rc),
                "evaluations": sum(r["evaluations"] for r in rc),
                "converged": all(bool(r["converged"]) for r in rc)}
        log(f"[plain zero-start solve] {cold}")
    for _ in range(args.warmup):
        r, e = one_solve(batch)
        log(f"[warmup] iters {max(x['iterations'] for x in r)} evals {e} "
            f"maxrel {max(x['max_rel_err'] for x in r):.3e}")
    torch.cuda.synchronize()
    batch.reset_stats()
    reps = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        r, e = one_solve(batch)
        reps.append((r, e, time.perf_counter() - ts))
        log(f"[step {k}] iters {max(x['iterations'] for x in r)} evals {e} "
            f"maxrel {max(x['max_rel_err'] for x in r):.3e} {reps[-1][2]:.3f}s")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = batch.stats()
    conv = all(x["converged"] for r, _, _ in reps for x in r)
    # dominant work of an evaluation: construction + integration passes of the full batch
    cell_avg = st["cell_seconds"] / max(st["cell_launches"], 1)
    integ_avg = st["integ_seconds"] / max(st["integ_launches"], 1)
    P = sum(p.rho.size for p in probs)
    N = sum(p.n for p in probs)
    alg_bytes = P * 4 + 48 * N  # FP32 density read once; sites, w, nu, mass, f
    achieved = alg_bytes / (cell_avg + integ_avg) / 1e9
    n_solves = len(probs) * args.steps
    out = {
        "metric": wl["metric"],
        "value": n_solves / dt if conv else 0.0,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded jittered grids + disks/rectangles or Gaussian-mixture targets)",
        "config": {
            "workload": wl["desc"],
            "start": args.start, "problems": len(probs), "n_sites": N,
            "target_pixels": P,
            "final_iterations_max": int(np.mean([max(x["iterations"] for x in r)
                                                 for r, _, _ in reps])),
            "evaluations_all_levels": int(np.mean([e for _, e, _ in reps])),
            "converged": conv,
            "max_rel_err": max(x["max_rel_err"] for r, _, _ in reps for x in r),
            "device_loop": args.device_loop,
            "newton_from": args.newton_from or None,
            "newton_coarse": args.newton_coarse,
            "newton_sweeps": args.newton_sweeps,
            "coarse_precision": args.coarse_precision,
            "final_level_newton_per_solve": {
                "steps": st["newton_steps"] / args.steps,
                "cg_iterations": st["cg_iterations"] / args.steps,
                "coarse_inverse_host_ms": 1e3 * st["coarse_seconds"] / args.steps},
            "newton_iterations": int(np.mean([max(x.get("newton_iterations", 0) for x in r)
                                              for r, _, _ in reps])),
            "launch_shapes": tuning,
        },
        "plain_zero_start": cold,
        "speedup_vs_zero_start": (cold["seconds"] / (dt / args.steps)) if cold else None,
        "roofline": {
            "kernel": "cell evaluation: cell_kernel (construction) + integrate_kernel (exact "
                      "mass integral), full-batch passes of the timed solves",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_BYTES.get(args.workload),
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
    if not args.no_cpu_baseline and args.workload == "shapes65k":
        n_eval = cold["evaluations"] if cold else out["config"]["evaluations_all_levels"]
        out["cpu_baseline"] = cpu_baseline(probs[0], n_eval, args.cpu_budget)
    if args.learned_seeds > 0 and os.path.exists(WS.DEFAULT_MODEL):
        # the metric's second part: the learned warm start against the zero start on held-out
        # N = 16,384 problems (128 x 128 jittered sites, 512 x 512 shapes targets)
        rows = WS.evaluate_warm_start(WS.WarmStartNet.load(WS.DEFAULT_MODEL),
                                      range(900000, 900000 + args.learned_seeds), log=log)
        ci, wi = sum(r["cold_iters"] for r in rows), sum(r["warm_iters"] for r in rows)
        cs, ws_ = sum(r["cold_s"] for r in rows), sum(r["warm_s"] for r in rows)
        out["learned_vs_zero"] = {
            "workload": "N=16,384 (128x128 jittered grid), 512x512 shapes target, mixed "
                        "precision, L-BFGS to max-rel 1 then Newton to 1e-5",
            "held_out_seeds": [r["seed"] for r in rows],
            "zero_iterations": ci, "learned_iterations": wi, "iteration_ratio": ci / max(wi, 1),
            "zero_seconds": cs, "learned_seconds": ws_, "speedup": cs / ws_,
            "learned_used": [r["learned_used"] for r in rows],
            "converged": all(r["converged"] for r in rows)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
