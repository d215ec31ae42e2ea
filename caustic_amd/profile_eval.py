"""Profiling driver for the evaluation kernels: solve a workload close to its solution, then
evaluate it repeatedly at nearby weights, so that a profiler sees the construction, integration
and conformity-check kernels of a late-phase solve in isolation.

    python -m caustic_amd.profile_eval [--workload shapes65k|curves262k] [--evals 50]

Prints the in-kernel timers' average launch durations (construction, integration) and the
algorithmic bytes of one evaluation (FP32 density once, 48 bytes per site).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="shapes65k", choices=["shapes65k", "curves262k"])
    ap.add_argument("--evals", type=int, default=50)
    ap.add_argument("--spread", type=float, default=1e-2,
                    help="perturbation of the solved weights (normal, per site), in units of the "
                         "mean cell area -- a late-phase step moves edges by a small fraction of "
                         "the site spacing")
    args = ap.parse_args(argv)

    import torch

    if not torch.cuda.is_available():
        raise SystemExit("no GPU visible")
    from . import warmstart as WS
    from . import workloads as W
    from .ot import OTBatch

    prob = W.large_shapes() if args.workload == "shapes65k" else W.tilted_curves()
    levels = (16, 32, 64, 128) if args.workload == "shapes65k" else (16, 32, 64, 128, 256)
    t0 = time.perf_counter()
    ws, _, rep = WS.multiscale_solve([prob], levels=levels, tol=1e-5, newton_from=1.0,
                                     log=lambda s: print(s, flush=True))
    print(f"[profile_eval] solved in {time.perf_counter() - t0:.2f}s: {rep[0]['max_rel_err']:.2e}",
          flush=True)
    w = ws[0]
    rng = np.random.default_rng(0)
    x0, x1, y0, y1 = prob.proj_domain
    scale = args.spread * (x1 - x0) * (y1 - y0) / prob.n
    with OTBatch([prob], precision="mixed") as b:
        b.evaluate(w)
        b.reset_stats()
        t0 = time.perf_counter()
        for _ in range(args.evals):
            b.evaluate(w + rng.normal(0.0, scale, w.shape[0]), allow_unverified=True)
        wall = time.perf_counter() - t0
        st = b.stats()
    n = max(st["cell_launches"], 1)
    out = {
        "workload": args.workload, "evaluations": args.evals, "wall_s": wall,
        "cell_kernel_us": 1e6 * st["cell_seconds"] / n,
        "integrate_kernel_us": 1e6 * st["integ_seconds"] / max(st["integ_launches"], 1),
        "repair_s": st["repair_seconds"], "repair_launches": st["repair_launches"],
        "alg_bytes_per_eval": int(prob.rho.size * 4 + 48 * prob.n),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
