"""Autotuner of the evaluation kernels' launch shapes, measured on the problems to be solved.

Knobs (fields of cot_options):

  int_lanes         integration lanes per site (4, 8, 16).  Lane q of a site walks edges q,
                    q + L, ... across the pixel lattice, so more lanes mean shorter serial walks
                    and more waves in flight, at the price of idle lanes on cells with few edges
                    and a longer shuffle reduction.  Large cells (coarse sites over a fine image)
                    favour more lanes.
  repair_blocks     grid of the repair passes that walk the work list of flagged sites
                    (128 ... 1024 blocks: short lists want few blocks, long ones one per CU or
                    more).
  sites_per_bucket  not tuned (kept at 1).  Sites are stored in 8x8-bucket tiles, so this sets
                    the shape of the patch of plane under one 64-site block: one site per bucket
                    makes a block one square tile, two make it half a tile, 2:1.  The Newton
                    steps' preconditioner works on those blocks (block-local Jacobi sweeps, patches
                    of whole blocks), and the elongated blocks doubled the CG iterations of the
                    N=65,536 shapes solve (583 -> 1,176, 195 -> 212 ms a solve) while the
                    evaluations it measures here ran 8% faster.
  full_pass_ring    not tuned (kept at 2): candidates of the FP32 full pass, the previous cells'
                    one-ring (1) or two-ring (2).  The two-ring costs more clipping per cell and
                    leaves fewer unmatched edges to the repair passes.  On the tuner's stretches
                    the one-ring ranked within a few percent of the two-ring at N=4,096 and
                    N=16,384, but in the coarse-to-fine solve's own levels (the start a coarser
                    level hands over moves the cells further) it failed the conformity check at
                    1,000 of 4,096 and 3,500 of 16,384 sites per evaluation against 55 and 150
                    with the two-ring, and made those levels 13-24% slower.
  pass_mode         construction of the full passes: one lane per cell (0) or one wave per cell
                    (1, candidates loaded 64 at a time, cooperative clipping).  The same cells
                    either way; one wave per cell puts many more waves in flight and wins on
                    small problems (a few thousand cells), where one lane per cell leaves most
                    of the chip idle.

Objective: elapsed time per full evaluation over a stretch of the real solve -- every kernel of
the evaluation (construction, integration, conformity checks, repair rounds, adjacency saves)
and of the L-BFGS steps around it, with the gaps between them.  The in-kernel timers
(cot_get_stats) see only the construction, integration and repair kernels, and only while they
have work: ranked by them, the one-ring full pass won at N=16,384 (cheaper construction) while
the conformity checks and repair rounds it caused -- 21% of the sites failing per evaluation,
each round's check over every site -- made that level of the coarse-to-fine solve 18% slower
than with the two-ring.  The stretch: `iters` L-BFGS
iterations from the caller's weights, or with `stop_tol` the whole L-BFGS phase until the Newton
hand-off (max-rel stop_tol, no empty cell; a 40-iteration stretch from the start ranked two
bucket densities the other way round from the full phase).  Per evaluation, not per phase: the
phase's evaluation count depends on the trajectory, which a shape changes chaotically through
rounding (ranking by whole-phase time picked shapes by luck and made the next level's Newton
phase 3x longer).  Each candidate is replayed, the smaller of
`repeats` measurements (the repair passes' work lists fill in a different order from run to run,
which moves sites between blocks; the resulting spread is one-sided, up to ~1.8x slower in rare
runs, and one such outlier is enough to make the descent reject the right shape).  Coordinate
descent over the knobs (a change must win by 5%), at most two sweeps.  Results are cached per device,
precision and problem shape in a JSON file ($CAUSTIC_TUNE_CACHE, default
~/.cache/caustic_amd/tuning.json).
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

KNOBS = {
    "int_lanes": (4, 8, 16),
    "repair_blocks": (128, 256, 512, 1024),
    "pass_mode": (0, 1),
}
DEFAULT = {"int_lanes": 8, "repair_blocks": 256, "sites_per_bucket": 1.0, "full_pass_ring": 2,
           "pass_mode": 0}
# a change must win by 5%: shapes within a few percent of each other (8 or 16 integration lanes,
# 256 or 512 repair blocks) trade places from run to run, and each flip sends the solve down a
# different (equally fast on average) trajectory, so near-ties keep the incumbent
MIN_GAIN = 0.05


def cache_path() -> str:
    return os.environ.get("CAUSTIC_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "caustic_amd",
                                       "tuning.json"))


def device_key() -> str:
    """The GPU's architecture and compute-unit count (the marketing name is not reported
    reliably under every tool)."""
    import torch

    if not torch.cuda.is_available():
        return "none"
    pr = torch.cuda.get_device_properties(0)
    arch = (getattr(pr, "gcnArchName", "") or "gpu").split(":")[0]
    return f"{arch}-{pr.multi_processor_count}cu"


def shape_key(problems, precision: str) -> str:
    dev = device_key()
    n = sum(int(p.n) for p in problems)
    px = sum(int(p.rho.size) for p in problems)
    return f"{dev}|{precision}|problems={len(problems)}|sites={n}|pixels={px}"


def _load_cache() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key: str, entry: dict) -> None:
    data = _load_cache()
    data[key] = entry
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a re-tune


def measure(problems, precision: str, cfg: dict, w0, iters: int, solve_kw: dict) -> dict:
    """Seconds per full evaluation (elapsed) of `iters` L-BFGS iterations from w0 under cfg,
    with the in-kernel timers' share of it."""
    from .ot import OTBatch

    kw = dict(solve_kw, raise_on_fail=False)
    with OTBatch(problems, precision=precision, tuning=cfg) as b:
        b.solve(w0, **dict(kw, max_iter=2))  # adjacency of the start, graph capture
        b.reset_stats()
        t0 = time.perf_counter()
        b.solve(w0, **dict(kw, max_iter=iters))
        wall = time.perf_counter() - t0
        st = b.stats()
    evals = max(int(st["cell_launches"]), 1)
    dev = st["cell_seconds"] + st["integ_seconds"] + st["repair_seconds"]
    return {"per_eval": wall / evals, "total": wall, "evaluations": evals, "wall": wall,
            "timed_kernels": dev / evals, "cell": st["cell_seconds"] / evals,
            "integrate": st["integ_seconds"] / evals, "repair": st["repair_seconds"] / evals}


def tune(problems, precision: str = "mixed", w0=None, iters: int = 40, solve_kw: dict = None,
         use_cache: bool = True, log=None, repeats: int = 2, stop_tol: float = None):
    """Launch shapes for `problems` (a list of workloads.Problem).  w0: weights to start the
    measured iterations from (zeros by default; a warm start is more representative of a real
    solve).  stop_tol: measure L-BFGS from w0 until max-rel stop_tol (at most `iters`
    iterations) instead of exactly `iters` iterations.  Returns (config dict for
    OTBatch(tuning=...), report dict)."""
    key = shape_key(problems, precision) + (f"|to={stop_tol:g}" if stop_tol else f"|it={iters}")
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            return dict(hit["config"]), dict(hit, cached=True)
    solve_kw = dict(solve_kw or {})
    solve_kw.setdefault("history", 10)
    solve_kw.setdefault("tol", stop_tol if stop_tol else 1e-12)  # 1e-12: the full stretch
    solve_kw.setdefault("stop", "handoff" if stop_tol else "maxrel")
    solve_kw.setdefault("h0_threshold", 0.0)
    n = sum(int(p.n) for p in problems)
    w0 = np.zeros(n) if w0 is None else np.asarray(w0, dtype=np.float64)
    seen = {}

    # objective: elapsed time per evaluation (a property of the shape; the phase's evaluation
    # count is a property of the trajectory, which the shape changes chaotically through
    # rounding, and does not carry over to other inputs of the same size)
    obj = "per_eval"

    def run(cfg):
        k = tuple(sorted(cfg.items()))
        if k not in seen:
            ms = [measure(problems, precision, cfg, w0, iters, solve_kw)
                  for _ in range(max(int(repeats), 1))]
            seen[k] = min(ms, key=lambda m: m.get(obj, m["per_eval"]))
            if log:
                log(f"[autotune] {cfg} -> "
                    + " / ".join(f"{m.get(obj, m['per_eval']) * 1e6:.1f}" for m in ms)
                    + " us/eval"
                    + f" ({seen[k]['evaluations']} evaluations)")
        return seen[k].get(obj, seen[k]["per_eval"])

    t_start = time.perf_counter()
    cfg = dict(DEFAULT)
    best = run(cfg)
    for _ in range(2):
        changed = False
        for knob, values in KNOBS.items():
            for v in values:
                if v == cfg[knob]:
                    continue
                trial = dict(cfg, **{knob: v})
                t = run(trial)
                if t < best * (1.0 - MIN_GAIN):
                    best, cfg, changed = t, trial, True
        if not changed:
            break
    base = seen[tuple(sorted(DEFAULT.items()))].get(obj, seen[tuple(sorted(DEFAULT.items()))]["per_eval"])
    report = {
        "config": cfg,
        "objective": obj,
        "per_eval_us": seen[tuple(sorted(cfg.items()))]["per_eval"] * 1e6,
        "objective_us": best * 1e6,
        "default_objective_us": base * 1e6,
        "gain_vs_default": base / best if best > 0 else 1.0,
        "candidates": len(seen),
        "seconds": time.perf_counter() - t_start,
        "cached": False,
    }
    if use_cache:
        _store_cache(key, report)
    return dict(cfg), report
