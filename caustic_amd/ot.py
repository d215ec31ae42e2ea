"""Python host side of the MI355X OT solver: a batch of problems resident on the GPU.

    batch = OTBatch([problem], precision="fp64")
    w, reports = batch.solve(history=10, tol=1e-5)

Each problem is given in the reference's conventions (sites = ``proj_xy``, nu = carry energy,
rho = normalised picture in the layout rho[iy, ix], proj_domain = (x0, x1, y0, y1)); see
``caustic_amd.workloads.Problem``.  ``direct_ot`` mirrors the reference operator of the same
name (Optimal_Transport_rl/ot_op.cuh:66-70) for the parallel-light, t = 0 path.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib as L

_dp = ctypes.POINTER(ctypes.c_double)
_up = ctypes.POINTER(ctypes.c_ubyte)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _problem_array(problems):
    """(cot_problem array, the numpy arrays it points into) for workloads.Problem-like objects."""
    arr = (L.cot_problem * len(problems))()
    keep = []
    for k, p in enumerate(problems):
        sites = np.ascontiguousarray(p.sites, dtype=np.float64).reshape(-1, 2)
        nu = np.ascontiguousarray(p.nu, dtype=np.float64)
        rho = np.ascontiguousarray(p.rho, dtype=np.float64)
        if nu.shape[0] != sites.shape[0]:
            raise ValueError("nu and sites differ in length")
        keep += [sites, nu, rho]
        arr[k].n_sites = sites.shape[0]
        arr[k].sites = _ptr(sites)
        arr[k].nu = _ptr(nu)
        arr[k].rho = _ptr(rho)
        arr[k].pn, arr[k].pm = rho.shape
        for q in range(4):
            arr[k].proj_domain[q] = float(p.proj_domain[q])
    return arr, keep


def ladder_defaults(problems):
    """The library's default coarse-to-fine schedule for these problems
    (cot_default_ladder_opts): {"sides", "floors", "final_floors", "coarse_tol", "handoff"}."""
    arr, keep = _problem_array(list(problems))
    lo = L.cot_ladder_opts()
    L.load().cot_default_ladder_opts(arr, len(problems), ctypes.byref(lo))
    del keep
    return {"sides": [lo.sides[k] for k in range(lo.n_levels)],
            "floors": [lo.floors[k] for k in range(lo.n_levels)],
            "final_floors": [lo.final_floors[k] for k in range(lo.n_final_floors)],
            "coarse_tol": lo.coarse_tol, "handoff": lo.handoff}


class OTBatch:
    """A context of one or more OT problems on the current GPU."""

    def __init__(self, problems: Sequence, precision: str = "fp64", sites_per_bucket: float = 1.0,
                 ring_init: int = 2, ring_max: int = 32, verify_tol: float = 1e-4,
                 max_repairs: int = 12, int_lanes: int = 8, repair_blocks: int = 256,
                 full_pass_ring: int = 0, tuning: dict = None, repair_mode: str = "wave",
                 pass_mode: int = 0):
        """tuning: launch-shape options as chosen by caustic_amd.autotune (int_lanes,
        repair_blocks, sites_per_bucket, full_pass_ring, pass_mode); they override the keyword
        arguments.  repair_mode: "wave" (one wave per repaired cell) or "lane" (one lane per
        cell); pass_mode: the full passes' construction, 0 one lane per cell, 1 one wave per
        cell.  Either choice builds the same cells."""
        if tuning:
            int_lanes = int(tuning.get("int_lanes", int_lanes))
            repair_blocks = int(tuning.get("repair_blocks", repair_blocks))
            sites_per_bucket = float(tuning.get("sites_per_bucket", sites_per_bucket))
            full_pass_ring = int(tuning.get("full_pass_ring", full_pass_ring))
            pass_mode = int(tuning.get("pass_mode", pass_mode))
        lib = L.load()
        self.problems = list(problems)
        arr, self._keep = _problem_array(self.problems)
        self.sizes = [int(arr[k].n_sites) for k in range(len(self.problems))]
        opt = L.cot_options()
        lib.cot_default_options(ctypes.byref(opt))
        opt.precision = {"fp64": L.COT_PREC_FP64, "mixed": L.COT_PREC_MIXED}[precision]
        opt.sites_per_bucket = sites_per_bucket
        opt.ring_init = ring_init
        opt.ring_max = ring_max
        opt.verify_tol = verify_tol
        opt.max_repairs = max_repairs
        opt.int_lanes = int_lanes
        opt.repair_blocks = repair_blocks
        opt.full_pass_ring = full_pass_ring
        opt.repair_mode = {"wave": 0, "lane": 1}[repair_mode]
        opt.pass_mode = int(pass_mode)
        ctx = ctypes.c_void_p()
        L.check(lib.cot_create(arr, len(problems), ctypes.byref(opt), ctypes.byref(ctx)))
        self._keep = []  # the library copied everything
        self._opt = opt
        self._ctx = ctx
        self._lib = lib
        self.precision = precision
        self.tuning = {"int_lanes": int_lanes, "repair_blocks": repair_blocks,
                       "sites_per_bucket": sites_per_bucket, "full_pass_ring": full_pass_ring,
                       "pass_mode": int(pass_mode)}
        self._img_shapes = [tuple(np.shape(p.rho)) for p in problems]
        self.n_total = int(sum(self.sizes))
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n_problems = len(problems)

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.cot_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _w(self, w):
        if w is None:
            return np.zeros(self.n_total)
        w = np.ascontiguousarray(np.asarray(w, dtype=np.float64).ravel())
        if w.shape[0] != self.n_total:
            raise ValueError(f"expected {self.n_total} weights, got {w.shape[0]}")
        return w

    def split(self, v):
        return [v[self.offsets[k]:self.offsets[k + 1]] for k in range(self.n_problems)]

    def evaluate(self, w=None, allow_unverified: bool = False):
        """Masses, per-problem objective, rho-weighted centroids and existence flags at w.
        allow_unverified: return the cells as built (with "verified": False) when the
        conformity check still fails after the repair rounds, instead of raising."""
        w = self._w(w)
        mass = np.zeros(self.n_total)
        f = np.zeros(self.n_problems)
        cen = np.zeros((self.n_total, 2))
        ex = np.zeros(self.n_total, dtype=np.uint8)
        rc = self._lib.cot_evaluate(self._ctx, _ptr(w), _ptr(mass), _ptr(f), _ptr(cen),
                                    ex.ctypes.data_as(_up))
        L.check(rc, allow=(L.COT_EUNVERIFIED,) if allow_unverified else ())
        return {"mass": mass, "f": f, "centroid": cen, "exist": ex.astype(bool),
                "verified": rc == L.COT_OK}

    # ---- torch tensors on the GPU (cot_*_dev: device arrays, the caller's current stream) ----

    def _tensor_w(self, w):
        import torch

        if w is None:
            return torch.zeros(self.n_total, dtype=torch.float64, device="cuda")
        if not (isinstance(w, torch.Tensor) and w.is_cuda):
            raise TypeError("expected a CUDA (HIP) tensor of weights")
        w = w.detach().reshape(-1).to(torch.float64).contiguous()
        if w.numel() != self.n_total:
            raise ValueError(f"expected {self.n_total} weights, got {w.numel()}")
        return w

    @staticmethod
    def _stream(t):
        import torch

        return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)

    def evaluate_tensor(self, w=None):
        """evaluate() on GPU tensors: masses (N,), per-problem f (P,), rho-weighted centroids
        (N, 2) float64 and existence (N,) uint8, all on w's device, ordered on the current
        stream (cot_evaluate_dev)."""
        import torch

        w = self._tensor_w(w)
        kw = dict(dtype=torch.float64, device=w.device)
        mass, f = torch.empty(self.n_total, **kw), torch.empty(self.n_problems, **kw)
        cen = torch.empty((self.n_total, 2), **kw)
        ex = torch.empty(self.n_total, dtype=torch.uint8, device=w.device)
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        L.check(self._lib.cot_evaluate_dev(self._ctx, p(w), p(mass), p(f), p(cen), p(ex),
                                           self._stream(w)))
        return {"mass": mass, "f": f, "centroid": cen, "exist": ex}

    def centroids_tensor(self, w=None, weighted: bool = True):
        """centroids() on GPU tensors (cot_centroids_dev): ((N, 2) float64, (N,) uint8)."""
        import torch

        w = self._tensor_w(w)
        cen = torch.empty((self.n_total, 2), dtype=torch.float64, device=w.device)
        ex = torch.empty(self.n_total, dtype=torch.uint8, device=w.device)
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        L.check(self._lib.cot_centroids_dev(self._ctx, p(w), int(weighted), p(cen), p(ex),
                                            self._stream(w)))
        return cen, ex

    def solve_tensor(self, w0=None, history: int = 10, max_iter: int = 50001, tol: float = 1e-5,
                     initial_step: float = 1.0, stop: str = "maxrel", h0_threshold: float = 1e-5,
                     accept_rule: str = "auto", device_loop=True, raise_on_fail: bool = True):
        """L-BFGS from the tensor w0 (zeros: None) on the GPU (cot_solve_dev); returns (a new
        float64 tensor of weights, reports)."""
        import torch

        w = self._tensor_w(w0).clone()
        o = self._solve_opts(history, max_iter, tol, initial_step, stop, h0_threshold, 0,
                             device_loop, accept_rule, 3)
        reps = (L.cot_report * self.n_problems)()
        rc = self._lib.cot_solve_dev(self._ctx, ctypes.c_void_p(w.data_ptr()), ctypes.byref(o),
                                     reps, self._stream(w))
        if raise_on_fail:
            L.check(rc, allow=(L.COT_ENOTCONV,))
        return w, [{k: getattr(r, k) for k, _ in L.cot_report._fields_} for r in reps]

    def _solve_opts(self, history, max_iter, tol, initial_step, stop, h0_threshold, verbose,
                    device_loop, accept_rule, repair_rounds):
        if accept_rule == "auto":
            accept_rule = "derivative" if self.precision == "mixed" else "decrease"
        o = L.cot_solve_opts()
        self._lib.cot_default_solve_opts(ctypes.byref(o))
        o.history = history
        o.max_iter = max_iter
        o.tol = tol
        o.initial_step = initial_step
        o.stop_rule = {"reference": L.COT_STOP_REFERENCE, "maxrel": L.COT_STOP_MAXREL,
                       "handoff": L.COT_STOP_HANDOFF}[stop]
        o.h0_threshold = h0_threshold
        o.verbose = verbose
        # True/1: graph-replayed device rounds; 2: the same rounds launched directly; False/0:
        # host-driven rounds
        o.device_loop = int(device_loop)
        o.accept_rule = {"decrease": 0, "derivative": 1}[accept_rule]
        o.repair_rounds = repair_rounds
        return o

    def centroids(self, w=None, weighted: bool = True):
        w = self._w(w)
        cen = np.zeros((self.n_total, 2))
        ex = np.zeros(self.n_total, dtype=np.uint8)
        L.check(self._lib.cot_centroids(self._ctx, _ptr(w), int(weighted), _ptr(cen),
                                        ex.ctypes.data_as(_up)))
        return cen, ex.astype(bool)

    def render(self, w=None):
        """Per-problem images (H, W): each site's source mass spread uniformly over its cell
        (the reference's diagnostic render)."""
        w = self._w(w)
        total = sum(int(np.prod(sh)) for sh in self._img_shapes)
        img = np.zeros(total)
        L.check(self._lib.cot_render(self._ctx, _ptr(w), _ptr(img)))
        out, off = [], 0
        for sh in self._img_shapes:
            n = int(np.prod(sh))
            out.append(img[off:off + n].reshape(sh))
            off += n
        return out

    def cell_polygons(self, w=None):
        """(counts (N,), vertices (N, COT_MAX_VERTS, 2)) of the cells at w, CCW; count 0 for an
        empty cell."""
        w = self._w(w)
        counts = np.zeros(self.n_total, dtype=np.int32)
        verts = np.zeros((self.n_total, L.COT_MAX_VERTS, 2))
        L.check(self._lib.cot_cell_polygons(self._ctx, _ptr(w),
                                            counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                            _ptr(verts)))
        return counts, verts

    def hessian_product(self, w, x):
        """H x with H = d mass / d w at w (the Newton steps' matrix)."""
        w = self._w(w)
        x = self._w(x)
        y = np.zeros(self.n_total)
        L.check(self._lib.cot_hessian_product(self._ctx, _ptr(w), _ptr(x), _ptr(y)))
        return y

    def cell_areas(self, w=None):
        w = self._w(w)
        a = np.zeros(self.n_total)
        L.check(self._lib.cot_cell_areas(self._ctx, _ptr(w), _ptr(a)))
        return a

    def cell_solid_angles(self, light, z: float, w=None):
        """Solid angle under which a point light at light (x, y, z) sees each cell placed in
        the plane z (cot_cell_solid_angles; solid_angle, basic_function.cuh:635-651)."""
        w = self._w(w)
        lt = np.ascontiguousarray(light, dtype=np.float64).reshape(3)
        om = np.zeros(self.n_total)
        L.check(self._lib.cot_cell_solid_angles(self._ctx, _ptr(w), _ptr(lt), float(z),
                                                _ptr(om)))
        return om

    def stats(self):
        """Kernel time and counts accumulated since the last reset_stats()."""
        s = L.cot_stats()
        L.check(self._lib.cot_get_stats(self._ctx, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in L.cot_stats._fields_}

    def reset_stats(self):
        L.check(self._lib.cot_reset_stats(self._ctx))

    def solve(self, w0=None, history: int = 10, max_iter: int = 50001, tol: float = 1e-5,
              initial_step: float = 1.0, stop: str = "maxrel", h0_threshold: float = 1e-5,
              verbose: int = 0, raise_on_fail: bool = True, device_loop=True,
              accept_rule: str = "auto", repair_rounds: int = 3, newton_from: float = None,
              newton_max_iter: int = 50, newton_sweeps: int = 2, newton_coarse: int = 256,
              newton_small_cg: bool = True):
        """L-BFGS dual solve from w0 (zeros by default).  Returns (w, reports).

        accept_rule: "decrease" (f_trial < f, the reference rule), "derivative" (also accept a
        trial whose directional derivative is <= 0 -- for FP32 geometry, whose f is too coarse
        for the last decades), or "auto" (derivative for mixed precision, decrease for fp64).
        newton_from (stop="maxrel" only): L-BFGS until max-rel newton_from with no empty cell,
        then damped Newton steps (cot_solve_newton) to tol; problems the Newton steps leave unconverged finish
        with L-BFGS.  Reports then add "newton_iterations".  newton_sweeps: block-local
        Jacobi sweeps of the conjugate gradients' preconditioner (0: plain Jacobi);
        newton_coarse: at most this many patches per problem in its two-level coarse term
        (0: none); newton_small_cg: solve a system of at most 8 blocks in one workgroup."""
        kw = dict(history=history, max_iter=max_iter, initial_step=initial_step, stop=stop,
                  h0_threshold=h0_threshold, verbose=verbose, device_loop=device_loop,
                  accept_rule=accept_rule, repair_rounds=repair_rounds)
        if newton_from is not None and stop == "maxrel" and newton_from > tol:
            # L-BFGS until max-rel newton_from with no empty cell (Newton's starting condition)
            w1, r1 = self.solve(w0, tol=newton_from, raise_on_fail=False,
                                **dict(kw, stop="handoff"))
            w2, r2 = self._newton(w1, tol, newton_max_iter, verbose, repair_rounds, device_loop,
                                  newton_sweeps, newton_coarse, newton_small_cg)
            if all(r["converged"] for r in r2):
                w3, r3 = w2, None
            else:
                w3, r3 = self.solve(w2, tol=tol, raise_on_fail=raise_on_fail, **kw)
            out = []
            for k in range(self.n_problems):
                last = r3[k] if r3 is not None else r2[k]
                rep = dict(last)
                rep["iterations"] = r1[k]["iterations"] + r2[k]["iterations"] + (
                    r3[k]["iterations"] if r3 is not None else 0)
                rep["evaluations"] = r1[k]["evaluations"] + r2[k]["evaluations"] + (
                    r3[k]["evaluations"] if r3 is not None else 0)
                rep["newton_iterations"] = r2[k]["iterations"]
                rep["lbfgs_seconds"] = r1[k]["seconds"]
                rep["newton_seconds"] = r2[k]["seconds"]
                rep["seconds"] = r1[k]["seconds"] + r2[k]["seconds"] + (
                    r3[k]["seconds"] if r3 is not None else 0.0)
                out.append(rep)
            if raise_on_fail and not all(r["converged"] for r in out):
                raise RuntimeError("solve did not converge")
            return w3, out
        w = self._w(w0).copy()
        o = self._solve_opts(history, max_iter, tol, initial_step, stop, h0_threshold, verbose,
                             device_loop, accept_rule, repair_rounds)
        reps = (L.cot_report * self.n_problems)()
        rc = self._lib.cot_solve(self._ctx, _ptr(w), ctypes.byref(o), reps)
        if raise_on_fail:
            L.check(rc, allow=(L.COT_ENOTCONV,))
        out = [{k: getattr(r, k) for k, _ in L.cot_report._fields_} for r in reps]
        return w, out

    def ladder_options(self, frames=None, level_tunings=None, **overrides):
        """cot_ladder_opts for this batch: the library's defaults (cot_default_ladder_opts),
        then overrides (n_levels/sides/floors/final_floors as sequences, coarse_tol, handoff,
        seed).  frames: per-problem (x0, x1, y0, y1) regions the sites sample (default: each
        problem's frame).  level_tunings: launch shapes per coarse level (autotune dicts)."""
        arr, keep = _problem_array(self.problems)
        lo = L.cot_ladder_opts()
        self._lib.cot_default_ladder_opts(arr, len(self.problems), ctypes.byref(lo))
        if "sides" in overrides:
            sides = list(overrides.pop("sides"))
            lo.n_levels = len(sides)
            floors = list(overrides.pop("floors", [0.0] * len(sides)))
            for k, (sd, fl) in enumerate(zip(sides, floors)):
                lo.sides[k], lo.floors[k] = int(sd), float(fl)
        if "final_floors" in overrides:
            ff = list(overrides.pop("final_floors"))
            lo.n_final_floors = len(ff)
            for k, v in enumerate(ff):
                lo.final_floors[k] = float(v)
        for k, v in overrides.items():
            setattr(lo, k, v)
        fr = np.ascontiguousarray([p.frame if frames is None else frames[k]
                                   for k, p in enumerate(self.problems)], dtype=np.float64)
        keep.append(fr)
        lo.frames = _ptr(fr)
        if level_tunings:
            lvl = (L.cot_options * lo.n_levels)()
            for k in range(lo.n_levels):
                ctypes.memmove(ctypes.byref(lvl[k]), ctypes.byref(self._opt),
                               ctypes.sizeof(L.cot_options))
                t = level_tunings[k] if k < len(level_tunings) and level_tunings[k] else {}
                for name in ("int_lanes", "repair_blocks", "sites_per_bucket", "full_pass_ring",
                             "pass_mode"):
                    if name in t:
                        setattr(lvl[k], name, type(getattr(lvl[k], name))(t[name]))
            keep.append(lvl)
            lo.level_options = ctypes.cast(lvl, ctypes.POINTER(L.cot_options))
        return lo, keep

    def solve_ladder(self, history: int = 10, max_iter: int = 20000, tol: float = 1e-5,
                     stop: str = "maxrel", h0_threshold: float = 0.0, verbose: int = 0,
                     device_loop=True, raise_on_fail: bool = True, ladder=None, **ladder_kw):
        """Coarse-to-fine solve from nothing but the problems (cot_solve_ladder): coarse
        jittered grids, density-floor stages for targets with zero-density regions, then the
        exact target to `stop`/`tol` ("maxrel" or "reference").  ladder: (cot_ladder_opts,
        keep-alive list) from ladder_options(); else ladder_options(**ladder_kw).  Returns
        (w, reports)."""
        lo, keep = ladder if ladder is not None else self.ladder_options(**ladder_kw)
        arr, keep2 = _problem_array(self.problems)
        o = L.cot_solve_opts()
        self._lib.cot_default_solve_opts(ctypes.byref(o))
        o.history = history
        o.max_iter = max_iter
        o.tol = tol
        o.stop_rule = {"reference": L.COT_STOP_REFERENCE, "maxrel": L.COT_STOP_MAXREL}[stop]
        o.h0_threshold = h0_threshold
        o.verbose = verbose
        o.device_loop = int(device_loop)
        o.accept_rule = 1 if self.precision == "mixed" else 0
        w = np.zeros(self.n_total)
        reps = (L.cot_report * self.n_problems)()
        rc = self._lib.cot_solve_ladder(self._ctx, arr, _ptr(w), ctypes.byref(o), ctypes.byref(lo),
                                        reps)
        del keep, keep2
        if raise_on_fail:
            L.check(rc, allow=(L.COT_ENOTCONV,))
        return w, [{k: getattr(r, k) for k, _ in L.cot_report._fields_} for r in reps]

    def _newton(self, w0, tol, max_iter, verbose, repair_rounds, device_loop=True, sweeps=2,
                coarse=256, small_cg=True):
        """Damped Newton steps (cot_solve_newton) from w0 to max-rel tol; device_loop == 2
        launches the conjugate-gradient kernels directly instead of replaying a graph."""
        w = self._w(w0).copy()
        o = L.cot_solve_opts()
        self._lib.cot_default_solve_opts(ctypes.byref(o))
        o.tol = tol
        o.max_iter = max_iter
        o.stop_rule = L.COT_STOP_MAXREL
        o.verbose = verbose
        o.repair_rounds = repair_rounds
        o.device_loop = int(device_loop)
        o.newton_sweeps = int(sweeps)
        o.newton_coarse = int(coarse)
        o.small_cg = int(bool(small_cg))
        reps = (L.cot_report * self.n_problems)()
        rc = self._lib.cot_solve_newton(self._ctx, _ptr(w), ctypes.byref(o), reps)
        L.check(rc, allow=(L.COT_ENOTCONV,))
        return w, [{k: getattr(r, k) for k, _ in L.cot_report._fields_} for r in reps]

    def solve_reflection(self, w0, lens, history: int = 7, max_iter: int = 50001,
                         tol: float = 1e-2, initial_step: float = 1.0,
                         h0_threshold: float = 1e-5, aa_history_size: int = 7,
                         aa_fixed_para: float = -2.0, if_aa_op_initial: bool = True,
                         device_loop=True, verbose: int = 0, raise_on_fail: bool = True):
        """The t >= 1 solve (aarl_direct_ot, ot_op.cuh:591-1323) of a one-problem batch: source
        masses from the lens's Fresnel transmittance at the current cells' geometric centroids,
        L-BFGS with those masses frozen (if_aa_op_initial), then Anderson acceleration of
        w + aa_fixed_para * (mass - nu) with the masses recomputed at every trial, until
        |mass - nu|_2 < tol.  lens: dict with points (N, 2), ins (N, 3), height (N,),
        clip_area (N,), z_of_len, z_of_receiver, refraction_index, area_scalar, and for a
        point light incident (N, 3), the unit directions of the light at the points
        (main.cu:218-231; parallel light along +z without it).  Returns (w, report)."""
        if self.n_problems != 1:
            raise ValueError("the reflection solve takes one-problem batches")
        w = self._w(w0).copy()
        n = self.n_total
        pts = np.ascontiguousarray(lens["points"], dtype=np.float64).reshape(n, 2)
        ins = np.ascontiguousarray(lens["ins"], dtype=np.float64).reshape(n, 3)
        hgt = np.ascontiguousarray(lens["height"], dtype=np.float64).reshape(n)
        clip = np.ascontiguousarray(lens["clip_area"], dtype=np.float64).reshape(n)
        r = L.cot_reflection()
        r.points, r.inside_dir, r.height, r.clip_area = _ptr(pts), _ptr(ins), _ptr(hgt), _ptr(clip)
        r.z_of_len = float(lens["z_of_len"])
        r.z_of_receiver = float(lens["z_of_receiver"])
        r.refraction_index = float(lens["refraction_index"])
        r.area_scalar = float(lens["area_scalar"])
        r.aa_history_size = int(aa_history_size)
        r.aa_fixed_para = float(aa_fixed_para)
        r.if_aa_op_initial = int(bool(if_aa_op_initial))
        if lens.get("incident") is not None:
            inc = np.ascontiguousarray(lens["incident"], dtype=np.float64).reshape(n, 3)
            r.incident_dir = _ptr(inc)
        o = L.cot_solve_opts()
        self._lib.cot_default_solve_opts(ctypes.byref(o))
        o.history = history
        o.max_iter = max_iter
        o.tol = tol
        o.initial_step = initial_step
        o.h0_threshold = h0_threshold
        o.verbose = verbose
        o.device_loop = int(device_loop)
        rep = L.cot_report()
        rc = self._lib.cot_solve_reflection(self._ctx, _ptr(w), ctypes.byref(o), ctypes.byref(r),
                                            ctypes.byref(rep))
        if raise_on_fail:
            L.check(rc, allow=(L.COT_ENOTCONV,))
        return w, {k: getattr(rep, k) for k, _ in L.cot_report._fields_}


def direct_ot(proj_xy, weights, picture, picture_size, total_iter, convergence_tolerance,
              initial_step, history_size, proj_domain, carry_energy, precision="fp64",
              stop="reference", h0_threshold=1e-5):
    """Mirror of direct_ot (ot_op.cuh:66-70) for one problem: solves in place on `weights`
    and returns (output_points, if_exist) -- the rho-weighted centroids of the final cells.

    picture is the normalised target in the reference's flat layout (picture[ix + iy * m],
    picture_size = (m, n)); carry_energy the source masses."""
    from .workloads import Problem

    m, n = int(picture_size[0]), int(picture_size[1])
    rho = np.asarray(picture, dtype=np.float64).reshape(n, m)
    prob = Problem(sites=np.asarray(proj_xy, dtype=np.float64).reshape(-1, 2),
                   nu=np.asarray(carry_energy, dtype=np.float64), rho=rho,
                   proj_domain=tuple(proj_domain))
    with OTBatch([prob], precision=precision) as b:
        w, rep = b.solve(np.asarray(weights, dtype=np.float64), history=history_size,
                         max_iter=total_iter, tol=convergence_tolerance,
                         initial_step=initial_step, stop=stop, h0_threshold=h0_threshold)
        cen, ex = b.centroids(w, weighted=True)
    weights[...] = w
    return cen.ravel(), ex
