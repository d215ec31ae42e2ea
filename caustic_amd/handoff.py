"""In-memory hand-off between the lens-design loop and the OT solve.

The reference's loop (spline_irra.py:67-85) writes seed files and t/height.txt, runs the
caustic_GPU process (run_cpp_program, utils/basic_function.py:14-34: subprocess with the run's
directory as cwd) and reads in.txt and t/out.txt back (load_ot_data, utils/data_utils.py:19-24)
as float32 CUDA tensors with requires_grad; the next run starts from the weights the previous
one wrote (main.cu:193-198).  OTHandoff does the same steps in memory, on the same GPU:

    ot = OTHandoff(loader.data, picture_text, points, proj_points, inside_light_direct)
    for total_iter in range(T):
        xy_tensor, proj_xy_tensor = ot.run(total_iter, height)     # in place of the file trip

with the executable's arithmetic (caustic_amd/csrc/caustic_gpu_main.cpp, which follows
main.cu:51-342): source masses of the sample points' Voronoi cells in the sample domain (parallel
light) or of their solid angles from the point light, the picture normalised to the sample
domain's area, L-BFGS from w = 0 with the reference's stopping rule at t = 0, the reflection-aware
solve (Fresnel masses, L-BFGS then Anderson) from the previous weights at t >= 1, rho-weighted
centroids of the existing cells.  By default the values pass through the same 6-significant-digit
text the files carry (weights between runs, in/out points), so a run reproduces the executable's
outputs; text_precision=False keeps full precision instead.
"""
from __future__ import annotations

import numpy as np

from . import workloads as W
from .ot import OTBatch


def _as_text(a):
    """Values as C++ ostream prints them (6 significant digits) and a reader parses them back."""
    return np.array([float("%g" % v) for v in np.ravel(a)]).reshape(np.shape(a))


class OTHandoff:
    """One lens-design run's OT solves, kept in memory between ot_time 0, 1, ...

    settings: the run's configuration (configs/_py_sp_settings.json as a dict: the keys
    main.cu:89-121 reads); picture: the target's text matrix (data/pic_data/<name>.txt, m rows);
    points, proj_xy: (N, 2) seed points and their projections (utils/data_utils.py:374-435);
    inside_light_direct: (N, 3) directions inside the lens.  precision: "fp64" (the executable's)
    or "mixed"; solver: "lbfgs" (the reference's algorithm at t = 0) or "ladder"
    (cot_solve_ladder, the same stopping rule).  diagnostics: as the executable by default, the
    render and the cell polygons at the start weights before each solve and at the final weights
    after it (ot_op.cuh:214-219, 499-533; kept in self.diagnostics[t]); they evaluate the cells
    like the executable does, so a run follows the executable's evaluations exactly."""

    def __init__(self, settings, picture, points, proj_xy, inside_light_direct,
                 precision="fp64", solver="lbfgs", text_precision=True, diagnostics=True):
        sp, lp, li = settings["sample_point_para"], settings["len_para"], settings["light_para"]
        self.total_iter = int(sp["total_iter"])
        self.tol = float(sp["convergence_tolerance"])
        self.history = int(sp["history_size"])
        self.initial_step = float(sp["initial_step"])
        self.aa_history = int(sp["aa_history_size"])
        self.aa_fixed = float(sp["aa_fixed_para"])
        self.aa_initial = bool(sp["if_aa_op_initial"])
        self.sample_domain = tuple(float(v) for v in sp["sample_domain"])
        self.proj_domain = tuple(float(v) for v in lp["proj_domain"])
        self.z_of_len = float(lp["z_of_len"])
        self.z_of_receiver = float(lp["z_of_receiver"])
        self.refraction_index = float(lp["refraction_index"])
        self.light_type = li["light_type"]
        if self.light_type not in ("parallel_light", "point_light"):
            raise ValueError(f"unknown light_type {self.light_type!r}")
        self.light = np.asarray(li["point_light_coordinate"], dtype=np.float64).reshape(3)
        if solver not in ("lbfgs", "ladder"):
            raise ValueError("solver must be 'lbfgs' or 'ladder'")
        self.solver, self.precision, self.text = solver, precision, text_precision
        self.diagnose = bool(diagnostics)
        self.diagnostics = {}
        self.points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
        self.proj_xy = np.ascontiguousarray(proj_xy, dtype=np.float64).reshape(-1, 2)
        n = self.points.shape[0]
        if self.proj_xy.shape[0] != n:
            raise ValueError("points and proj_xy differ in length")
        self.inside = np.ascontiguousarray(inside_light_direct, dtype=np.float64).reshape(n, 3)
        x0, x1, y0, y1 = self.sample_domain
        self.area_scalar = (x1 - x0) * (y1 - y0)
        pic = W.picture_from_text_rows(np.asarray(picture, dtype=np.float64))
        # rho = picture / (sum / area_scalar) (ot_op.cuh:81-88,154); sums in the executable's
        # order (np.cumsum adds left to right, as its loops do)
        self.rho = pic / (np.cumsum(pic.ravel())[-1] / self.area_scalar)
        self.area, self.nu, self.incident = self._source_masses()
        self.weights = None      # the last run's final weights (what the next run starts from)
        self.in_points = None    # in.txt: the t = 0 source points of the existing cells
        self.reports = {}

    def _source_masses(self):
        """(area, nu, incident): the source cells' areas in the sample domain (parallel light) or
        their solid angles from the light scaled to the domain's area (point light), and nu, the
        same normalised to area_scalar (ot_op.cuh:90-206); incident: the light's unit directions
        at the points for a point light (main.cu:218-231), None for a parallel light."""
        n = self.points.shape[0]
        src = W.Problem(sites=self.points, nu=np.ones(n), rho=np.ones((1, 1)),
                        proj_domain=self.sample_domain)
        with OTBatch([src], precision="fp64") as b:
            if self.light_type == "parallel_light":
                area = b.cell_areas(np.zeros(n))
                incident = None
            else:
                omega = b.cell_solid_angles(self.light, self.z_of_len, np.zeros(n))
                x0, x1, y0, y1 = self.sample_domain
                whole = W.Problem(sites=np.array([[0.5 * (x0 + x1), 0.5 * (y0 + y1)]]),
                                  nu=np.ones(1), rho=np.ones((1, 1)),
                                  proj_domain=self.sample_domain)
                with OTBatch([whole], precision="fp64") as wb:
                    dom = float(wb.cell_solid_angles(self.light, self.z_of_len, np.zeros(1))[0])
                if not dom > 0.0:
                    raise ValueError("the light sees the sample domain under no solid angle")
                area = omega / dom * self.area_scalar
                d = np.column_stack([self.points, np.full(n, self.z_of_len)]) - self.light
                r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
                incident = d / r[:, None]
        return area, area / np.cumsum(area)[-1] * self.area_scalar, incident

    def run(self, ot_time, height):
        """The solve of ot_time (0, 1, ...) with the back-surface heights of this stage (N,):
        returns (origin_xy, aim_xy), float32 tensors on cuda:0 with requires_grad, as
        load_ot_data returns them (in.txt and <ot_time>/out.txt)."""
        import torch

        t = int(ot_time)
        if t > 0 and self.weights is None:
            raise RuntimeError("ot_time >= 1 starts from the previous run's weights: run 0 first")
        prob = W.Problem(sites=self.proj_xy, nu=self.nu, rho=self.rho,
                         proj_domain=self.proj_domain)
        with OTBatch([prob], precision=self.precision) as b:
            diag = {}
            if self.diagnose:
                w0 = np.zeros(prob.n) if t == 0 else self.weights
                diag["render_initial"] = b.render(w0)[0]
                diag["mesh_initial"] = b.cell_polygons(w0)
            if t == 0 and self.solver == "ladder":
                w, rep = b.solve_ladder(history=self.history, max_iter=self.total_iter,
                                        tol=self.tol, stop="reference", h0_threshold=1e-5,
                                        raise_on_fail=False)
                rep = rep[0]
            elif t == 0:
                w, rep = b.solve(None, history=self.history, max_iter=self.total_iter,
                                 tol=self.tol, initial_step=self.initial_step,
                                 stop="reference", h0_threshold=1e-5, raise_on_fail=False)
                rep = rep[0]
            else:
                lens = {"points": self.points, "ins": self.inside,
                        "height": np.asarray(height, dtype=np.float64).reshape(-1),
                        "clip_area": self.area, "z_of_len": self.z_of_len,
                        "z_of_receiver": self.z_of_receiver,
                        "refraction_index": self.refraction_index,
                        "area_scalar": self.area_scalar, "incident": self.incident}
                w, rep = b.solve_reflection(self.weights, lens, history=self.history,
                                            max_iter=self.total_iter, tol=self.tol,
                                            initial_step=self.initial_step,
                                            aa_history_size=self.aa_history,
                                            aa_fixed_para=self.aa_fixed,
                                            if_aa_op_initial=self.aa_initial,
                                            raise_on_fail=False)
            if self.diagnose:
                diag["render_final"] = b.render(w)[0]
                diag["mesh_final"] = b.cell_polygons(w)
                self.diagnostics[t] = diag
            cen, ex = b.centroids(w, weighted=True)
        self.reports[t] = rep
        # the next run reads the weights file (main.cu:193-198, written at 6 digits, :295-299)
        self.weights = _as_text(w) if self.text else w
        self.full_weights = w
        if t == 0:
            self.in_points = self.points[ex]
        aim = cen[ex]
        origin = self.in_points
        if self.text:
            origin, aim = _as_text(origin), _as_text(aim)
        dev = torch.device("cuda:0")
        return (torch.as_tensor(origin, dtype=torch.float32, device=dev).requires_grad_(),
                torch.as_tensor(aim, dtype=torch.float32, device=dev).requires_grad_())
