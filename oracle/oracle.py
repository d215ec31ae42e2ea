"""ORACLE -- test infrastructure only.

FP64 CPU restatement of the reference's semi-discrete OT dual solve (rmk523/caustic,
Optimal_Transport_rl/), used as the checker for the HIP product path.  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it.

Pieces and what they follow:
  * ``regular_neighbours`` -- the regular (weighted Delaunay) triangulation by the textbook lifting
    map: sites lifted to (x, y, x^2 + y^2 - w); the lower convex hull (scipy's qhull) gives the
    triangulation.  Stands in for CGAL's Regular_triangulation_2 (power_diagram.cpp:51-52),
    which is not available here.
  * ``power_cells`` -- restricted power cells; the C function ``orc_power_cells`` clips the box by
    each neighbour's half-plane (semantics of power_diagram.cpp:8-134 inside proj_domain).
  * ``evaluate`` -- compute_inte / compute_f / compute_center (basic_function.cuh:743-1191).
  * ``lbfgs`` -- the L-BFGS driver of direct_ot (ot_op.cuh:289-537) with the reference's rules;
    ``stop="maxrel"`` swaps only the stopping test for max_i |mass_i - nu_i| / nu_i < tol.
  * ``anderson`` / ``reflection_solve`` -- the t >= 1 path of aarl_direct_ot (ot_op.cuh:591-1323):
    Fresnel-weighted source masses, L-BFGS with frozen masses, then Anderson acceleration.
  * ``solid_angle`` / ``point_light_source`` -- the point-light source masses (compute_solid_angle,
    solid_angle, len_solid_angle: basic_function.cuh:558-588, 635-658; ot_op.cuh:160-206) and the
    light's directions at the points (main.cu:218-231).
  * ``irra_face_areas`` -- the pixel pieces of a triangle mesh, for the irra_render extension's
    image and gradients (render_exe_program/render.cu:93-288, 422-542).
  * ``gaussian_splat`` -- the receiver splat of utils/basic_function.py:674-709.
  * ``read_pic_txt`` / ``format_weight`` -- text formats (basic_function.cuh:145-204,
    main.cu:295-299).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# COT_ORACLE_LIB names another build of ot_oracle.cpp (the sanitizer build of the CPU tests)
_LIB_PATH = os.environ.get("COT_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_ubyte)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_power_cells.restype = ctypes.c_longlong
        L.orc_power_cells.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip, _ip, ctypes.c_longlong,
                                      _ip, _ip, _dp]
        L.orc_cell_integrals.restype = None
        L.orc_cell_integrals.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip, _ip, _dp, _dp,
                                         ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _up,
                                         _dp, _dp]
        L.orc_clip_area.restype = None
        L.orc_clip_area.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp]
        L.orc_render.restype = None
        L.orc_render.argtypes = [ctypes.c_int, _dp, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int,
                                 _dp, _dp]
        L.orc_fresnel_rate.restype = None
        L.orc_fresnel_rate.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, _dp]
        L.orc_edge_green_moment.restype = ctypes.c_double
        L.orc_edge_green_moment.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_fan_moment.restype = ctypes.c_double
        L.orc_fan_moment.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_set_centroid_min_piece.restype = None
        L.orc_set_centroid_min_piece.argtypes = [ctypes.c_double]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------------------------
# Power diagram
# --------------------------------------------------------------------------------------------

def regular_neighbours(xy: np.ndarray, w: np.ndarray):
    """CSR neighbour lists of the regular triangulation of weighted sites (lifting map)."""
    from scipy.spatial import ConvexHull

    xy = _f64(xy)
    w = _f64(w)
    n = xy.shape[0]
    # Centre and scale for qhull's conditioning; the lower hull is invariant under this
    # similarity (weights scale with the square of the length scale, plus a gauge shift).
    c = xy.mean(axis=0)
    s = max(float(np.ptp(xy[:, 0])), float(np.ptp(xy[:, 1])), 1e-300)
    p = (xy - c) / s
    ww = (w - w.mean()) / (s * s)
    lifted = np.column_stack([p, (p * p).sum(axis=1) - ww])
    hull = ConvexHull(lifted, qhull_options="Qt Qbb")
    lower = hull.equations[:, 2] < -1e-12
    tri = hull.simplices[lower]
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e = np.concatenate([e, e[:, ::-1]])
    e = np.unique(e, axis=0)
    counts = np.bincount(e[:, 0], minlength=n)
    ptr = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(counts, out=ptr[1:])
    return ptr, np.ascontiguousarray(e[:, 1], dtype=np.int32)


def power_cells(xy, w, box, method: str = "qhull"):
    """Restricted power cells of sites `xy` (N,2) with weights `w` in box (x0,x1,y0,y1).

    Returns (start, length, verts) in the reference's flattened CCW layout."""
    xy = _f64(xy)
    w = _f64(w)
    box = _f64(box)
    n = xy.shape[0]
    start = np.zeros(n, dtype=np.int32)
    length = np.zeros(n, dtype=np.int32)
    cap = 64 * n + 64
    verts = np.zeros(2 * cap, dtype=np.float64)
    if method == "qhull" and n >= 4:
        ptr, idx = regular_neighbours(xy, w)
        tot = lib().orc_power_cells(n, _d(xy), _d(w), _d(box), _i(ptr), _i(idx), cap, _i(start),
                                    _i(length), _d(verts))
    else:
        tot = lib().orc_power_cells(n, _d(xy), _d(w), _d(box), None, None, cap, _i(start),
                                    _i(length), _d(verts))
    if tot < 0:
        raise RuntimeError("oracle polygon buffer overflow")
    return start, length, verts[: 2 * tot].copy()


def cell_polygon(start, length, verts, i):
    return verts[2 * start[i]: 2 * (start[i] + length[i])].reshape(-1, 2)


# --------------------------------------------------------------------------------------------
# Per-cell integrals
# --------------------------------------------------------------------------------------------

class Evaluation:
    __slots__ = ("mass", "f_cell", "f", "centroid", "exist", "area", "geo", "max_len", "cells")


def evaluate(xy, w, nu, rho, proj, method: str = "qhull", box=None,
             centroid_min_piece: float = 1e-10) -> Evaluation:
    """Masses, objective, centroids of the power diagram of (xy, w) over image rho.

    rho: (H, W) mass per pixel (reference layout, x fast); proj = (x0, x1, y0, y1).
    box: bounding box of the cells before the domain clip (default: proj).
    centroid_min_piece: pixel pieces of at most this area are left out of the rho-weighted
    centroid (the reference uses 1e-10, basic_function.cuh:1153; 0 keeps every piece)."""
    lib().orc_set_centroid_min_piece(float(centroid_min_piece))
    xy = _f64(xy)
    w = _f64(w)
    nu = _f64(nu)
    rho = _f64(rho)
    proj = _f64(proj)
    n = xy.shape[0]
    H, W = rho.shape
    start, length, verts = power_cells(xy, w, proj if box is None else box, method)
    ev = Evaluation()
    ev.mass = np.zeros(n)
    ev.f_cell = np.zeros(n)
    ev.centroid = np.zeros((n, 2))
    ev.exist = np.zeros(n, dtype=np.uint8)
    ev.area = np.zeros(n)
    ev.geo = np.zeros((n, 2))
    lib().orc_cell_integrals(n, _d(xy), _d(w), _d(nu), _i(start), _i(length), _d(verts),
                             _d(rho), W, H, _d(proj), _d(ev.mass), _d(ev.f_cell),
                             _d(ev.centroid), ev.exist.ctypes.data_as(_up), _d(ev.area),
                             _d(ev.geo))
    ev.f = float(np.sum(ev.f_cell))
    ev.max_len = int(length.max()) if n else 0
    ev.cells = (start, length, verts)
    return ev


def clip_area(cells, dom):
    start, length, verts = cells
    out = np.zeros(start.shape[0])
    lib().orc_clip_area(start.shape[0], _i(start), _i(length), _d(verts), _d(_f64(dom)),
                        _d(out))
    return out


def render(cells, nu, rho_shape, proj):
    start, length, verts = cells
    H, W = rho_shape
    img = np.zeros(H * W)
    lib().orc_render(start.shape[0], _d(_f64(nu)), _i(start), _i(length), _d(verts), W, H,
                     _d(_f64(proj)), _d(img))
    return img.reshape(H, W)


def irra_face_areas(points, faces, pm, pn, proj, signed=False):
    """|T_f n p| for every face f of a triangle mesh lying inside proj and every pixel p of the
    pm x pn grid: [F, pn, pm] (pixel (ix, iy) at [f, iy, ix]), the pixel pieces of the
    irra_render extension's kernels (render_exe_program/render.cu:93-171, 422-542) by the same
    Sutherland-Hodgman clipping as ``render``.  signed: each piece carries the face's
    orientation (negative for a clockwise face), as computer_grad_inte sums them
    (compute_grad_area, render_exe_program/clip.h:277-292: the fan area without abs, and
    clipping keeps a polygon's orientation)."""
    pts = _f64(points)
    out = np.zeros((len(faces), pn, pm))
    one = np.zeros(1, dtype=np.int32)
    three = np.full(1, 3, dtype=np.int32)
    for f, tri in enumerate(faces):
        v = np.ascontiguousarray(pts[np.asarray(tri)].reshape(-1))
        s2 = (v[2] - v[0]) * (v[5] - v[1]) - (v[4] - v[0]) * (v[3] - v[1])
        a = 0.5 * abs(s2)
        if a > 0.0:
            out[f] = render((one, three, v), np.array([a]), (pn, pm), proj)
            if signed and s2 < 0.0:
                out[f] = -out[f]
    return out


def gaussian_splat(points, energy, res_y, res_x, proj, sigma=0.5, radius=2, window_sums=None):
    """_gaussian_splat_raw_chunk (utils/basic_function.py:674-709) in float64 numpy: each
    point's (2r+1)^2 window around its pixel, Gaussian weights normalised over the window,
    pixels outside the image dropped; image [res_y, res_x].  window_sums: the normalising sums
    to use instead (the reference detaches them, :699, so its gradient holds them fixed)."""
    x0, x1, y0, y1 = proj
    sx, sy = (x1 - x0) / res_x, (y1 - y0) / res_y
    img = np.zeros((res_y, res_x))
    off = np.arange(-radius, radius + 1)
    for k, ((x, y), e) in enumerate(zip(_f64(points), _f64(energy))):
        px, py = (x - x0) / sx, (y - y0) / sy
        cx, cy = int(np.floor(px)), int(np.floor(py))
        xs, ys = cx + off, cy + off
        w = np.exp(-((xs[None, :] + 0.5 - px) ** 2 + (ys[:, None] + 0.5 - py) ** 2)
                   / (2 * sigma * sigma))
        w *= e / max(w.sum() if window_sums is None else window_sums[k], 1e-12)
        for a, yy in enumerate(ys):
            if 0 <= yy < res_y:
                for b, xx in enumerate(xs):
                    if 0 <= xx < res_x:
                        img[yy, xx] += w[a, b]
    return img


def splat_window_sums(points, res_y, res_x, proj, sigma=0.5, radius=2):
    """The normalising sum of each point's window weights (gaussian_splat)."""
    x0, x1, y0, y1 = proj
    sx, sy = (x1 - x0) / res_x, (y1 - y0) / res_y
    off = np.arange(-radius, radius + 1)
    out = []
    for x, y in _f64(points):
        px, py = (x - x0) / sx, (y - y0) / sy
        cx, cy = np.floor(px), np.floor(py)
        out.append(np.exp(-((cx + off[None, :] + 0.5 - px) ** 2
                            + (cy + off[:, None] + 0.5 - py) ** 2) / (2 * sigma * sigma)).sum())
    return np.array(out)


def source_masses(points, sample_domain):
    """nu_i = |Vor_i n sample_domain| / sum * area_scalar for a parallel light
    (ot_op.cuh:90-92,158,199-206): the unweighted Voronoi cells are built in the sample domain
    widened by 1/40 of its size, then clipped to the sample domain."""
    sd = _f64(sample_domain)
    xs, ys = sd[1] - sd[0], sd[3] - sd[2]
    corner = np.array([sd[0] - xs / 40, sd[1] + xs / 40, sd[2] - ys / 40, sd[3] + ys / 40])
    cells = power_cells(points, np.zeros(points.shape[0]), corner)
    a = clip_area(cells, sd)
    return a / a.sum() * (xs * ys)


def solid_angle(poly, light, z):
    """Solid angle under which a point light sees the polygon poly (k, 2) placed in the plane
    z, as compute_solid_angle does it (basic_function.cuh:558-588): fan from vertex 0, each
    spherical triangle 2 atan(a.(b x c) / (1 + a.b + b.c + c.a)) of the unit vectors a, b, c
    from the light to its corners (Van Oosterom and Strackee 1983), triangles with
    |denominator| < 1e-6 skipped, the magnitude of the sum."""
    poly = _f64(poly).reshape(-1, 2)
    if poly.shape[0] < 3:
        return 0.0
    lt = _f64(light)
    v = np.column_stack([poly[:, 0] - lt[0], poly[:, 1] - lt[1],
                         np.full(poly.shape[0], z - lt[2])])
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    total = 0.0
    a = v[0]
    for k in range(1, poly.shape[0] - 1):
        b, c = v[k], v[k + 1]
        num = float(np.dot(a, np.cross(b, c)))
        den = 1.0 + float(np.dot(a, b) + np.dot(b, c) + np.dot(a, c))
        if abs(den) < 1e-6:
            continue
        total += 2.0 * np.arctan(num / den)
    return abs(total)


def domain_solid_angle(sample_domain, light, z):
    """The sample domain's solid angle, its corners taken in len_solid_angle's order
    (basic_function.cuh:653-658)."""
    x0, x1, y0, y1 = _f64(sample_domain)
    return solid_angle(np.array([[x0, y0], [x0, y1], [x1, y1], [x1, y0]]), light, z)


def cell_solid_angles(cells, light, z):
    start, length, verts = cells
    return np.array([solid_angle(cell_polygon(start, length, verts, i), light, z)
                     for i in range(start.shape[0])])


def point_light_source(points, sample_domain, light, z_of_len):
    """Point-light source data (ot_op.cuh:160-206; main.cu:218-231): the Voronoi cells of the
    points inside the sample domain, each weighted by omega_i / omega_domain * area_scalar
    (returned as clip_area, which the t >= 1 masses take), nu_i = that weight normalised to sum
    area_scalar, and the unit directions from the light to the points in the plane z_of_len.
    The cells are built in the sample domain directly: inside it they equal the widened
    domain's cells clipped to it."""
    sd = _f64(sample_domain)
    area_scalar = (sd[1] - sd[0]) * (sd[3] - sd[2])
    pts = _f64(points)
    cells = power_cells(pts, np.zeros(pts.shape[0]), sd)
    clip = cell_solid_angles(cells, light, z_of_len) / domain_solid_angle(sd, light, z_of_len) \
        * area_scalar
    d = np.column_stack([pts[:, 0] - light[0], pts[:, 1] - light[1],
                         np.full(pts.shape[0], z_of_len - light[2])])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return clip / clip.sum() * area_scalar, clip, d


def fresnel_rate(points, in_dir, ins, height, targets, z_len, z_recv, n_idx):
    n = points.shape[0]
    out = np.zeros(n)
    lib().orc_fresnel_rate(n, _d(_f64(points)), _d(_f64(in_dir)), _d(_f64(ins)),
                           _d(_f64(height)), _d(_f64(targets)), z_len, z_recv, n_idx, _d(out))
    return out


# --------------------------------------------------------------------------------------------
# L-BFGS driver (ot_op.cuh:289-537)
# --------------------------------------------------------------------------------------------

class LbfgsResult:
    def __init__(self):
        self.w = None
        self.iterations = 0
        self.evaluations = 0
        self.converged = False
        self.history = []  # per-iteration (f, |g|_2, maxrel, step)


def max_rel_error(mass, nu):
    return float(np.max(np.abs(mass - nu) / nu))


def lbfgs(evaluate_fn, w0, nu, m=10, total_iter=50001, tol=1e-2, initial_step=1.0,
          stop="reference", h0_threshold=1e-5, max_vertices=200, callback=None,
          accept="decrease") -> LbfgsResult:
    """evaluate_fn(w) -> (f, mass, max_len).  Rules follow ot_op.cuh:289-537:

    * d = -g, two-loop recursion over <= m pairs, skipped when `restart` is set (:298-337)
    * H0 = (y.s)/(y.y) if y.y > h0_threshold else 1; rho = 1/(y.s) unchecked (:455-463)
    * g.d >= 0 -> d = -g, history cleared (:339-351)
    * trials halve the step on max_len > max_vertices or f >= f_prev (:360-411)
    * step <= 1e-4*step0 -> revert w, g, f; restart; clear history; skip the rest (:413-431)
    * stop (reference): |g_start|_2 < tol and no zero mass at the start of the iteration (:510)
    * stop (maxrel):    max |g|/nu < tol on the current iterate (checked after each update).
    * accept="derivative" (not in the reference): a trial is also accepted when the directional
      derivative (mass_trial - nu).d is <= 0, which for the convex objective certifies a decrease.
    """
    w = _f64(w0).copy()
    nu = _f64(nu)
    f, mass, _ = evaluate_fn(w)
    g = mass - nu
    res = LbfgsResult()
    res.evaluations = 1
    S = [None] * m
    Y = [None] * m
    rho = [0.0] * m
    alpha = [0.0] * m
    H0 = 1.0
    restart = False
    loop_iter = 0
    if stop == "maxrel" and max_rel_error(mass, nu) < tol:
        res.w, res.converged = w, True
        return res
    for it in range(total_iter):
        res.iterations = it + 1
        max_idx = min(loop_iter, m)
        d = -g
        mass_start = mass.copy()
        g_start = g.copy()
        if not restart:
            for i in range(1, max_idx + 1):
                idx = (loop_iter - i) % m
                alpha[idx] = float(np.dot(S[idx], d)) * rho[idx]
                d = d - alpha[idx] * Y[idx]
            d = d * H0
            for i in range(max_idx, 0, -1):
                idx = (loop_iter - i) % m
                beta = rho[idx] * float(np.dot(Y[idx], d))
                d = d + (alpha[idx] - beta) * S[idx]
        if float(np.dot(g, d)) >= 0:
            d = -g
            loop_iter = 0
            H0 = 1.0
        g_prev, w_prev, f_prev = g, w, f
        step = initial_step
        while True:
            w_try = w_prev + step * d
            f_try, mass_try, max_len = evaluate_fn(w_try)
            res.evaluations += 1
            if max_len > max_vertices:
                step *= 0.5
            elif f_try < f_prev or (accept == "derivative" and
                                    float(np.dot(mass_try - nu, d)) <= 0.0):
                w, f, mass = w_try, f_try, mass_try
                g = mass - nu
                restart = False
                break
            else:
                step *= 0.5
            if step <= 1e-4 * initial_step:
                restart = True
                g = g_prev
                loop_iter = 0
                H0 = 1.0
                w = w_prev
                f = f_prev
                break
        if step <= 1e-4 * initial_step:
            continue
        hist = loop_iter % m
        S[hist] = step * d
        Y[hist] = g - g_prev
        ys = float(np.dot(Y[hist], S[hist]))
        yy = float(np.dot(Y[hist], Y[hist]))
        rho[hist] = 1.0 / ys
        H0 = ys / yy if yy > h0_threshold else 1.0
        loop_iter += 1
        mr = max_rel_error(mass, nu)
        res.history.append((f, float(np.linalg.norm(g_start)), mr, step))
        if callback:
            callback(it, w, f, mass)
        if stop == "reference":
            if np.linalg.norm(g_start) < tol and not np.any(mass_start == 0):
                res.converged = True
                break
        else:
            if mr < tol:
                res.converged = True
                break
    res.w = w
    return res


def solve(problem, w0=None, m=10, stop="maxrel", tol=None, total_iter=5000, method="qhull",
          h0_threshold=0.0):
    """Solve a workloads.Problem on the CPU oracle (unit-square conventions)."""
    proj = np.asarray(problem.proj_domain, dtype=np.float64)
    nu = problem.nu

    def ev(w):
        e = evaluate(problem.sites, w, nu, problem.rho, proj, method=method)
        return e.f, e.mass, e.max_len

    w0 = np.zeros(problem.n) if w0 is None else w0
    return lbfgs(ev, w0, nu, m=m, total_iter=total_iter,
                 tol=problem.tol if tol is None else tol, stop=stop, h0_threshold=h0_threshold)


# --------------------------------------------------------------------------------------------
# Anderson acceleration and the t >= 1 driver (ot_op.cuh:591-1323)
# --------------------------------------------------------------------------------------------

def anderson_coefficients(A):
    """alpha = M^-1 1 / (1 . M^-1 1), M = A + 1e-10 I, A the Gram matrix of the stored
    residuals (ot_op.cuh:1147-1163; the reference factorises M by LDL^T)."""
    A = np.asarray(A, dtype=np.float64)
    H = A.shape[0]
    a = np.linalg.solve(A + 1e-10 * np.eye(H), np.ones(H))
    return a / a.sum()


class AndersonResult:
    def __init__(self):
        self.w = None
        self.g = None
        self.iterations = 0
        self.evaluations = 0
        self.converged = False
        self.restarts = 0
        self.history = []  # per-iteration (|g|_2 after the step, step)


def anderson(residual_fn, w0, g0, zeros0, m, beta, total_iter, tol, initial_step=1.0,
             max_vertices=200) -> AndersonResult:
    """Anderson acceleration of the fixed point G(w) = w + beta g(w) (ot_op.cuh:1112-1323).

    residual_fn(w) -> (g, zeros, max_len): the residual mass - nu at w (nu recomputed from
    w's own cells), the number of cells of zero mass and the largest cell's vertex count.
    (g0, zeros0) are the residual and zero count at w0.  Each iteration:

    * stores g and G(w) = w + beta g in slot aa % m and uses the H = min(aa + 1, m) slots
      0..H-1 (:1131-1145);
    * alpha from the H x H Gram matrix of the stored g (anderson_coefficients);
    * d = sum_h alpha_h G_h - w; trials w + t d from t = 1, halved while a cell exceeds
      max_vertices; the first admissible trial is accepted, residual and all (:1176-1243);
    * t <= 1e-4 initial_step keeps w and g and restarts the history (aa = 0) (:1244-1253);
    * stop when |g|_2 < tol after the step and -- on every 10th iteration only, counted at the
      start of the iteration -- no cell has zero mass (:1257-1268, 1296).
    """
    w = _f64(w0).copy()
    g = _f64(g0).copy()
    zeros = int(zeros0)
    Gs = [None] * m
    Fs = [None] * m
    aa = 0
    res = AndersonResult()
    for it in range(total_iter):
        res.iterations = it + 1
        H = min(aa + 1, m)
        slot = aa % m
        Gs[slot] = g.copy()
        Fs[slot] = w + beta * g
        A = np.array([[float(np.dot(Gs[i], Gs[j])) for j in range(H)] for i in range(H)])
        alpha = anderson_coefficients(A)
        d = sum(alpha[h] * Fs[h] for h in range(H)) - w
        zeros_start = zeros
        step = 1.0
        while True:
            w_try = w + step * d
            g_try, z_try, max_len = residual_fn(w_try)
            res.evaluations += 1
            if max_len > max_vertices:
                step *= 0.5
            else:
                w, g, zeros = w_try, _f64(g_try), int(z_try)
                aa += 1
                break
            if step <= 1e-4 * initial_step:
                aa = 0
                res.restarts += 1
                break
        gn = float(np.linalg.norm(g))
        res.history.append((gn, step))
        degenerate = zeros_start if it % 10 == 0 else 0
        if gn < tol and degenerate == 0:
            res.converged = True
            break
    res.w = w
    res.g = g
    return res


class ReflectionResult:
    def __init__(self):
        self.w = None
        self.nu = None
        self.stage_a = None
        self.stage_b = None


def fresnel_masses(ev, lens):
    """nu_i = clip_area_i T1_i T2_i / sum * area_scalar, the transmittances taken at the cells'
    geometric centroids ev.geo (ot_op.cuh:779-795); the light arrives along lens["incident"]
    (a point light) or along wi = (0, 0, 1) without it (parallel light)."""
    n = lens["points"].shape[0]
    wi = lens.get("incident")
    wi = np.tile([0.0, 0.0, 1.0], (n, 1)) if wi is None else _f64(wi).reshape(n, 3)
    rate = fresnel_rate(lens["points"], wi, lens["ins"], lens["height"], ev.geo,
                        lens["z_of_len"], lens["z_of_receiver"], lens["refraction_index"])
    we = _f64(lens["clip_area"]) * rate
    return we / we.sum() * lens["area_scalar"]


def reflection_solve(xy, w0, rho, proj, lens, m=7, total_iter=50001, tol=1e-2,
                     initial_step=1.0, h0_threshold=1e-5, aa_history_size=7, aa_fixed_para=-2.0,
                     if_aa_op_initial=True, max_vertices=200) -> ReflectionResult:
    """The t >= 1 solve of aarl_direct_ot (ot_op.cuh:591-1323) for a parallel light.

    lens: dict of points (n, 2) lens-plane source points, ins (n, 3) unit directions inside
    the lens, height (n,), clip_area (n,), z_of_len, z_of_receiver, refraction_index,
    area_scalar.  Stage A (if_aa_op_initial): L-BFGS with the Fresnel masses of w0's cells
    frozen, the reference's stopping rule (:787-1073).  Then the masses are recomputed at the
    result (:1076-1103) and stage B runs ``anderson`` with masses recomputed at every trial."""
    xy = _f64(xy)
    rho = _f64(rho)
    proj = _f64(proj)
    w = _f64(w0).copy()
    out = ReflectionResult()
    ev = evaluate(xy, w, np.zeros(xy.shape[0]), rho, proj)
    nu = fresnel_masses(ev, lens)
    if if_aa_op_initial:
        def ev_a(wv):
            e = evaluate(xy, wv, nu, rho, proj)
            return e.f, e.mass, e.max_len

        out.stage_a = lbfgs(ev_a, w, nu, m=m, total_iter=total_iter, tol=tol,
                            initial_step=initial_step, stop="reference",
                            h0_threshold=h0_threshold, max_vertices=max_vertices)
        w = out.stage_a.w

    def residual(wv):
        e = evaluate(xy, wv, np.zeros(xy.shape[0]), rho, proj)
        nu_v = fresnel_masses(e, lens)
        return e.mass - nu_v, int(np.sum(e.mass == 0.0)), e.max_len

    g, zeros, _ = residual(w)
    out.stage_b = anderson(residual, w, g, zeros, aa_history_size, aa_fixed_para, total_iter, tol,
                           initial_step, max_vertices)
    out.w = out.stage_b.w
    out.nu = fresnel_masses(evaluate(xy, out.w, np.zeros(xy.shape[0]), rho, proj), lens)
    return out


# --------------------------------------------------------------------------------------------
# Text formats
# --------------------------------------------------------------------------------------------

def read_pic_txt(path):
    """Whitespace matrix of m rows and n columns -> (m, n, data) with data[iy*m + ix] =
    text[m-1-ix][iy] (basic_function.cuh:145-204).  Returned as a (n, m) array A[iy, ix]."""
    rows = []
    with open(path) as fh:
        for line in fh:
            vals = line.split()
            if vals:
                rows.append([float(v) for v in vals])
    mat = np.array(rows, dtype=np.float64)
    m, n = mat.shape
    # element (ix, iy) = mat[m-1-ix, iy]; A[iy, ix]
    return mat[::-1, :].T.copy()


def format_weight(x: float) -> str:
    """C++ default ostream formatting of a double (precision 6, %g style)."""
    return "%g" % x
