"""Seeded synthetic OT problems: sites on jittered grids, piecewise-constant target images.

Every problem lives on the unit square: proj_domain = sample_domain = [0, 1] x [0, 1], so the
reference's area_scalar is 1 and the target is normalised to total mass 1 (sum of pixel values).
The image layout is the reference's: ``rho[iy, ix]`` flattened row-major, i.e. element
``ix + iy * m`` with x the fast index (Optimal_Transport_rl/basic_function.cuh:195-201).

Random streams are numpy PCG64 seeded with ``base_seed + problem_index``.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

UNIT_DOMAIN = (0.0, 1.0, 0.0, 1.0)


@dataclasses.dataclass
class Problem:
    """One semi-discrete OT problem in the reference's conventions.

    sites:  (N, 2) float64 site positions (the reference's ``proj_xy``)
    nu:     (N,) float64 source masses (``carry_energy``), summing to ``area_scalar``
    rho:    (H, W) float64 target, mass per pixel, summing to ``area_scalar``
            (``picture / gray_multiple``, ot_op.cuh:81-88,154); W pixels along x
    proj_domain: (x0, x1, y0, y1) of the target plane
    tol:    max relative cell-mass error at which the solve stops
    site_box: (x0, x1, y0, y1) of the region the sites sample when it is not proj_domain (a
            source plane seen at a tilt); coarse problems of the warm start sample it too
    """

    sites: np.ndarray
    nu: np.ndarray
    rho: np.ndarray
    proj_domain: tuple = UNIT_DOMAIN
    tol: float = 1e-5
    name: str = ""
    seed: int = 0
    site_box: tuple = None

    @property
    def frame(self) -> tuple:
        """The region the sites sample: site_box, else proj_domain."""
        return tuple(self.site_box) if self.site_box is not None else tuple(self.proj_domain)

    @property
    def n(self) -> int:
        return int(self.sites.shape[0])

    @property
    def pic_size(self) -> tuple:
        """(m, n) = (pixels along x, pixels along y), the reference's pic_size."""
        return int(self.rho.shape[1]), int(self.rho.shape[0])


def rng_for(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def jittered_grid(n_side: int, rng: np.random.Generator, jitter: float = 0.25) -> np.ndarray:
    """n_side^2 sites ((i + 1/2 + u) / n, (j + 1/2 + v) / n), u, v ~ U[-jitter, jitter]."""
    j, i = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="ij")
    u = rng.uniform(-jitter, jitter, size=(n_side, n_side))
    v = rng.uniform(-jitter, jitter, size=(n_side, n_side))
    x = (i + 0.5 + u) / n_side
    y = (j + 0.5 + v) / n_side
    return np.stack([x.ravel(), y.ravel()], axis=1).astype(np.float64)


def pixel_centres(res: int):
    c = (np.arange(res) + 0.5) / res
    return np.meshgrid(c, c, indexing="xy")  # X[iy, ix], Y[iy, ix]


def gaussian_mixture(res: int, rng: np.random.Generator, k: int, sigma_range, amp_range,
                     centre_range, floor: float) -> np.ndarray:
    X, Y = pixel_centres(res)
    img = np.full((res, res), floor, dtype=np.float64)
    for _ in range(k):
        cx, cy = rng.uniform(*centre_range, size=2)
        s = rng.uniform(*sigma_range)
        a = rng.uniform(*amp_range)
        img += a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * s * s))
    return img / img.sum()


def disks_and_rects(res: int, rng: np.random.Generator, count: int = 40,
                    size_range=(0.01, 0.1), high: float = 1.0, floor: float = 0.01) -> np.ndarray:
    """`count` shapes, each a disk (radius) or axis-aligned rectangle (two sides) drawn from
    size_range, centres U[0,1]^2, value `high` on a `floor` background (pixel-centre inclusion)."""
    X, Y = pixel_centres(res)
    img = np.full((res, res), floor, dtype=np.float64)
    for _ in range(count):
        cx, cy = rng.uniform(0.0, 1.0, size=2)
        if rng.uniform() < 0.5:
            r = rng.uniform(*size_range)
            img[(X - cx) ** 2 + (Y - cy) ** 2 <= r * r] = high
        else:
            wx, wy = rng.uniform(*size_range, size=2)
            img[(np.abs(X - cx) <= 0.5 * wx) & (np.abs(Y - cy) <= 0.5 * wy)] = high
    return img / img.sum()


def make_problem(sites: np.ndarray, rho: np.ndarray, tol: float, name: str, seed: int) -> Problem:
    n = sites.shape[0]
    return Problem(sites=sites, nu=np.full(n, 1.0 / n), rho=rho, tol=tol, name=name, seed=seed)


def plain_fp64(seed: int = 1000, n_side: int = 64, res: int = 256) -> Problem:
    """N = 4,096 sites on a 64x64 jittered grid; 256^2 target of 3 Gaussians
    (centres U[0.2,0.8]^2, sigma U[0.05,0.15], amplitude U[0.5,1]) on a 0.05 floor; tol 1e-4."""
    rng = rng_for(seed)
    sites = jittered_grid(n_side, rng)
    rho = gaussian_mixture(res, rng, 3, (0.05, 0.15), (0.5, 1.0), (0.2, 0.8), 0.05)
    return make_problem(sites, rho, 1e-4, f"gauss3_n{n_side * n_side}_{res}px", seed)


def large_shapes(seed: int = 2000, n_side: int = 256, res: int = 1024) -> Problem:
    """N = 65,536 sites on a 256x256 jittered grid; 1024^2 target of 40 disks / rectangles at 1.0
    on a 0.01 floor (about 100:1 contrast); tol 1e-5."""
    rng = rng_for(seed)
    sites = jittered_grid(n_side, rng)
    rho = disks_and_rects(res, rng)
    return make_problem(sites, rho, 1e-5, f"shapes40_n{n_side * n_side}_{res}px", seed)


def batched_uneven(seed: int = 3000, count: int = 64, res: int = 512,
                   n_range=(2048, 32768)) -> list:
    """`count` problems, N_k = round(sqrt(exp(U[ln 2048, ln 32768])))^2 sites each, 512^2 targets
    of K ~ U{1..8} Gaussians (sigma U[0.02,0.2], amplitude U[0.5,1], centres U[0.1,0.9]^2) on a
    0.02 floor; tol 1e-5 each."""
    out = []
    for k in range(count):
        rng = rng_for(seed + k)
        n_side = int(round(math.sqrt(math.exp(rng.uniform(math.log(n_range[0]),
                                                          math.log(n_range[1]))))))
        sites = jittered_grid(n_side, rng)
        kk = int(rng.integers(1, 9))
        rho = gaussian_mixture(res, rng, kk, (0.02, 0.2), (0.5, 1.0), (0.1, 0.9), 0.02)
        out.append(make_problem(sites, rho, 1e-5, f"batch{k}_n{n_side * n_side}", seed + k))
    return out


def warm_start_sample(seed: int, n_side: int = 128, res: int = 512) -> Problem:
    """A shapes-style problem on a 128x128 jittered grid (N = 16,384), for warm-start data."""
    rng = rng_for(seed)
    sites = jittered_grid(n_side, rng)
    rho = disks_and_rects(res, rng)
    return make_problem(sites, rho, 1e-5, f"warm_n{n_side * n_side}_{res}px_s{seed}", seed)


def box_downsample(rho: np.ndarray, out: int) -> np.ndarray:
    """Box-filter a square image to out x out (sum-preserving)."""
    f = rho.shape[0] // out
    return rho.reshape(out, f, out, f).sum(axis=(1, 3))


def lens_problem(seed: int = 4000, n_side: int = 16, res: int = 64, tilt: float = 0.3,
                 z_of_receiver: float = 1.0, refraction_index: float = 1.5):
    """A t >= 1 (refraction-aware) problem: the plain_fp64-style target on a res^2 image and
    n_side^2 jittered sites, plus the lens data the Fresnel-weighted masses need.

    Lens plane z = 0, receiver z = z_of_receiver; source points = the sites (parallel light);
    inside directions (u, v, 1) normalised with u, v ~ U[-tilt, tilt]; back-surface heights
    0.1 + U[0, 0.02]; clip areas (1 + U[-0.1, 0.1]) / N.  Returns (Problem, lens dict)."""
    rng = rng_for(seed)
    sites = jittered_grid(n_side, rng)
    rho = gaussian_mixture(res, rng, 3, (0.05, 0.15), (0.5, 1.0), (0.2, 0.8), 0.05)
    n = sites.shape[0]
    ins = np.concatenate([rng.uniform(-tilt, tilt, size=(n, 2)), np.ones((n, 1))], axis=1)
    ins /= np.linalg.norm(ins, axis=1, keepdims=True)
    lens = {
        "points": sites.copy(),
        "ins": ins,
        "height": 0.1 + rng.uniform(0.0, 0.02, size=n),
        "clip_area": (1.0 + rng.uniform(-0.1, 0.1, size=n)) / n,
        "z_of_len": 0.0,
        "z_of_receiver": float(z_of_receiver),
        "refraction_index": float(refraction_index),
        "area_scalar": 1.0,
    }
    return make_problem(sites, rho, 1e-6, f"lens_n{n}_{res}px", seed), lens


def bezier_strokes(res: int, rng: np.random.Generator, count: int = 30,
                   width_px=(2.0, 6.0), high: float = 1.0, floor: float = 0.01) -> np.ndarray:
    """`count` cubic Bezier strokes (control points U[0,1]^2, width U[width_px] pixels) at
    `high` on a `floor` background: every pixel whose centre lies within half the width of the
    curve (sampled at <= 0.25-pixel steps) is set."""
    img = np.full((res, res), floor, dtype=np.float64)
    for _ in range(count):
        P = rng.uniform(0.0, 1.0, size=(4, 2)) * res
        r = 0.5 * rng.uniform(*width_px)
        length = (np.linalg.norm(P[1] - P[0]) + np.linalg.norm(P[2] - P[1]) +
                  np.linalg.norm(P[3] - P[2]))
        t = np.linspace(0.0, 1.0, int(np.ceil(4 * length)) + 2)[:, None]
        C = ((1 - t) ** 3 * P[0] + 3 * (1 - t) ** 2 * t * P[1] + 3 * (1 - t) * t ** 2 * P[2] +
             t ** 3 * P[3])
        k = int(np.ceil(r)) + 1
        base = np.floor(C).astype(np.int64)
        for dy in range(-k, k + 1):
            for dx in range(-k, k + 1):
                px = base[:, 0] + dx
                py = base[:, 1] + dy
                ok = (px >= 0) & (px < res) & (py >= 0) & (py < res)
                d2 = (px + 0.5 - C[:, 0]) ** 2 + (py + 0.5 - C[:, 1]) ** 2
                hit = ok & (d2 <= r * r)
                img[py[hit], px[hit]] = high
    return img / img.sum()


def tilted_curves(seed: int = 5000, n_side: int = 512, res: int = 2048,
                  tilt_deg: float = 30.0) -> Problem:
    """N = 262,144 sites: a 512x512 jittered grid on a source plane tilted by 30 degrees about
    the y axis and projected straight onto the target, x -> 1/2 + (x - 1/2) cos(30 deg), so the
    sites cover only the central 87% of the target's width and the border cells must stretch;
    2048^2 target of 30 cubic Bezier strokes 2-6 px wide at 1.0 on a 0.01 floor; tol 1e-5."""
    rng = rng_for(seed)
    sites = jittered_grid(n_side, rng)
    c = math.cos(math.radians(tilt_deg))
    sites[:, 0] = 0.5 + (sites[:, 0] - 0.5) * c
    rho = bezier_strokes(res, rng)
    p = make_problem(sites, rho, 1e-5, f"curves30_tilt{tilt_deg:g}_n{n_side * n_side}_{res}px",
                     seed)
    p.site_box = (0.5 - 0.5 * c, 0.5 + 0.5 * c, 0.0, 1.0)
    return p


# --------------------------------------------------------------------------------------------
# The reference's own configuration: the seeds spline_irra.py writes for caustic_GPU
# --------------------------------------------------------------------------------------------

def triangular_lattice(point_num: int, sample_domain) -> np.ndarray:
    """The "triangle" seed lattice of initial_xy_point (utils/data_utils.py:374-410).

    nx = int(sqrt(sqrt(3)/2 * point_num * X / Y)) + 1 points along x, ny rows (made odd) with
    ny = int(nx_real * 2/sqrt(3) * Y / X) + 1; even rows hold nx points at x = j * dx, odd rows
    the two ends plus nx - 1 midpoints at x = (j + 1/2) dx.  Coordinates are offsets from 0
    (the reference adds no sample_domain origin).  point_num 160,000 on 40 x 40 gives
    nx = 373, ny = 431 and 160,978 points."""
    x0, x1, y0, y1 = (float(v) for v in sample_domain)
    xs, ys = x1 - x0, y1 - y0
    nx_real = math.sqrt(math.sqrt(3.0) / 2.0 * float(point_num) * xs / ys)
    ny_real = nx_real * 2.0 / math.sqrt(3.0) * ys / xs
    nx, ny = int(nx_real) + 1, int(ny_real) + 1
    if ny % 2 == 0:
        ny += 1
    dx, dy = xs / float(nx - 1), ys / float(ny - 1)
    rows = []
    for i in range(ny):
        y = i * dy
        if i % 2 == 0:
            x = np.arange(nx) * dx
        else:
            x = np.concatenate([[0.0], np.arange(nx - 1) * dx + dx / 2.0, [xs]])
        rows.append(np.column_stack([x, np.full(x.shape[0], y)]))
    return np.concatenate(rows).astype(np.float64)


def similarity_map(points: np.ndarray, sample_domain, proj_domain) -> np.ndarray:
    """proj_xy of initial_proj_point (utils/data_utils.py:413-435): the sample domain's centre
    to the projection domain's centre, scaled by min(proj_w / sample_w, proj_h / sample_h)."""
    sx0, sx1, sy0, sy1 = (float(v) for v in sample_domain)
    px0, px1, py0, py1 = (float(v) for v in proj_domain)
    s = min((px1 - px0) / (sx1 - sx0), (py1 - py0) / (sy1 - sy0))
    out = np.empty_like(points, dtype=np.float64)
    out[:, 0] = (points[:, 0] - 0.5 * (sx0 + sx1)) * s + 0.5 * (px0 + px1)
    out[:, 1] = (points[:, 1] - 0.5 * (sy0 + sy1)) * s + 0.5 * (py0 + py1)
    return out


def lattice_boundary_indices(points: np.ndarray, point_num: int, sample_domain):
    """(left, right, bottom, top) index arrays of find_boundary_points (utils/data_utils.py:
    437-491) for a convex lattice: every point within tol = 0.1 sqrt(area / point_num) of a side
    of the points' bounding box lies on the Delaunay boundary, and no other point does."""
    x0, x1, y0, y1 = (float(v) for v in sample_domain)
    tol = 0.1 * math.sqrt((x1 - x0) * (y1 - y0) / float(point_num))
    x, y = points[:, 0], points[:, 1]
    return tuple(np.nonzero(np.abs(v - e) <= tol)[0]
                 for v, e in ((x, x.min()), (x, x.max()), (y, y.min()), (y, y.max())))


def picture_from_text_rows(T: np.ndarray) -> np.ndarray:
    """A picture text matrix (m rows, n columns, as in data/pic_data/*.txt) in the layout
    rho[iy, ix]: element (ix, iy) is text row m - 1 - ix, column iy
    (read_pic_txt_to_array, basic_function.cuh:145-204)."""
    return np.ascontiguousarray(np.asarray(T, dtype=np.float64)[::-1, :].T)
