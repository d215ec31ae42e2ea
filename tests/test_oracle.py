"""Pins the FP64 CPU oracle with closed forms, invariants and brute force (CPU only)."""
import ctypes

import numpy as np
import pytest

from caustic_amd import workloads as W
from oracle import oracle as O

UNIT = (0.0, 1.0, 0.0, 1.0)


def _poly(*pts):
    return np.ascontiguousarray(np.array(pts, dtype=np.float64).ravel())


def _moment(fn, p, s):
    n = p.size // 2
    return fn(p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, s[0], s[1])


def test_second_moment_closed_forms():
    L = O.lib()
    sq = _poly((0, 0), (1, 0), (1, 1), (0, 1))
    # integral of x^2 + y^2 over the unit square about the origin = 2/3 (CCW positive)
    assert _moment(L.orc_fan_moment, sq, (0, 0)) == pytest.approx(2 / 3, abs=1e-15)
    assert _moment(L.orc_edge_green_moment, sq, (0, 0)) == pytest.approx(2 / 3, abs=1e-15)
    cw = _poly((0, 0), (0, 1), (1, 1), (1, 0))
    assert _moment(L.orc_fan_moment, cw, (0, 0)) == pytest.approx(-2 / 3, abs=1e-15)
    # about the centre: polar moment of a unit square = 1/6
    assert _moment(L.orc_fan_moment, sq, (0.5, 0.5)) == pytest.approx(1 / 6, abs=1e-15)


def test_fan_moment_equals_reference_edge_formula():
    L = O.lib()
    rng = np.random.default_rng(0)
    for _ in range(50):
        k = rng.integers(3, 9)
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        r = rng.uniform(0.5, 2.0, k)
        c = rng.normal(size=2) * 3
        p = _poly(*np.stack([c[0] + r * np.cos(ang), c[1] + r * np.sin(ang)], 1))
        s = rng.normal(size=2) * 2
        a = _moment(L.orc_fan_moment, p, s)
        b = _moment(L.orc_edge_green_moment, p, s)
        assert a == pytest.approx(b, rel=1e-12, abs=1e-12)


def test_regular_grid_uniform_density_is_a_fixed_point():
    """Sites at the centres of a regular grid, uniform density, nu = 1/N: g = 0 at w = 0."""
    n = 16
    c = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(c, c)
    sites = np.stack([X.ravel(), Y.ravel()], 1)
    rho = np.full((64, 64), 1.0 / 64 ** 2)
    nu = np.full(n * n, 1.0 / (n * n))
    ev = O.evaluate(sites, np.zeros(n * n), nu, rho, UNIT)
    assert np.abs(ev.mass - nu).max() < 1e-15
    assert np.abs(ev.centroid - sites).max() < 1e-13
    # f = -sum int |x - s|^2 = -N * (1/6) h^4 / h^2 ... = -(1/6) h^2 with h = 1/n (unit mass)
    assert ev.f == pytest.approx(-(1.0 / 6.0) / n ** 2, rel=1e-12)


def test_two_sites_bisector_shift():
    """Two sites on a horizontal line: a weight difference dw moves the bisector by dw/(2d)."""
    sites = np.array([[0.25, 0.5], [0.75, 0.5]])
    rho = np.full((50, 50), 1.0 / 2500)
    nu = np.array([0.5, 0.5])
    for dw in (0.0, 0.05, -0.1):
        ev = O.evaluate(sites, np.array([dw, 0.0]), nu, rho, UNIT, method="brute")
        x_b = 0.5 + dw / (2 * 0.5)
        assert ev.area[0] == pytest.approx(x_b, abs=1e-14)
        assert ev.mass[0] == pytest.approx(x_b, abs=1e-13)


def test_qhull_and_brute_force_cells_agree():
    p = W.plain_fp64(seed=3, n_side=20, res=80)
    rng = np.random.default_rng(1)
    for scale in (0.0, 1e-3, 5e-3):
        w = rng.normal(0, scale, p.n)
        a = O.evaluate(p.sites, w, p.nu, p.rho, UNIT, method="qhull")
        b = O.evaluate(p.sites, w, p.nu, p.rho, UNIT, method="brute")
        assert np.array_equal(a.exist, b.exist)
        np.testing.assert_allclose(a.mass, b.mass, rtol=0, atol=1e-16)
        np.testing.assert_allclose(a.area, b.area, rtol=0, atol=1e-15)
        assert a.f == pytest.approx(b.f, rel=1e-13)


def test_cells_tile_the_domain_and_mass_is_conserved():
    p = W.plain_fp64(seed=4, n_side=24, res=96)
    rng = np.random.default_rng(2)
    w = rng.normal(0, 3e-3, p.n)
    ev = O.evaluate(p.sites, w, p.nu, p.rho, UNIT)
    assert (~ev.exist.astype(bool)).sum() > 0  # hidden cells present
    assert ev.area.sum() == pytest.approx(1.0, abs=1e-13)
    assert ev.mass.sum() == pytest.approx(p.rho.sum(), abs=1e-13)
    assert np.all(ev.mass[~ev.exist.astype(bool)] == 0)


def test_gauge_invariance():
    p = W.plain_fp64(seed=5, n_side=16, res=64)
    rng = np.random.default_rng(3)
    w = rng.normal(0, 1e-3, p.n)
    a = O.evaluate(p.sites, w, p.nu, p.rho, UNIT)
    b = O.evaluate(p.sites, w + 0.37, p.nu, p.rho, UNIT)
    np.testing.assert_allclose(a.mass, b.mass, rtol=0, atol=1e-15)


def test_gradient_is_mass_minus_nu_by_finite_differences():
    p = W.plain_fp64(seed=6, n_side=12, res=64)
    rng = np.random.default_rng(4)
    w = rng.normal(0, 5e-4, p.n)
    ev = O.evaluate(p.sites, w, p.nu, p.rho, UNIT)
    h = 1e-7
    for i in rng.choice(p.n, 8, replace=False):
        e = np.zeros(p.n)
        e[i] = h
        fp = O.evaluate(p.sites, w + e, p.nu, p.rho, UNIT).f
        fm = O.evaluate(p.sites, w - e, p.nu, p.rho, UNIT).f
        fd = (fp - fm) / (2 * h)
        assert fd == pytest.approx(ev.mass[i] - p.nu[i], abs=2e-8)


def test_pixel_clip_areas_closed_form():
    """A single cell equal to an axis-aligned rectangle covering fractions of pixels."""
    sites = np.array([[0.5, 0.5]])
    rho = np.arange(16, dtype=np.float64).reshape(4, 4)
    rho /= rho.sum()
    # one site: the cell is the whole domain, mass = sum rho
    ev = O.evaluate(sites, np.zeros(1), np.ones(1), rho, UNIT, method="brute")
    assert ev.mass[0] == pytest.approx(1.0, abs=1e-15)
    # weighted centroid of the whole image = rho-weighted pixel centres
    X, Y = W.pixel_centres(4)
    cx = (rho * X).sum()
    cy = (rho * Y).sum()
    assert ev.centroid[0, 0] == pytest.approx(cx, abs=1e-15)
    assert ev.centroid[0, 1] == pytest.approx(cy, abs=1e-15)


def test_lbfgs_rules_on_a_quadratic():
    """The L-BFGS driver on f(w) = 1/2 w'Aw - b'w (gradient Aw - b, as mass - nu) reaches the
    solution A^-1 b.  The plain-decrease line search cannot resolve f below ~eps*|f|, so the
    gradient tolerance stays well above sqrt(eps)."""
    rng = np.random.default_rng(0)
    n = 30
    Q = rng.normal(size=(n, n))
    A = Q @ Q.T / n + np.eye(n)
    b = rng.normal(size=n)

    def ev(w):
        return 0.5 * w @ A @ w - b @ w, A @ w, 0

    r = O.lbfgs(ev, np.zeros(n), b, m=10, tol=1e-6, stop="reference", h0_threshold=0.0,
                total_iter=500)
    assert r.converged
    np.testing.assert_allclose(r.w, np.linalg.solve(A, b), atol=1e-5)


def test_lbfgs_restart_after_failed_line_search():
    """A function that never decreases along the first direction forces the 1e-4 step floor:
    the driver reverts, restarts and keeps the iterate."""
    calls = {"n": 0}

    def ev(w):
        calls["n"] += 1
        return 1.0, np.ones_like(w), 0  # f never decreases, gradient constant

    r = O.lbfgs(ev, np.zeros(4), np.zeros(4), m=3, tol=1e-12, total_iter=2)
    assert not r.converged
    np.testing.assert_array_equal(r.w, np.zeros(4))
    # 1 initial + 14 halvings per iteration (1 -> 2^-14 <= 1e-4) x 2 iterations
    assert calls["n"] == 1 + 2 * 14


def test_fresnel_normal_incidence():
    """Normal incidence, n = 1.5: R = 0.04 per surface, T1 T2 = 0.96^2 = 0.9216."""
    n = 3
    pts = np.zeros((n, 2))
    up = np.tile([0.0, 0.0, 1.0], (n, 1))
    out = O.fresnel_rate(pts, up, up, np.zeros(n), pts, 0.0, 10.0, 1.5)
    np.testing.assert_allclose(out, 0.9216, atol=1e-15)


def test_source_masses_of_a_lattice():
    """Interior cells of a regular lattice have area step_x * step_y; sum nu = area."""
    xs = np.linspace(0, 40, 11)
    X, Y = np.meshgrid(xs, xs)
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    nu = O.source_masses(pts, (0.0, 40.0, 0.0, 40.0))
    assert nu.sum() == pytest.approx(1600.0, rel=1e-14)
    interior = (pts[:, 0] > 0) & (pts[:, 0] < 40) & (pts[:, 1] > 0) & (pts[:, 1] < 40)
    np.testing.assert_allclose(nu[interior], 16.0, rtol=1e-12)


def test_reference_centroid_piece_rule_is_below_1e8():
    """The reference leaves pixel pieces of area <= 1e-10 out of the weighted centroid
    (basic_function.cuh:1153); on this workload that moves centroids by < 1e-8."""
    c1 = W.plain_fp64()
    a = O.evaluate(c1.sites, np.zeros(c1.n), c1.nu, c1.rho, c1.proj_domain)
    b = O.evaluate(c1.sites, np.zeros(c1.n), c1.nu, c1.rho, c1.proj_domain,
                   centroid_min_piece=0.0)
    assert np.abs(a.centroid - b.centroid).max() < 1e-8


def test_fresnel_brewster_angle():
    """Air -> glass at Brewster's angle tan(i) = n: Rp = 0, so T1 = 1 - Rs/2 with
    Rs = ((cos i - n cos t) / (cos i + n cos t))^2; the exit ray leaves along the inside
    direction (normal incidence on the back surface): T2 = 1 - ((n - 1)/(n + 1))^2."""
    n_idx = 1.5
    i = np.arctan(n_idx)
    t = np.arcsin(np.sin(i) / n_idx)
    wi = np.array([[np.sin(i), 0.0, np.cos(i)]])
    d = np.array([[np.sin(t), 0.0, np.cos(t)]])
    z_recv = 10.0
    tgt = np.array([[z_recv * np.tan(t), 0.0]])  # on the inside ray from the origin
    out = O.fresnel_rate(np.zeros((1, 2)), wi, d, np.zeros(1), tgt, 0.0, z_recv, n_idx)
    rs = ((np.cos(i) - n_idx * np.cos(t)) / (np.cos(i) + n_idx * np.cos(t))) ** 2
    t2 = 1.0 - ((n_idx - 1.0) / (n_idx + 1.0)) ** 2
    assert out[0] == pytest.approx((1.0 - 0.5 * rs) * t2, rel=1e-13)
    # the same rays without the Brewster cancellation: Rp > 0 lowers T1
    wi2 = np.array([[np.sin(0.5 * i), 0.0, np.cos(0.5 * i)]])
    out2 = O.fresnel_rate(np.zeros((1, 2)), wi2, d, np.zeros(1), tgt, 0.0, z_recv, n_idx)
    assert out2[0] < t2


def test_anderson_coefficients_closed_forms():
    """One residual: alpha = 1.  Orthogonal residuals of norms a, b: the minimum-norm affine
    combination has alpha proportional to (1/a^2, 1/b^2)."""
    np.testing.assert_allclose(O.anderson_coefficients([[4.0]]), [1.0])
    a, b = 2.0, 0.5
    al = O.anderson_coefficients([[a * a, 0.0], [0.0, b * b]])
    np.testing.assert_allclose(al, np.array([1 / a**2, 1 / b**2]) / (1 / a**2 + 1 / b**2),
                               rtol=1e-9)


def _linear_residual(K, b):
    def fn(w):
        return b - K @ w, 0, 3
    return fn


def test_anderson_history_one_is_the_damped_fixed_point_iteration():
    """m = 1: every step is w <- w + beta g(w)."""
    rng = np.random.default_rng(3)
    M = rng.normal(size=(5, 5))
    K = M @ M.T / 5 + np.eye(5)
    b = rng.normal(size=5)
    beta = 0.1
    fn = _linear_residual(K, b)
    w0 = np.zeros(5)
    r = O.anderson(fn, w0, fn(w0)[0], 0, 1, beta, 6, 0.0)
    w = w0.copy()
    for _ in range(6):
        w = w + beta * (b - K @ w)
    np.testing.assert_allclose(r.w, w, rtol=1e-13, atol=1e-15)
    assert r.iterations == 6 and not r.converged


def test_anderson_solves_a_linear_problem_in_n_plus_one_steps():
    """On an affine residual in R^n, Anderson acceleration with an untruncated history is
    equivalent to GMRES (Walker & Ni, SIAM J. Numer. Anal. 49 (2011)) and reaches the solution
    at iteration n + 1, far ahead of the fixed-point iteration.  The residuals are scaled to
    ~1e3 so that the Gram matrix's fixed 1e-10 regularisation stays negligible."""
    rng = np.random.default_rng(4)
    n = 5
    M = rng.normal(size=(n, n))
    K = M @ M.T / n + 0.5 * np.eye(n)
    b = 1e3 * rng.normal(size=n)
    fn = _linear_residual(K, b)
    w0 = np.zeros(n)
    beta = 1.0 / np.linalg.norm(K, 2)
    r = O.anderson(fn, w0, fn(w0)[0], 0, n + 3, beta, 50, 1e-6)
    assert r.converged and r.iterations == n + 1
    np.testing.assert_allclose(r.w, np.linalg.solve(K, b), rtol=1e-9)
    slow = O.anderson(fn, w0, fn(w0)[0], 0, 1, beta, 50, 1e-6)
    assert not slow.converged


def test_anderson_rejected_trials_restart_and_the_zero_count_rule():
    """Trials whose cells exceed max_vertices halve the step down to 1e-4 (14 halvings from 1),
    then w is kept and the history restarts.  The zero-mass count at the start of an iteration
    blocks convergence only on every 10th iteration."""
    calls = {"n": 0}

    def too_big(w):
        calls["n"] += 1
        return np.zeros(2), 0, 500

    r = O.anderson(too_big, np.ones(2), np.full(2, 1.0), 0, 3, -1.0, 2, 1e-12, max_vertices=24)
    assert calls["n"] == 2 * 14 and r.restarts == 2
    np.testing.assert_array_equal(r.w, np.ones(2))

    def tiny(w):
        return np.full(2, 1e-9), 1, 3

    r = O.anderson(tiny, np.zeros(2), np.full(2, 1e-9), 1, 3, -1.0, 10, 1e-6)
    assert r.converged and r.iterations == 2  # iteration 0 sees a zero cell; 1 does not count


def test_solid_angle_of_a_rectangle_closed_form():
    """A rectangle [0, a] x [0, b] seen from the point h below its corner subtends
    atan(a b / (h sqrt(a^2 + b^2 + h^2))); four such rectangles around the foot of the light
    add up to the centred rectangle; the fan's first vertex does not matter."""
    for a, b, h in ((1.0, 1.0, 1.0), (0.3, 2.0, 0.7), (5.0, 0.1, 3.0)):
        rect = np.array([[0, 0], [a, 0], [a, b], [0, b]], dtype=np.float64)
        want = np.arctan(a * b / (h * np.sqrt(a * a + b * b + h * h)))
        for k in range(4):
            got = O.solid_angle(np.roll(rect, k, axis=0), (0.0, 0.0, -h), 0.0)
            assert abs(got - want) < 1e-14 * max(want, 1.0), (a, b, h, k, got, want)
        # clockwise order, light above the plane: same magnitude
        assert abs(O.solid_angle(rect[::-1], (0.0, 0.0, h), 0.0) - want) < 1e-14
    a, b, h = 0.8, 0.5, 0.6
    whole = O.domain_solid_angle((-a, a, -b, b), (0.0, 0.0, -h), 0.0)
    want = 4.0 * np.arctan(a * b / (h * np.sqrt(a * a + b * b + h * h)))
    assert abs(whole - want) < 1e-13
    # the reference's atan (not atan2) and its 1e-6 denominator rule hold while every fan
    # triangle subtends less than pi: a huge square's two triangles approach pi each, their
    # denominators vanish and they are skipped (the formula's limit, kept as the reference has it)
    assert O.domain_solid_angle((-1e7, 1e7, -1e7, 1e7), (0.0, 0.0, -1.0), 0.0) < 1e-6


def test_point_light_cells_add_up_and_tend_to_areas():
    """The cells' solid angles add up to the sample domain's; a light far away weights the
    cells by area (the parallel-light masses)."""
    rng = W.rng_for(77)
    sd = (0.0, 2.0, 0.0, 1.5)
    pts = np.array([0.0, 0.0]) + W.jittered_grid(12, rng) * [2.0, 1.5]
    light = (0.7, 0.4, -1.3)
    nu, clip, d = O.point_light_source(pts, sd, light, 0.25)
    area_scalar = 3.0
    assert abs(clip.sum() - area_scalar) < 1e-12 * area_scalar
    assert abs(nu.sum() - area_scalar) < 1e-12
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, rtol=0, atol=1e-15)
    par = O.source_masses(pts, sd)
    assert np.abs(nu - par).max() > 0.05 * par.max()  # a near light reweights the cells
    far, _, _ = O.point_light_source(pts, sd, (1.0, 0.75, -1e5), 0.0)
    np.testing.assert_allclose(far, par, rtol=1e-8)


def test_gaussian_splat_conserves_interior_energy():
    """A point whose window lies inside the image deposits all its energy, symmetric about the
    pixel centre it sits on; at a corner only the window's in-image weights remain."""
    img = O.gaussian_splat(np.array([[0.55, 0.45]]), np.array([2.0]), 10, 10, (0, 1, 0, 1))
    assert abs(img.sum() - 2.0) < 1e-14
    np.testing.assert_allclose(img[4, 3:8], img[4, 3:8][::-1], rtol=1e-14)
    np.testing.assert_allclose(img[2:7, 5], img[2:7, 5][::-1], rtol=1e-14)
    corner = O.gaussian_splat(np.array([[0.05, 0.05]]), np.array([1.0]), 10, 10, (0, 1, 0, 1))
    w = np.exp(-(np.arange(-2, 3)[:, None] ** 2 + np.arange(-2, 3)[None, :] ** 2) / 0.5)
    assert abs(corner.sum() - w[2:, 2:].sum() / w.sum()) < 1e-14


def test_tilted_plane_cost_reduces_to_the_projected_sites():
    """The N = 262,144 workload's tilt: sites on a plane tilted by theta about the y axis, at
    space points p = (a cos theta, b, a sin theta), with the quadratic cost |q - p|^2 to target
    points q = (q1, q2, Z).  Brute force over the 3-D distances assigns every sample point to
    the same site as the power diagram of the projected sites p' = (a cos theta, b) with
    weights w - (Z - a sin theta)^2 -- the reduction tilted_curves relies on (its sites are
    p' + (1/2) (1 - cos theta, 0)); and on the oracle, the cells' masses agree."""
    rng = np.random.default_rng(31)
    theta, Z = np.radians(30.0), 0.7
    n = 40
    ab = rng.uniform(0.0, 1.0, (n, 2))
    w = rng.normal(0.0, 3e-3, n)
    P3 = np.stack([ab[:, 0] * np.cos(theta), ab[:, 1], ab[:, 0] * np.sin(theta)], axis=1)
    Pp = P3[:, :2]
    h = (Z - P3[:, 2]) ** 2
    g = (np.arange(150) + 0.5) / 150
    Q = np.stack(np.meshgrid(g, g, indexing="xy"), axis=-1).reshape(-1, 2)
    d3 = ((Q[:, None, 0] - P3[None, :, 0]) ** 2 + (Q[:, None, 1] - P3[None, :, 1]) ** 2
          + (Z - P3[None, :, 2]) ** 2) - w[None, :]
    d2 = ((Q[:, None, :] - Pp[None, :, :]) ** 2).sum(-1) - (w - h)[None, :]
    a3, a2 = d3.argmin(1), d2.argmin(1)
    gap = np.sort(d3, axis=1)[:, 1] - np.sort(d3, axis=1)[:, 0]
    clear = gap > 1e-12
    assert np.array_equal(a3[clear], a2[clear])
    # the same problem through the oracle: masses of the projected problem with shifted weights
    rho = np.ones((64, 64)) / 64 ** 2
    nu = np.full(n, 1.0 / n)
    e2 = O.evaluate(Pp, w - h, nu, rho, (0.0, 1.0, 0.0, 1.0))
    e2b = O.evaluate(Pp, w - h + 0.123, nu, rho, (0.0, 1.0, 0.0, 1.0))
    assert np.allclose(e2.mass, e2b.mass, atol=1e-14)  # the gauge: a constant changes no cell
    share = np.bincount(a3[clear], minlength=n) / clear.sum()
    assert np.abs(share - e2.mass).max() < 0.01  # sample-point shares against exact masses
