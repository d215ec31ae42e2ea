"""The oracle and the seed/format conventions pinned against the reference's shipped run
(data/point_data/ustc_centerL_tri_standard_sp_parallel_light_rl, data/weights/..., the only
output of this path the reference itself produced): N = 160,978 sites of the triangular lattice,
a 640 x 640 target with 76.6 % zero pixels, weights in [-512, 2588] at 6 significant digits.

The fixture (tests/golden/ustc_centerL_t0.npz) is written by tests/golden/make_ustc_fixture.py
from the reference's data files only."""
import hashlib

import numpy as np
import pytest

import shipped
from caustic_amd import workloads as W
from oracle import oracle as O


@pytest.fixture(scope="module")
def S():
    return shipped.load()


def test_lattice_reproduces_shipped_in_txt(S):
    """initial_xy_point's lattice (utils/data_utils.py:374-410), written with "%f" to points.txt
    (utils/basic_function.py:775) and printed back by C++ ostream into in.txt (main.cu:312-324,
    every cell exists): the text reproduces the shipped in.txt byte for byte."""
    assert S.n == S.in_lines == 160978
    rows = W.triangular_lattice(160000, (0.0, 40.0, 0.0, 40.0))
    # rows alternate 373 and 374 points; 216 * 373 + 215 * 374 = 160,978 (odd row count 431)
    assert np.count_nonzero(rows[:, 1] == 0.0) == 373
    assert np.count_nonzero(rows[:, 0] == 0.0) == 431
    text = "".join("%g %g\n" % (x, y) for x, y in S.points)
    assert hashlib.sha256(text.encode()).hexdigest() == S.in_sha256
    # a lattice off by one row or column would not reproduce it
    wrong = W.triangular_lattice(160000, (0.0, 40.0, 0.0, 40.5))
    assert wrong.shape[0] != S.n or not np.allclose(wrong, rows)


def test_similarity_map_closed_form(S):
    """initial_proj_point (utils/data_utils.py:413-435) at the shipped domains: scale
    min(160/40, 160/40) = 4 about the centres, proj = 4 (p - (20, 20)) + (80, 20)."""
    p = S.points
    expect = 4.0 * (p - np.array([20.0, 20.0])) + np.array([80.0, 20.0])
    np.testing.assert_allclose(W.similarity_map(p, S.sample_domain, S.proj_domain), expect,
                               rtol=0, atol=1e-12)
    corners = W.similarity_map(np.array([[0.0, 0.0], [40.0, 40.0]]), S.sample_domain,
                               S.proj_domain)
    np.testing.assert_allclose(corners, [[0.0, -60.0], [160.0, 100.0]], atol=1e-12)
    # sites fill proj_domain edge to edge: the corners the reference pushes out by one unit
    # (main.cu:149-187) lie outside the target, where no cell is clipped
    assert S.proj_xy[:, 0].min() == 0.0 and S.proj_xy[:, 0].max() == 160.0


def test_boundary_indices_match_delaunay(S):
    """find_boundary_points (utils/data_utils.py:437-491): points on a boundary edge of the
    Delaunay triangulation (scipy's qhull, a library routine) and within tol of a side of the
    bounding box -- on the convex lattice exactly the closed-form rows and columns."""
    from scipy.spatial import Delaunay

    p = S.points
    tri = Delaunay(p).simplices
    e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    ue, cnt = np.unique(e, axis=0, return_counts=True)
    on_boundary = np.zeros(p.shape[0], dtype=bool)
    on_boundary[ue[cnt == 1].ravel()] = True
    tol = 0.1 * np.sqrt(40.0 * 40.0 / 160000)
    x, y = p[:, 0], p[:, 1]
    masks = [np.isclose(x, x.min(), atol=tol), np.isclose(x, x.max(), atol=tol),
             np.isclose(y, y.min(), atol=tol), np.isclose(y, y.max(), atol=tol)]
    closed = W.lattice_boundary_indices(p, 160000, S.sample_domain)
    for m, c in zip(masks, closed):
        np.testing.assert_array_equal(np.nonzero(m & on_boundary)[0], c)
    assert [len(c) for c in closed] == [431, 431, 373, 373]


def test_shipped_picture_and_weights(S):
    """The shipped target (640 x 640, max 200, sum 17,181,952, 76.6 % zero) and the Sigma w = 0
    gauge of the shipped weights (g sums to zero, so L-BFGS from w = 0 keeps sum(w) = 0)."""
    assert S.text.shape == (640, 640)
    assert S.text.max() == 200 and S.text.sum() == 17181952
    assert abs(np.mean(S.text == 0) - 0.7657) < 1e-4
    assert abs(S.rho.sum() - 1600.0) < 1e-9
    assert abs(S.weights.sum()) < 1.0 and abs(S.weights.mean()) < 1e-5 * np.abs(S.weights).max()
    assert S.weights.min() < -500 and S.weights.max() > 2500


def _evaluate_at_shipped_weights(S, rho, nu):
    ev = O.evaluate(S.proj_xy, S.weights, nu, rho, S.proj_domain)
    return ev, shipped.centroid_agreement(ev.centroid, S.centroids)


def test_oracle_reproduces_shipped_run(S):
    """The oracle at the shipped 6-digit t = 0 weights, with the source masses of the lattice's
    Voronoi cells (ot_op.cuh:90-206), against the shipped run:
      * every one of the 160,978 cells exists and holds mass, and |mass - nu|_2 = 0.08: the
        shipped stopping rule |mass - nu|_2 < 0.01 (ot_op.cuh:510) held at full precision, and
        the weights' 6-digit truncation (+-0.005 on values ~2000 moves a bisector by ~0.006,
        a few percent of a small cell's mass) leaves about this much;
      * the rho-weighted centroids (compute_center, basic_function.cuh:1056-1191) reproduce
        0/out.txt: half of them within 3e-4 units (domain 160), 99 % within 1e-2 (the worst are
        large cells lying mostly on zero density, whose weighted centroid hangs on a few bright
        pixels).
    With the picture transposed or its text rows taken in file order (the orientation
    read_pic_txt_to_array undoes, basic_function.cuh:145-204), the residual is 29-40 with tens
    of thousands of empty cells and the centroids' 99th percentile misses by ~2 units: the
    masses pin orientation and normalisation (at fixed weights a cell inside a uniform bright
    region has the same centroid under any image, so the centroids alone would not)."""
    nu = O.source_masses(S.points, S.sample_domain)
    assert abs(nu.sum() - S.area_scalar) < 1e-9
    ev, (med, p99, mx) = _evaluate_at_shipped_weights(S, S.rho, nu)
    assert ev.exist.all() and int(np.count_nonzero(ev.mass == 0.0)) == 0
    assert 0.01 < np.linalg.norm(ev.mass - nu) < 0.2
    assert med < 1e-3 and p99 < 1e-2 and mx < 0.2, (med, p99, mx)
    assert ev.max_len <= 24
    for wrong in (W.picture_from_text_rows(S.text.T), np.ascontiguousarray(S.text.T)):
        pic = wrong / wrong.sum() * S.area_scalar
        evw, (_, wp99, _) = _evaluate_at_shipped_weights(S, pic, nu)
        assert np.linalg.norm(evw.mass - nu) > 10.0 and wp99 > 0.5
        assert np.count_nonzero(evw.mass == 0.0) > 10000
