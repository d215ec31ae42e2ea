"""The cell and conformity-check kernel code, built for the host (tests/hostemu), against the
FP64 oracle -- the same per-site code the GPU runs, checked without a GPU."""
import os
import sys

import numpy as np
import pytest

from caustic_amd import workloads as W
from oracle import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "hostemu"))
import hostemu as E  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _built():
    E.build()


def _check(p, w, fp32, mass_tol, cen_tol):
    orc = O.evaluate(p.sites, w, p.nu, p.rho, p.proj_domain, centroid_min_piece=0.0)
    e = E.evaluate(p, w, fp32=fp32)
    assert e["fails"] == 0
    assert np.array_equal(e["exist"], orc.exist.astype(bool))
    dm = np.abs(e["mass"] - orc.mass) / np.maximum(orc.mass, p.nu)
    assert dm.max() <= mass_tol, dm.max()
    big = e["exist"] & (orc.mass >= 1e-3 * p.nu)
    assert np.abs(e["centroid"][big] - orc.centroid[big]).max() <= cen_tol
    return e, orc


@pytest.mark.parametrize("fp32,mass_tol,cen_tol", [(False, 1e-12, 1e-12), (True, 1e-6, 1e-6)])
@pytest.mark.parametrize("scale", [0.0, 2e-4])
def test_hostemu_matches_oracle(fp32, mass_tol, cen_tol, scale):
    p = W.plain_fp64()
    w = np.random.default_rng(7).normal(0.0, scale, p.n)
    e, orc = _check(p, w, fp32, mass_tol, cen_tol)
    if not fp32:
        assert abs(e["f"] - orc.f) <= 1e-12 * abs(orc.f)


def test_hostemu_large_shapes_mixed():
    p = W.large_shapes(n_side=128, res=512)
    w = np.random.default_rng(5).normal(0.0, 4e-5, p.n)
    _check(p, w, True, 1e-6, 1e-6)


def _displaced(p):
    x, y = p.sites[:, 0], p.sites[:, 1]
    return 0.08 * (x - 0.5) ** 2 + 0.05 * np.sin(6 * y)  # smooth, large: displaced cells


def test_hostemu_displaced_cells_need_repairs():
    """Weights that move cells several site spacings away: the first rings miss neighbours,
    the conformity check catches it and the repaired cells equal the oracle's."""
    p = W.plain_fp64(seed=9, n_side=32, res=128)
    w = _displaced(p)
    orc = O.evaluate(p.sites, w, p.nu, p.rho, p.proj_domain, centroid_min_piece=0.0)
    e = E.evaluate(p, w, fp32=False, ring_init=1)
    assert e["repairs"] > 0 and e["fails"] == 0
    np.testing.assert_allclose(e["mass"], orc.mass, rtol=0, atol=1e-12 * p.nu.max())


@pytest.mark.parametrize("fp32", [False, True])
def test_hostemu_adjacency_candidates_follow_moving_weights(fp32):
    """Candidates from the previous evaluation's cells: walking the weights from zero to the
    displaced field in steps changes the adjacency (edge flips, cells that vanish and return),
    the two-ring candidates, the check and its repairs (bucket rings) keep every evaluation equal
    to the oracle's."""
    p = W.plain_fp64(seed=9, n_side=32, res=128)
    w1 = _displaced(p)
    s = E.Session(p, fp32=fp32, ring_init=1, max_repairs=4)
    tol = 1e-6 if fp32 else 1e-12
    rebuilt = 0
    for k, a in enumerate([0.0, 0.1, 0.25, 0.5, 0.55, 0.8, 1.0, 0.7, 1.0]):
        w = a * w1
        e = s.evaluate(w)
        assert e["fails"] == 0, (k, a)
        orc = O.evaluate(p.sites, w, p.nu, p.rho, p.proj_domain, centroid_min_piece=0.0)
        assert np.array_equal(e["exist"], orc.exist.astype(bool)), (k, a)
        dm = np.abs(e["mass"] - orc.mass) / np.maximum(orc.mass, p.nu)
        assert dm.max() <= tol, (k, a, dm.max())
        rebuilt += e["rebuilt"] if k > 0 else 0
    assert rebuilt > 0  # the adjacency did change and was repaired
    s.close()


@pytest.mark.parametrize("fp32,accept", [(False, "decrease"), (True, "derivative")])
def test_hostemu_lbfgs_converges(fp32, accept):
    """The L-BFGS rules driving the host-built evaluation reach max-rel 1e-5 on the N = 4,096
    workload; FP32 geometry needs the derivative-aware acceptance."""
    p = W.plain_fp64()
    s = E.Session(p, fp32=fp32, max_repairs=3)

    def ev(w):
        e = s.evaluate(w)
        if e["fails"]:
            return np.inf, np.full(p.n, np.inf), 0
        return e["f"], e["mass"], 0

    r = O.lbfgs(ev, np.zeros(p.n), p.nu, m=10, tol=1e-5, stop="maxrel", h0_threshold=0.0,
                total_iter=3000, accept=accept)
    assert r.converged
    orc = O.evaluate(p.sites, r.w, p.nu, p.rho, p.proj_domain)
    assert O.max_rel_error(orc.mass, p.nu) < 1e-5 + 1e-6


def _patch_laplacian(n, seed):
    """Off-diagonal couplings of a random patch graph (a jittered grid's neighbours), as the
    device accumulates them: each row's entries carry their own rounding (a[i][j] != a[j][i])."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(n)))
    a = np.zeros((n, n))
    for i in range(n):
        for j in (i + 1, i + side):
            if j < n and (j != i + 1 or (i + 1) % side):
                v = rng.uniform(0.1, 3.0)
                a[i, j] = -v * (1 + 1e-13 * rng.normal())
                a[j, i] = -v * (1 + 1e-13 * rng.normal())
    return a


@pytest.mark.parametrize("n,seed", [(2, 1), (7, 2), (64, 3), (200, 4)])
def test_coarse_inverse_matches_numpy(n, seed):
    """The Newton preconditioner's coarse inverse: the pseudo-inverse of A_c, the symmetrised
    Laplacian (diagonal = minus the off-diagonal row sums) -- against numpy.linalg.pinv of the
    same matrix built independently, and through its defining properties."""
    a = _patch_laplacian(n, seed)
    got = E.coarse_inverse(a)
    sym = 0.5 * (a + a.T)
    np.fill_diagonal(sym, 0.0)
    lap = sym - np.diag(sym.sum(axis=1))
    want = np.linalg.pinv(lap)  # the graph is connected: null space = the constants
    assert np.allclose(got, got.T, rtol=0, atol=1e-14 * np.abs(want).max())
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    # nothing along the constant vector; an inverse on its complement
    one = np.ones(n)
    assert np.abs(got @ one).max() <= 1e-12 * np.abs(want).max() * n
    x = np.random.default_rng(seed).normal(size=n)
    x -= x.mean()
    assert np.allclose(got @ (lap @ x), x, rtol=0, atol=1e-9 * np.abs(x).max())


def test_coarse_inverse_of_zero_and_disconnected():
    """A zero matrix (an inactive problem) and a matrix that is not positive definite after the
    rank-one term (two disconnected groups: a second null vector) leave zeros -- no coarse
    correction rather than a wrong one."""
    assert not E.coarse_inverse(np.zeros((5, 5))).any()
    a = np.zeros((4, 4))
    a[0, 1] = a[1, 0] = -1.0
    a[2, 3] = a[3, 2] = -1.0
    assert not E.coarse_inverse(a).any()


def test_coarse_inverse_wide_band_and_disconnected_banded():
    """A randomly renumbered patch graph (bandwidth ~n: the dense route) still gives the
    pseudo-inverse; a banded graph with two components (the grounded matrix is singular) gives
    zeros."""
    n = 100
    a = _patch_laplacian(n, 5)
    perm = np.random.default_rng(6).permutation(n)
    ap = a[np.ix_(perm, perm)]
    got = E.coarse_inverse(ap)
    sym = 0.5 * (ap + ap.T)
    np.fill_diagonal(sym, 0.0)
    want = np.linalg.pinv(sym - np.diag(sym.sum(axis=1)))
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    b = _patch_laplacian(n, 7)
    b[:50, 50:] = 0.0
    b[50:, :50] = 0.0
    assert not E.coarse_inverse(b).any()


def test_cover_check_rejects_a_consistent_double_cover():
    """The last step of the conformity argument (edges matched => the cells cover the target a
    constant number of times; sum of areas = |D| => once; power_diagram.cpp:8-134 semantics:
    the cells tile the domain).  Four sites on y = 1/2 whose candidates pair them (A-B, C-D):
    each pair tiles the square with one matched bisector edge, so the edge check finds nothing,
    but the square is covered twice.  The cover check flags the evaluation, and the next one
    (bucket-ring candidates) tiles the square."""
    from caustic_amd.workloads import Problem

    xs = [0.125, 0.375, 0.625, 0.875]
    p = Problem(sites=np.array([[x, 0.5] for x in xs]), nu=np.full(4, 0.25),
                rho=np.full((8, 8), 1.0 / 64))
    s = E.Session(p, False, 2, 12, 1e-4)
    try:
        s.set_adjacency([[1], [0], [3], [2]])
        out = s.evaluate(np.zeros(4))
        st = s.state()
        # every bisector edge present on both of its cells: nothing for the edge check
        assert not st["redo"].any() and out["repairs"] == 0
        np.testing.assert_allclose(out["mass"], [0.25, 0.75, 0.75, 0.25], atol=1e-14)
        assert not out["covered"] and out["fails"] >= 1
        out2 = s.evaluate(np.zeros(4))
        assert out2["covered"] and out2["fails"] == 0
        np.testing.assert_allclose(out2["mass"], [0.25] * 4, atol=1e-14)
    finally:
        s.close()


def test_xcd_slot_is_a_bijection_with_contiguous_ranges_per_xcd():
    """Workgroup w runs on XCD w % 8; xcd_slot must renumber the G workgroups onto [0, G) with
    each XCD's slots contiguous and in increasing order of w."""
    for G in list(range(1, 70)) + [255, 256, 257, 512, 1023, 1024, 4096 + 3]:
        slots = [E.xcd_slot(w, G) for w in range(G)]
        assert sorted(slots) == list(range(G)), G
        for x in range(8):
            mine = [slots[w] for w in range(x, G, 8)]
            assert mine == list(range(mine[0], mine[0] + len(mine))) if mine else True, (G, x)


@pytest.mark.parametrize("fp32", [True, False])
@pytest.mark.parametrize("shape", [(5, 77), (3, 32), (4, 8), (2, 1), (6, 100)])
def test_density_groups_reproduce_the_sequential_prefix(fp32, shape):
    """The compact density groups (a prefix per group of 8 or 4 pixels plus the group's
    densities) give, at every pixel, the same bits as summing the row's chunk sequentially --
    rows whose length is no multiple of the group or the chunk included."""
    rho = np.random.default_rng(11).uniform(0.0, 3.0, shape) ** 3
    rho[0, : shape[1] // 2] = 0.0
    assert E.prefix_mismatches(rho, fp32) == 0
