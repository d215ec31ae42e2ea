"""GPU parity of the HIP OT path against the FP64 CPU oracle (through the C ABI)."""
import numpy as np
import pytest

from caustic_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _batch(problems, **kw):
    from caustic_amd.ot import OTBatch

    return OTBatch(problems, **kw)


def _exact(p, w):
    """Oracle evaluation keeping every pixel piece in the weighted centroid (the GPU integrates
    the centroid exactly; the reference's 1e-10 piece rule is checked separately)."""
    return O.evaluate(p.sites, w, p.nu, p.rho, p.proj_domain, centroid_min_piece=0.0)


def _compare(ev_gpu, ev_orc, nu, mass_tol, cen_tol):
    """Masses within mass_tol relative to each cell's own mass for every cell holding at least
    a tenth of its source mass, and relative to max(mass, nu) for all (slivers judged against
    the mass scale of the problem); centroids of cells holding at least 1e-3 nu within
    cen_tol."""
    err = np.abs(ev_gpu["mass"] - ev_orc.mass)
    held = ev_orc.mass >= 0.1 * nu
    own = (err[held] / ev_orc.mass[held]).max() if held.any() else 0.0
    assert own <= mass_tol, f"mass error {own:.3e} (relative to the cell's own mass)"
    dm = err / np.maximum(ev_orc.mass, nu)
    assert dm.max() <= mass_tol, f"mass error {dm.max():.3e} (relative to max(mass, nu))"
    ex = ev_gpu["exist"]
    assert np.array_equal(ex, ev_orc.exist.astype(bool))
    big = ex & (ev_orc.mass >= 1e-3 * nu)
    dc = np.abs(ev_gpu["centroid"][big] - ev_orc.centroid[big]).max() if big.any() else 0.0
    assert dc <= cen_tol, f"centroid error {dc:.3e}"
    return dm.max()


@pytest.fixture(scope="module")
def c1():
    return W.plain_fp64()


@pytest.mark.parametrize("precision,mass_tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_eval_zero_weights(gpu_available, c1, precision, mass_tol):
    print(f"[case] eval c1 w=0 {precision}", flush=True)
    w = np.zeros(c1.n)
    orc = _exact(c1, w)
    with _batch([c1], precision=precision) as b:
        ev = b.evaluate(w)
    _compare(ev, orc, c1.nu, mass_tol, 1e-11 if precision == "fp64" else 1e-6)
    f_tol = 1e-12 if precision == "fp64" else 1e-7
    assert abs(ev["f"][0] - orc.f) <= f_tol * abs(orc.f)
    assert abs(ev["mass"].sum() - 1.0) < (1e-12 if precision == "fp64" else 1e-7)


@pytest.mark.parametrize("precision,mass_tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_eval_random_weights_hidden_cells(gpu_available, c1, precision, mass_tol):
    print(f"[case] eval c1 random w {precision}", flush=True)
    rng = np.random.default_rng(7)
    w = rng.normal(0.0, 2e-4, c1.n)
    orc = _exact(c1, w)
    assert (~orc.exist.astype(bool)).sum() > 10  # the case exercises hidden cells
    with _batch([c1], precision=precision) as b:
        ev = b.evaluate(w)
    _compare(ev, orc, c1.nu, mass_tol, 1e-11 if precision == "fp64" else 1e-6)
    rel_f = abs(ev["f"][0] - orc.f) / abs(orc.f)
    assert rel_f <= (1e-12 if precision == "fp64" else 1e-7)


@pytest.mark.parametrize("device_loop", [False, True])
def test_solve_c1_maxrel(gpu_available, c1, device_loop):
    print(f"[case] solve c1 fp64 maxrel 1e-4 device_loop={device_loop}", flush=True)
    with _batch([c1], precision="fp64") as b:
        w, rep = b.solve(history=10, tol=1e-4, stop="maxrel", h0_threshold=0.0, max_iter=5000,
                         device_loop=device_loop)
    assert rep[0]["converged"] == 1, rep
    orc = O.evaluate(c1.sites, w, c1.nu, c1.rho, c1.proj_domain)
    mr = O.max_rel_error(orc.mass, c1.nu)
    assert mr < 1.0001e-4, mr
    assert abs(w.mean()) < 1e-12  # sum(g) = 0 keeps the gauge sum(w) = 0


def test_device_and_host_loops_follow_the_same_iterates(gpu_available, c1):
    """Both drivers apply the same rules to the same FP64 evaluations; only the summation order
    of the dot products differs.  So the iterates agree to rounding -- until a line-search
    decision falls within that rounding (a trial whose f is within an ulp or two of the current
    one), after which the drivers take different, equally valid paths, which shows in their
    evaluation counts.  Checked after 5, 10, ... 40 iterations: every stretch with equal
    evaluation counts agrees to 1e-9 of the weights' scale, and the first 10 iterations do."""
    print("[case] device vs host loop, 5..40 iterations", flush=True)
    agreed = 0
    for k in range(5, 41, 5):
        out = []
        for dl in (False, True):
            with _batch([c1], precision="fp64") as b:
                w, rep = b.solve(history=10, tol=1e-12, stop="maxrel", h0_threshold=0.0,
                                 max_iter=k, device_loop=dl, raise_on_fail=False)
            out.append((w, rep[0]))
        assert out[0][1]["iterations"] == out[1][1]["iterations"] == k
        if out[0][1]["evaluations"] != out[1][1]["evaluations"]:
            break
        scale = np.abs(out[0][0]).max()
        assert np.abs(out[0][0] - out[1][0]).max() <= 1e-9 * scale, k
        agreed = k
    assert agreed >= 10, agreed


@pytest.mark.parametrize("device_loop", [False, True])
def test_solve_c1_mixed_precision(gpu_available, c1, device_loop):
    print(f"[case] solve c1 mixed maxrel 1e-5 device_loop={device_loop}", flush=True)
    with _batch([c1], precision="mixed") as b:
        w, rep = b.solve(history=10, tol=1e-5, stop="maxrel", h0_threshold=0.0, max_iter=5000,
                         device_loop=device_loop)
    assert rep[0]["converged"] == 1, rep
    orc = O.evaluate(c1.sites, w, c1.nu, c1.rho, c1.proj_domain)
    # FP32 geometry and image: the FP64 oracle's masses at the returned weights meet the stopping
    # value to within 1e-6 relative
    assert O.max_rel_error(orc.mass, c1.nu) < 1e-5 + 1e-6


def test_solve_batch_of_uneven_problems(gpu_available):
    print("[case] batched solve of 3 uneven problems", flush=True)
    probs = [W.plain_fp64(seed=21, n_side=24, res=96), W.plain_fp64(seed=22, n_side=48, res=192),
             W.plain_fp64(seed=23, n_side=32, res=128)]
    with _batch(probs) as b:
        w, rep = b.solve(history=10, tol=1e-4, stop="maxrel", h0_threshold=0.0, max_iter=5000)
    assert all(r["converged"] for r in rep), rep
    for k, p in enumerate(probs):
        wk = w[int(b.offsets[k]):int(b.offsets[k + 1])]
        orc = O.evaluate(p.sites, wk, p.nu, p.rho, p.proj_domain)
        assert O.max_rel_error(orc.mass, p.nu) < 1.0001e-4


def test_batch_matches_single(gpu_available, c1):
    print("[case] batch of 3 vs single", flush=True)
    small = W.plain_fp64(seed=11, n_side=24, res=96)
    other = W.plain_fp64(seed=12, n_side=40, res=128)
    rng = np.random.default_rng(3)
    ws = [rng.normal(0, 1e-4, p.n) for p in (small, c1, other)]
    with _batch([small, c1, other]) as b:
        ev = b.evaluate(np.concatenate(ws))
    for k, p in enumerate((small, c1, other)):
        with _batch([p]) as s:
            ev1 = s.evaluate(ws[k])
        sl = slice(int(b.offsets[k]), int(b.offsets[k + 1]))
        np.testing.assert_allclose(ev["mass"][sl], ev1["mass"], rtol=0, atol=1e-15)
        assert abs(ev["f"][k] - ev1["f"][0]) <= 1e-14 * abs(ev1["f"][0])


def test_eval_large_shapes_full_size(gpu_available):
    """N = 65,536 sites, 1024^2 target of disks/rectangles at 100:1 contrast."""
    print("[case] eval c2 full size", flush=True)
    p = W.large_shapes()
    rng = np.random.default_rng(5)
    w = rng.normal(0.0, 2e-5, p.n)
    orc = _exact(p, w)
    for prec, tol in (("fp64", 1e-9), ("mixed", 1e-6)):
        with _batch([p], precision=prec) as b:
            ev = b.evaluate(w)
        _compare(ev, orc, p.nu, tol, 1e-8 if prec == "fp64" else 1e-6)


@pytest.fixture(scope="module")
def c5():
    return W.tilted_curves()


def test_eval_tilted_curves_full_size(gpu_available, c5):
    """N = 262,144 sites on a 30-degree tilted source plane, 2048^2 target of Bezier strokes."""
    print("[case] eval tilted curves N=262144 2048^2", flush=True)
    rng = np.random.default_rng(6)
    w = rng.normal(0.0, 1e-6, c5.n)
    orc = _exact(c5, w)
    assert (~orc.exist.astype(bool)).sum() > 0  # hidden cells at this weight spread
    for prec, tol in (("fp64", 1e-9), ("mixed", 1e-6)):
        with _batch([c5], precision=prec) as b:
            ev = b.evaluate(w)
        _compare(ev, orc, c5.nu, tol, 1e-8 if prec == "fp64" else 1e-6)
        # mixed: the image is held in FP32, whose rounding of the two grey levels shifts the
        # total by up to ~3e-8
        assert abs(ev["mass"].sum() - 1.0) < (1e-10 if prec == "fp64" else 1e-7)


def test_solve_tilted_curves_coarse_to_fine(gpu_available, c5):
    """The full N = 262,144 solve to max-rel 1e-5 in mixed precision from the coarse-to-fine
    start (64, 128, 256 levels), each level L-BFGS to 1e-2 and damped Newton steps after; the
    FP64 oracle confirms the masses at the returned weights."""
    from caustic_amd import warmstart as WS

    print("[case] solve tilted curves N=262144 coarse-to-fine", flush=True)
    ws, _, rep = WS.multiscale_solve([c5], levels=(64, 128, 256), tol=1e-5, max_iter=20000,
                                     log=lambda s: print(s, flush=True), verbose=1,
                                     newton_from=1e-2)
    assert rep[0]["converged"], rep
    orc = O.evaluate(c5.sites, ws[0], c5.nu, c5.rho, c5.proj_domain)
    assert O.max_rel_error(orc.mass, c5.nu) < 1e-5 + 1e-6


@pytest.mark.parametrize("int_lanes,repair_blocks,spb", [(4, 128, 0.5), (16, 1024, 2.0)])
def test_launch_shapes_give_the_same_cells(gpu_available, c1, int_lanes, repair_blocks, spb):
    """Every launch shape the autotuner may pick computes the same evaluation (the lane count
    only reorders the FP64 sums): masses within 1e-12 of max(mass, nu) of the defaults'.  A
    solve under the shape converges, and the oracle confirms its cells and its max-rel error.
    (The L-BFGS trajectories of two shapes need not coincide step for step: 30 iterations with
    the bucket density at 2 have ended elsewhere than with the default.)"""
    print(f"[case] launch shape lanes={int_lanes} repair_blocks={repair_blocks} spb={spb}",
          flush=True)
    rng = np.random.default_rng(7)
    w = rng.normal(0.0, 2e-4, c1.n)
    with _batch([c1], precision="fp64") as b:
        ref = b.evaluate(w)
    tuning = {"int_lanes": int_lanes, "repair_blocks": repair_blocks, "sites_per_bucket": spb}
    with _batch([c1], precision="fp64", tuning=tuning) as b:
        ev = b.evaluate(w)
        wt, rep = b.solve(None, tol=1e-4)
        evt = b.evaluate(wt)
    dm = np.abs(ev["mass"] - ref["mass"]) / np.maximum(ref["mass"], c1.nu)
    assert dm.max() < 1e-12
    assert np.array_equal(ev["exist"], ref["exist"])
    assert rep[0]["converged"]
    orc = _exact(c1, wt)
    _compare(evt, orc, c1.nu, 1e-10, 1e-11)
    assert O.max_rel_error(orc.mass, c1.nu) < 1e-4


def test_render_and_cell_polygons_match_oracle(gpu_available, c1):
    """The diagnostic render (each nu_i spread uniformly over its cell) against the oracle's
    restatement of the reference's render kernel; the cell polygons tile the domain."""
    print("[case] render + cell polygons c1 fp64", flush=True)
    rng = np.random.default_rng(8)
    w = rng.normal(0.0, 2e-4, c1.n)
    orc = O.evaluate(c1.sites, w, c1.nu, c1.rho, c1.proj_domain)
    ref = O.render(orc.cells, c1.nu, c1.rho.shape, c1.proj_domain)
    with _batch([c1], precision="fp64") as b:
        img = b.render(w)[0]
        counts, verts = b.cell_polygons(w)
    assert img.shape == c1.rho.shape
    # pixel pieces of a cell are clipped in another order than the oracle's: slivers differ by
    # FP64 cancellation, ~1e-10 of the largest pixel
    np.testing.assert_allclose(img, ref, rtol=0, atol=1e-9 * ref.max())
    ex = orc.exist.astype(bool)
    assert abs(img.sum() - c1.nu[ex].sum()) < 1e-12
    assert np.array_equal(counts >= 3, ex)
    areas = np.zeros(c1.n)
    for i in np.nonzero(counts)[0]:
        x, y = verts[i, :counts[i], 0], verts[i, :counts[i], 1]
        areas[i] = 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)
    assert (areas[ex] > 0).all()  # counter-clockwise
    np.testing.assert_allclose(areas, orc.area, rtol=0, atol=1e-14)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cell_solid_angles_match_oracle(gpu_available, c1, precision):
    """Point-light source weights: the solid angle of each cell seen from a light below the
    plane (cot_cell_solid_angles) against the oracle's restatement of compute_solid_angle on
    its own cells; hidden cells get 0 and the cells add up to the domain's solid angle."""
    print(f"[case] cell solid angles c1 {precision}", flush=True)
    rng = np.random.default_rng(9)
    w = rng.normal(0.0, 2e-4, c1.n)
    light, z = (0.35, 0.6, -0.45), 0.1
    orc = O.evaluate(c1.sites, w, c1.nu, c1.rho, c1.proj_domain)
    ref = O.cell_solid_angles(orc.cells, light, z)
    with _batch([c1], precision=precision) as b:
        om = b.cell_solid_angles(light, z, w)
    ex = orc.exist.astype(bool)
    assert (~ex).any() and np.all(om[~ex] == 0.0)
    # FP64 vertices in both precisions; vertex order and fan start differ from the oracle's
    np.testing.assert_allclose(om, ref, rtol=1e-11, atol=1e-16)
    whole = O.domain_solid_angle(c1.proj_domain, light, z)
    assert abs(om.sum() - whole) < 1e-11 * whole
