"""The wave-cooperative repair construction (one wave per listed cell, the default) against the
one-lane-per-cell construction: both build the same cells, so evaluations agree to the last
bits the arithmetic order allows, and solves take the same path."""
import numpy as np
import pytest

from caustic_amd import workloads as W

pytestmark = pytest.mark.gpu


def _eval(problems, w, mode, **kw):
    from caustic_amd.ot import OTBatch

    with OTBatch(problems, repair_mode=mode, **kw) as b:
        ev = b.evaluate(w)
        st = b.stats()
        counts, verts = b.cell_polygons(w)
    return ev, st, counts, verts


def _agree(a, b, tag):
    ev_w, st_w, n_w, v_w = a
    ev_l, st_l, n_l, v_l = b
    assert st_w["repair_launches"] > 0 and st_l["repair_launches"] > 0, (tag, st_w, st_l)
    assert np.array_equal(ev_w["exist"], ev_l["exist"]), tag
    scale = np.maximum(np.abs(ev_l["mass"]), 1e-300)
    dm = np.abs(ev_w["mass"] - ev_l["mass"]) / scale
    held = ev_l["mass"] > 0
    assert dm[held].max() <= 1e-12, f"{tag}: mass differs by {dm[held].max():.3e}"
    assert np.abs(ev_w["f"] - ev_l["f"]).max() <= 1e-13 * np.abs(ev_l["f"]).max(), tag
    assert np.abs(ev_w["centroid"] - ev_l["centroid"]).max() <= 1e-12, tag
    assert np.array_equal(n_w, n_l), f"{tag}: vertex counts differ"
    k = np.arange(v_w.shape[1])[None, :] < n_w[:, None]
    dv = np.abs(v_w - v_l)[k].max()
    assert dv <= 1e-13, f"{tag}: vertices differ by {dv:.3e}"
    print(f"{tag}: repairs {st_w['repair_launches']} (wave) {st_l['repair_launches']} (lane), "
          f"exact masses {np.mean(ev_w['mass'] == ev_l['mass']):.4f}, vertex diff {dv:.1e}",
          flush=True)


@pytest.fixture(scope="module")
def c1():
    return W.plain_fp64()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wave_repair_hidden_cells(gpu_available, c1, precision):
    """Random weights hide cells and leave conformity failures for the repair rounds."""
    print(f"[case] wave vs lane repair, c1 random w, {precision}", flush=True)
    w = np.random.default_rng(7).normal(0.0, 2e-4, c1.n)
    _agree(_eval([c1], w, "wave", precision=precision),
           _eval([c1], w, "lane", precision=precision), f"hidden {precision}")


def test_wave_repair_bucket_rings(gpu_available, c1):
    """A cold start with ring 1 at widely spread weights: the full pass misses neighbours, the
    two adjacency rounds leave failures, and the repairs go on to the bucket rings, grown
    2r + 1 up to ring 31 (rings of more than 64 buckets).  The CPU build of the lane
    construction takes 6 repair rounds on this case."""
    print("[case] wave vs lane repair, ring_init 1, w ~ N(0, 1.6e-3)", flush=True)
    w = np.random.default_rng(100).normal(0.0, 1.6e-3, c1.n)
    a = _eval([c1], w, "wave", ring_init=1)
    b = _eval([c1], w, "lane", ring_init=1)
    assert a[1]["repair_rounds"] >= 3 and b[1]["repair_rounds"] >= 3, (a[1], b[1])
    _agree(a, b, "bucket rings")


def test_wave_repair_batch(gpu_available):
    """Several problems of different sizes in one context (problem offsets in the work list)."""
    print("[case] wave vs lane repair, 3 uneven problems, mixed", flush=True)
    probs = [W.plain_fp64(seed=s, n_side=n, res=128) for s, n in ((11, 20), (12, 33), (13, 47))]
    rng = np.random.default_rng(9)
    w = np.concatenate([rng.normal(0.0, 3e-4, p.n) for p in probs])
    _agree(_eval(probs, w, "wave", precision="mixed"), _eval(probs, w, "lane", precision="mixed"),
           "batch")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wave_repair_device_solve(gpu_available, c1, precision):
    """The device-driven L-BFGS (repairs inside its graph rounds, trial steps) converges the same
    way with either construction."""
    from caustic_amd.ot import OTBatch

    print(f"[case] wave vs lane repair, device solve c1 {precision}", flush=True)
    out = {}
    for mode in ("wave", "lane"):
        with OTBatch([c1], precision=precision, repair_mode=mode) as b:
            w, rep = b.solve(history=10, tol=1e-4, stop="maxrel", h0_threshold=0.0,
                             max_iter=5000, device_loop=True)
            ev = b.evaluate(w)
        assert rep[0]["converged"] == 1, (mode, rep)
        out[mode] = (w, rep[0], ev)
    (ww, rw, ew), (wl, rl, el) = out["wave"], out["lane"]
    for ev in (ew, el):
        assert (np.abs(ev["mass"] - c1.nu) / c1.nu).max() < 1e-4
    print(f"iterations {rw['iterations']} (wave) {rl['iterations']} (lane), "
          f"max |dw| {np.abs(ww - wl).max():.2e}", flush=True)
    assert abs(rw["iterations"] - rl["iterations"]) <= max(3, rl["iterations"] // 10)


# ---- the full construction pass with one wave per cell (pass_mode 1) against one lane per cell


def _eval_pass(problems, w, pass_mode, evals=1, **kw):
    from caustic_amd.ot import OTBatch

    with OTBatch(problems, pass_mode=pass_mode, **kw) as b:
        for _ in range(evals):  # later evaluations take the previous cells' adjacency
            ev = b.evaluate(w)
        st = b.stats()
        counts, verts = b.cell_polygons(w)
    return ev, st, counts, verts


def _agree_pass(a, b, tag):
    ev_w, st_w, n_w, v_w = a
    ev_l, st_l, n_l, v_l = b
    assert st_w["cell_launches"] > 0 and st_l["cell_launches"] > 0, tag
    assert np.array_equal(ev_w["exist"], ev_l["exist"]), tag
    assert np.array_equal(ev_w["mass"], ev_l["mass"]), \
        f"{tag}: masses differ ({np.mean(ev_w['mass'] != ev_l['mass']):.4f} of cells)"
    assert np.array_equal(ev_w["f"], ev_l["f"]), tag
    assert np.array_equal(ev_w["centroid"], ev_l["centroid"]), tag
    assert np.array_equal(n_w, n_l), f"{tag}: vertex counts differ"
    k = np.arange(v_w.shape[1])[None, :] < n_w[:, None]
    assert np.array_equal(v_w[k], v_l[k]), f"{tag}: vertices differ"
    print(f"{tag}: identical; repair launches {st_w['repair_launches']} (wave) "
          f"{st_l['repair_launches']} (lane)", flush=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("evals", [1, 2])
def test_wave_pass_hidden_cells(gpu_available, c1, precision, evals):
    """Cold start (bucket rings) and, with two evaluations, the adjacency full pass; the mixed
    case covers FP32 clipping, the FP64 vertex refinement and the hand-off of FP32 cells."""
    print(f"[case] wave vs lane full pass, c1 random w, {precision}, {evals} evals", flush=True)
    w = np.random.default_rng(7).normal(0.0, 2e-4, c1.n)
    _agree_pass(_eval_pass([c1], w, 1, evals, precision=precision),
                _eval_pass([c1], w, 0, evals, precision=precision), f"pass {precision} {evals}")


def test_wave_pass_bucket_rounds(gpu_available, c1):
    print("[case] wave vs lane full pass, ring_init 1, w ~ N(0, 1.6e-3), mixed", flush=True)
    w = np.random.default_rng(100).normal(0.0, 1.6e-3, c1.n)
    _agree_pass(_eval_pass([c1], w, 1, 2, ring_init=1, precision="mixed"),
                _eval_pass([c1], w, 0, 2, ring_init=1, precision="mixed"), "pass rings")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wave_pass_device_solve(gpu_available, c1, precision):
    """The device-driven L-BFGS takes the same path with either full-pass construction."""
    from caustic_amd.ot import OTBatch

    print(f"[case] wave vs lane full pass, device solve c1 {precision}", flush=True)
    out = {}
    for pm in (1, 0):
        with OTBatch([c1], precision=precision, pass_mode=pm) as b:
            w, rep = b.solve(history=10, tol=1e-4, stop="maxrel", h0_threshold=0.0,
                             max_iter=5000, device_loop=True)
        assert rep[0]["converged"] == 1, (pm, rep)
        out[pm] = (w, rep[0])
    print(f"iterations {out[1][1]['iterations']} (wave) {out[0][1]['iterations']} (lane)",
          flush=True)
    assert out[1][1]["iterations"] == out[0][1]["iterations"]
    assert np.array_equal(out[1][0], out[0][0])
