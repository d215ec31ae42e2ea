"""Warm-start host logic without a GPU: the network library's exported symbols, the parameter
layout against the CPU oracle's, features, weight normalisation and field resampling."""
import ctypes
import os
import re

import numpy as np

from caustic_amd import warmstart as WS
from caustic_amd import workloads as W
from oracle import unet_oracle as U

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_network_library_exports_every_declared_symbol():
    text = open(os.path.join(ROOT, "include", "caustic_wsn.h")).read()
    names = sorted(set(re.findall(r"\b(wsn_[a-z_]+)\s*\(", text)))
    assert len(names) >= 10
    lib = ctypes.CDLL(WS.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    WS._load()  # every declared function has a ctypes signature


def test_layer_table_and_flat_layout_match_the_oracle():
    assert [L[:4] for L in WS.LAYERS] == [L[:4] for L in U.LAYERS]
    P = U.init_params(4)
    flat = WS.dict_to_flat(P)
    assert flat.size == U.n_params()
    back = WS.flat_to_dict(flat)
    for k in P:
        assert np.array_equal(back[k], P[k])
    assert np.array_equal(WS.init_params(4), flat)


def test_features_are_log_of_mean_normalised_box_average():
    rho = np.ones((256, 256))
    rho[:128] = 3.0
    f = WS.features(rho, 64)
    assert f.shape == (64, 64) and f.dtype == np.float32
    np.testing.assert_allclose(np.exp(f[:32]), 1.5, rtol=1e-6)
    np.testing.assert_allclose(np.exp(f[32:]), 0.5, rtol=1e-6)


def test_field_round_trip_on_a_regular_grid_removes_the_gauge():
    """Sites exactly at the nodes: U = -w and w = -U(s), up to the mean."""
    rng = np.random.default_rng(0)
    S = 64
    sites = WS.grid_nodes(S)
    w = rng.normal(0, 1e-4, S * S) + 3.0
    U = WS.field_from_solution(sites, w, rng.uniform(0, 1, (S * S, 2)), S)
    assert abs(float(U.mean())) < 1e-12
    np.testing.assert_allclose(WS.weights_from_field(U, sites), w - w.mean(), atol=1e-12)
    np.testing.assert_allclose(WS.target_to_field(WS.field_to_target(U)), U, rtol=1e-6,
                               atol=1e-6 * np.abs(U).max())


def test_field_correction_recovers_a_smooth_field_from_jittered_sites():
    """For w_i = -U(s_i) with a smooth U and centroids from its gradient (c = s + grad U / 2),
    the first-order correction gives U at the nodes to second order in the jitter."""
    S = 64
    rng = np.random.default_rng(1)
    sites = WS.grid_nodes(S) + rng.uniform(-0.25, 0.25, (S * S, 2)) / S
    Uf = lambda x, y: 0.01 * np.sin(3 * x) * np.cos(2 * y)
    gx = lambda x, y: 0.03 * np.cos(3 * x) * np.cos(2 * y)
    gy = lambda x, y: -0.02 * np.sin(3 * x) * np.sin(2 * y)
    w = -Uf(sites[:, 0], sites[:, 1])
    cen = sites + 0.5 * np.stack([gx(sites[:, 0], sites[:, 1]), gy(sites[:, 0], sites[:, 1])], 1)
    U = WS.field_from_solution(sites, w, cen, S)
    nodes = WS.grid_nodes(S)
    ref = Uf(nodes[:, 0], nodes[:, 1]).reshape(S, S)
    naive = (-w - (-w).mean()).reshape(S, S)
    err = np.abs(U - (ref - ref.mean())).max()
    # first order in the jitter uncorrected (|grad U| |ds| ~ 1e-4), second order corrected:
    # |U''| |ds|^2 / 2 with |U''| <= 0.13 and |ds|^2 <= 2 (0.25 / 64)^2
    assert np.abs(naive - (ref - ref.mean())).max() > 5e-5
    assert err < 0.13 * 2 * (0.25 / S) ** 2
    assert err < 2e-2 * np.abs(naive - (ref - ref.mean())).max()


def test_resample_field_reproduces_linear_fields_and_nodes():
    S = 32
    c = (np.arange(S) + 0.5) / S
    X, Y = np.meshgrid(c, c, indexing="xy")
    f = 2.0 * X - 3.0 * Y + 1.0
    rng = np.random.default_rng(1)
    pts = rng.uniform(0.5 / S, 1 - 0.5 / S, size=(500, 2))
    np.testing.assert_allclose(WS.resample_field(f, pts), 2 * pts[:, 0] - 3 * pts[:, 1] + 1,
                               atol=1e-12)
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    np.testing.assert_allclose(WS.resample_field(f, nodes), f.ravel(), atol=1e-12)


def test_coarse_problem_keeps_pixels_per_site_and_mass():
    p = W.warm_start_sample(5, n_side=64, res=256)
    cp = WS.coarse_problem(p, 32, seed=9)
    assert cp.n == 1024 and cp.rho.shape == (128, 128)
    assert abs(cp.rho.sum() - 1.0) < 1e-12 and abs(cp.nu.sum() - 1.0) < 1e-12


def test_field_round_trip_in_a_frame():
    """Sites sampling a sub-rectangle (a tilted source plane): the field lives on the frame's
    grid nodes, and exact node weights come back from it."""
    frame = (0.1, 0.8, 0.2, 1.0)
    S = 8
    nodes = WS.grid_nodes(S, frame)
    assert nodes[:, 0].min() > 0.1 and nodes[:, 0].max() < 0.8
    rng = np.random.default_rng(4)
    w = rng.normal(size=S * S)
    cen = rng.uniform(0, 1, (S * S, 2))
    U = WS.field_from_solution(nodes, w, cen, S, frame)
    np.testing.assert_allclose(WS.weights_from_field(U, nodes, frame), w - w.mean(), atol=1e-12)


def test_learned_start_low_pass_is_in_grid_nodes():
    """smooth_field's width is in grid nodes: a cosine of k cycles over 128 nodes keeps
    exp(-(2 pi k sigma / 128)^2 / 2) of its amplitude away from the edges; the mean goes."""
    from caustic_amd import warmstart as WS

    S, k, sigma = 128, 4, 4.0
    x = np.arange(S)
    U = np.cos(2 * np.pi * k * x / S)[None, :] * np.ones((S, 1)) + 3.0
    f = WS.smooth_field(U, sigma)
    want = np.exp(-0.5 * (2 * np.pi * k * sigma / S) ** 2)
    # crest minus trough away from the edges: 2 x the kept amplitude (removing the mean shifts
    # both by the same constant)
    amp = 0.5 * (f[S // 2, [32, 64]] - f[S // 2, [48, 80]])
    np.testing.assert_allclose(amp, want, rtol=1e-4)
    assert abs(f.mean()) < 1e-12
