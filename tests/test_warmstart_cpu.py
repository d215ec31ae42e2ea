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


def test_weight_normalisation_round_trip_removes_the_gauge():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 1e-4, 64 * 64) + 3.0
    t = WS.weights_to_target(w, 64)
    assert abs(float(t.mean())) < 1e-5
    np.testing.assert_allclose(WS.target_to_weights(t, 64).ravel(), w - w.mean(), atol=1e-10)


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
