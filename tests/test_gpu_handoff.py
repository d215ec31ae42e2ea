"""The in-memory hand-off (caustic_amd.handoff) against the executable's files on the same tree,
the torch-tensor entry points (cot_*_dev) against the host ones, and ot.direct_ot against the
FP64 oracle."""
import os

import numpy as np
import pytest

from caustic_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_handoff_reproduces_the_executable_t0_then_t1(gpu_available, tmp_path):
    """caustic_gpu 0 then 1 on a synthetic tree in the reference's layout, and OTHandoff.run(0),
    run(1) on the same inputs in memory: the tensors equal what load_ot_data would read from
    in.txt and <t>/out.txt, and the weights carried between the runs equal the 6-digit weights
    file (utils/basic_function.py:14-34 and utils/data_utils.py:19-24 replaced)."""
    import json

    import torch

    import test_gpu_cli as C
    from caustic_amd.handoff import OTHandoff

    root = str(tmp_path)
    T = C.make_tree(root)
    xi = C.XIUSHI
    with open(os.path.join(root, "Optimal_Transport_rl", "set_json", xi,
                           "_py_sp_settings.json")) as f:
        settings = json.load(f)
    picture = np.loadtxt(os.path.join(root, "data", "pic_data", C.PIC + ".txt"))
    pdir = os.path.join(root, "data", "point_data", xi)
    wdir = os.path.join(root, "data", "weights", xi)
    ot = OTHandoff(settings, picture, T["points"], T["points"], T["ins"])
    for t in (0, 1):
        print(f"[case] caustic_gpu {t} vs OTHandoff.run({t})", flush=True)
        r = C.run_cli(T["build"], t)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        origin, aim = ot.run(t, T["heights"][t])
        assert origin.dtype == aim.dtype == torch.float32 and origin.is_cuda
        assert origin.requires_grad and aim.requires_grad
        in_file = np.loadtxt(os.path.join(pdir, "in.txt"), dtype=np.float32)
        out_file = np.loadtxt(os.path.join(pdir, str(t), "out.txt"), dtype=np.float32)
        np.testing.assert_array_equal(origin.detach().cpu().numpy(), in_file)
        np.testing.assert_array_equal(aim.detach().cpu().numpy(), out_file)
        w_file = np.loadtxt(os.path.join(wdir, str(t), "final_proj_point_weights.txt"))
        np.testing.assert_array_equal(ot.weights, w_file)
    assert ot.reports[0]["converged"] and ot.reports[1]["converged"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tensor_entry_points_match_host_ones(gpu_available, precision):
    """evaluate / centroids / solve on CUDA tensors (device arrays, the current stream) give
    exactly the host entry points' results: the same sequence of calls in two fresh contexts
    (each evaluation starts from the previous one's cells, so the sequences must match)."""
    import torch

    from caustic_amd.ot import OTBatch

    print(f"[case] tensor entry points {precision}", flush=True)
    p = W.plain_fp64(seed=41, n_side=32, res=128)
    rng = np.random.default_rng(3)
    w = rng.normal(0.0, 2e-4, p.n)
    kw = dict(history=10, tol=1e-5, stop="maxrel", h0_threshold=0.0)
    with OTBatch([p], precision=precision) as b:
        host = b.evaluate(w)
        cen_h, _ = b.centroids(w, weighted=False)
        ws_h, rep_h = b.solve(w, **kw)
    wt = torch.tensor(w, dtype=torch.float64, device="cuda:0")
    with OTBatch([p], precision=precision) as b:
        dev = b.evaluate_tensor(wt)
        cen_d, _ = b.centroids_tensor(wt, weighted=False)
        ws_d, rep_d = b.solve_tensor(wt, **kw)
    np.testing.assert_array_equal(dev["mass"].cpu().numpy(), host["mass"])
    np.testing.assert_array_equal(dev["centroid"].cpu().numpy(), host["centroid"])
    np.testing.assert_array_equal(dev["exist"].cpu().numpy().astype(bool), host["exist"])
    np.testing.assert_array_equal(dev["f"].cpu().numpy(), host["f"])
    np.testing.assert_array_equal(cen_d.cpu().numpy(), cen_h)
    assert rep_d[0]["converged"] and rep_d[0]["iterations"] == rep_h[0]["iterations"]
    np.testing.assert_array_equal(ws_d.cpu().numpy(), ws_h)
    assert wt.cpu().numpy().tolist() == w.tolist()  # the input tensor is left alone


def test_direct_ot_matches_oracle(gpu_available):
    """ot.direct_ot (the in-process mirror of direct_ot, ot_op.cuh:66-70): the reference's
    stopping rule met at the weights it writes back in place, confirmed by the oracle, and the
    returned points are the oracle's rho-weighted centroids there."""
    from caustic_amd.ot import direct_ot

    print("[case] direct_ot", flush=True)
    p = W.plain_fp64(seed=43, n_side=24, res=96)
    weights = np.zeros(p.n)
    m, n = p.pic_size
    out, ex = direct_ot(p.sites.ravel(), weights, p.rho.ravel(), (m, n), 20000, 1e-4, 1.0, 10,
                        p.proj_domain, p.nu)
    ev = O.evaluate(p.sites, weights, p.nu, p.rho, p.proj_domain, centroid_min_piece=0.0)
    assert np.linalg.norm(ev.mass - p.nu) < 1e-4 and ex.all()
    np.testing.assert_allclose(out.reshape(-1, 2), ev.centroid, rtol=0, atol=1e-10)
