"""End to end: the drop-in executable `caustic_gpu <ot_time> <xiushi>` on a synthetic input tree
in the reference's layout (main.cu:51-342), t = 0 then t = 1, outputs checked against the FP64
CPU oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from caustic_amd import workloads as W
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "caustic_amd", "bin", "caustic_gpu")
XIUSHI = "synthetic_parallel_light"
PIC = "synthetic_target"
SAMPLE = (0.0, 2.0, 0.0, 2.0)
PROJ = (-0.5, 2.5, -0.25, 2.25)
TOL = 1e-5


def write_picture(path, A):
    """Inverse of the reference's reader: element (ix, iy) at text row m-1-ix, column iy."""
    np.savetxt(path, A.T[::-1], fmt="%.17g")


def make_tree(root, xiushi=XIUSHI, light=None):
    """Sites on a 16x16 jittered grid of the sample domain, projected straight down (proj_xy =
    points); a 64x48 target of three Gaussians; lens data for t = 0 and t = 1.  light: None for
    a parallel light, else the point light's coordinates."""
    rng = W.rng_for(5000)
    pts = np.array(SAMPLE[0::2]) + W.jittered_grid(16, rng) * (SAMPLE[1] - SAMPLE[0])
    n = pts.shape[0]
    A = W.gaussian_mixture(64, rng, 3, (0.05, 0.15), (0.5, 1.0), (0.2, 0.8), 0.05)[:48, :]
    A = A / A.sum() * 1234.5  # arbitrary grey scale: the solver normalises
    ins = np.concatenate([rng.uniform(-0.1, 0.1, (n, 2)), np.ones((n, 1))], 1)
    rl = os.path.join(root, "Optimal_Transport_rl")
    build = os.path.join(rl, "build")
    seed = os.path.join(rl, "seed_point", xiushi)
    for d in (build, os.path.join(rl, "set_json", xiushi), os.path.join(seed, "0"),
              os.path.join(seed, "1"), os.path.join(root, "data", "pic_data")):
        os.makedirs(d, exist_ok=True)
    settings = {
        "test_num": "synthetic",
        "picture_para": {"name": PIC, "type": "png"},
        "sample_point_para": {"data_name": "synthetic", "sample_domain": list(SAMPLE),
                              "total_iter": 20000, "convergence_tolerance": TOL,
                              "history_size": 7, "initial_step": 1, "if_aa_op_initial": True,
                              "aa_history_size": 7, "aa_fixed_para": -0.5},
        "len_para": {"z_of_len": 0, "z_of_receiver": 4.0, "proj_domain": list(PROJ),
                     "refraction_index": 1.5},
        "light_para": {"light_type": "parallel_light" if light is None else "point_light",
                       "point_light_coordinate": [1, 1, -1] if light is None else list(light)},
        "mi355x": {"full_precision_weights": True},
    }
    with open(os.path.join(rl, "set_json", xiushi, "_py_sp_settings.json"), "w") as f:
        json.dump(settings, f, indent=2)
    write_picture(os.path.join(root, "data", "pic_data", PIC + ".txt"), A)
    for name in ("points.txt", "proj_xy.txt"):
        np.savetxt(os.path.join(seed, name), pts, fmt="%.17g")
    for k, name in enumerate(("left_idx.txt", "right_idx.txt", "bottom_idx.txt", "top_idx.txt")):
        np.savetxt(os.path.join(seed, name), np.arange(k, n, 16), fmt="%d")
    np.savetxt(os.path.join(seed, "inside_light_direct.txt"), ins, fmt="%.17g")
    heights = {}
    for t in (0, 1):
        heights[t] = 0.1 + rng.uniform(0.0, 0.02, n)
        np.savetxt(os.path.join(seed, str(t), "height.txt"), heights[t], fmt="%.17g")
    return {"build": build, "points": pts, "A": A, "ins": ins, "heights": heights,
            "settings": settings}


def run_cli(build, t, xiushi=XIUSHI):
    print(f"[case] caustic_gpu {t} {xiushi}", flush=True)
    return subprocess.run([CLI, str(t), xiushi], cwd=build, capture_output=True, text=True,
                          timeout=300)


def read_rows(path):
    with open(path) as f:
        return [line.split() for line in f if line.strip()]


@pytest.mark.gpu
def test_cli_t0_then_t1(gpu_available, tmp_path):
    assert os.path.exists(CLI), "caustic_amd/bin/caustic_gpu is not built"
    root = str(tmp_path)
    T = make_tree(root)
    rho = T["A"] / T["A"].sum() * ((SAMPLE[1] - SAMPLE[0]) * (SAMPLE[3] - SAMPLE[2]))
    pts = T["points"]
    nu = O.source_masses(pts, SAMPLE)
    wdir = os.path.join(root, "data", "weights", XIUSHI)
    pdir = os.path.join(root, "data", "point_data", XIUSHI)

    # ---- t = 0: plain solve, reference stopping rule
    r = run_cli(T["build"], 0)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    w_full = np.loadtxt(os.path.join(wdir, "0", "final_proj_point_weights_full.txt"))
    w6 = [row[0] for row in read_rows(os.path.join(wdir, "0", "final_proj_point_weights.txt"))]
    assert w6 == [O.format_weight(x) for x in w_full]
    ev = O.evaluate(pts, w_full, nu, rho, PROJ)
    assert np.linalg.norm(ev.mass - nu) < 2 * TOL
    ex = ev.exist.astype(bool)
    out = np.array(read_rows(os.path.join(pdir, "0", "out.txt")), dtype=np.float64)
    assert out.shape == (ex.sum(), 2)
    # the text files hold 6 significant digits: relative rounding error <= 5e-6
    np.testing.assert_allclose(out, ev.centroid[ex], rtol=6e-6, atol=1e-9)
    inn = np.array(read_rows(os.path.join(pdir, "in.txt")), dtype=np.float64)
    np.testing.assert_allclose(inn, pts[ex], rtol=6e-6, atol=1e-9)
    with open(os.path.join(pdir, "0", "_py_sp_settings.json")) as f:
        assert json.load(f) == T["settings"]
    # diagnostics: the final render in picture units (x from the bottom text row up) and mesh
    gray = T["A"].sum() / ((SAMPLE[1] - SAMPLE[0]) * (SAMPLE[3] - SAMPLE[2]))
    rdir = os.path.join(root, "Optimal_Transport_rl", "render_result", XIUSHI, "0")
    mdir = os.path.join(root, "Optimal_Transport_rl", "mesh_output", XIUSHI, "0")
    assert os.path.exists(os.path.join(rdir, "render_initial.txt"))
    assert os.path.exists(os.path.join(mdir, "mesh_initial.obj"))
    finals = [f for f in os.listdir(rdir) if f != "render_initial.txt"]
    assert len(finals) == 1 and finals[0].startswith("render_"), finals
    img = O.read_pic_txt(os.path.join(rdir, finals[0]))
    ref = O.render(ev.cells, nu, rho.shape, PROJ) * gray
    np.testing.assert_allclose(img, ref, rtol=0, atol=6e-7 + 1e-9 * ref.max())
    faces = [line for line in open(os.path.join(mdir, finals[0].replace("render_", "mesh_")
                                                .replace(".txt", ".obj"))) if line[0] == "f"]
    assert len(faces) == ex.sum()

    # ---- t = 1: starts from t = 0's 6-digit weights; Fresnel-weighted masses, Anderson
    r = run_cli(T["build"], 1)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    w1 = np.loadtxt(os.path.join(wdir, "1", "final_proj_point_weights_full.txt"))
    lens = {"points": pts, "ins": T["ins"] / np.linalg.norm(T["ins"], axis=1, keepdims=True),
            "height": T["heights"][1], "clip_area": nu, "z_of_len": 0.0,
            "z_of_receiver": 4.0, "refraction_index": 1.5,
            "area_scalar": (SAMPLE[1] - SAMPLE[0]) * (SAMPLE[3] - SAMPLE[2])}
    ev1 = O.evaluate(pts, w1, nu, rho, PROJ)
    nu1 = O.fresnel_masses(ev1, lens)
    assert np.linalg.norm(ev1.mass - nu1) < TOL * (1 + 1e-6)
    assert np.abs(nu1 - nu).max() > 1e-3 * nu.max()  # the Fresnel weighting moved the masses
    ex1 = ev1.exist.astype(bool)
    out1 = np.array(read_rows(os.path.join(pdir, "1", "out.txt")), dtype=np.float64)
    np.testing.assert_allclose(out1, ev1.centroid[ex1], rtol=6e-6, atol=1e-9)


@pytest.mark.gpu
def test_cli_point_light_t0_then_t1(gpu_available, tmp_path):
    """light_type point_light (ot_op.cuh:160-184, 687-709; main.cu:218-231): source masses from
    the solid angles under which the light sees the source cells, and at t = 1 the Fresnel
    weights with the light's own direction at each point; both checked by the oracle."""
    assert os.path.exists(CLI), "caustic_amd/bin/caustic_gpu is not built"
    xiushi, light = "synthetic_point_light", (0.8, 1.3, -0.9)
    root = str(tmp_path)
    T = make_tree(root, xiushi, light)
    area_scalar = (SAMPLE[1] - SAMPLE[0]) * (SAMPLE[3] - SAMPLE[2])
    rho = T["A"] / T["A"].sum() * area_scalar
    pts = T["points"]
    nu, clip, inc = O.point_light_source(pts, SAMPLE, light, 0.0)
    par = O.source_masses(pts, SAMPLE)
    assert np.abs(nu - par).max() > 0.1 * par.max()  # the light is near: masses differ
    wdir = os.path.join(root, "data", "weights", xiushi)
    r = run_cli(T["build"], 0, xiushi)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    w0 = np.loadtxt(os.path.join(wdir, "0", "final_proj_point_weights_full.txt"))
    ev = O.evaluate(pts, w0, nu, rho, PROJ)
    assert np.linalg.norm(ev.mass - nu) < 2 * TOL
    r = run_cli(T["build"], 1, xiushi)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    w1 = np.loadtxt(os.path.join(wdir, "1", "final_proj_point_weights_full.txt"))
    lens = {"points": pts, "ins": T["ins"] / np.linalg.norm(T["ins"], axis=1, keepdims=True),
            "height": T["heights"][1], "clip_area": clip, "incident": inc, "z_of_len": 0.0,
            "z_of_receiver": 4.0, "refraction_index": 1.5, "area_scalar": area_scalar}
    ev1 = O.evaluate(pts, w1, nu, rho, PROJ)
    nu1 = O.fresnel_masses(ev1, lens)
    assert np.linalg.norm(ev1.mass - nu1) < TOL * (1 + 1e-6)
    # the light's slant changes the front-surface transmittance: parallel-light masses differ
    nu1_par = O.fresnel_masses(ev1, dict(lens, incident=None))
    assert np.abs(nu1 - nu1_par).max() > 1e-4 * nu1.max()
