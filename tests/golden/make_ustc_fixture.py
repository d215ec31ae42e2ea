"""Writes tests/golden/ustc_centerL_t0.npz from the reference's shipped run (data files only).

The reference ships one converged run of caustic_GPU at its own configuration
(configs/_py_sp_settings.json: 160,000 requested "triangle" seeds on a 40 x 40 sample domain,
proj_domain [0, 160, -60, 100], target data/pic_data/ustc_centerL.txt, history_size 7,
convergence_tolerance 0.01).  Its outputs at ot_time 0 are
  data/point_data/ustc_centerL_tri_standard_sp_parallel_light_rl/in.txt       (source points)
  data/point_data/ustc_centerL_tri_standard_sp_parallel_light_rl/0/out.txt    (centroids)
  data/weights/ustc_centerL_tri_standard_sp_parallel_light_rl/0/final_proj_point_weights.txt
all printed by C++ ostream at 6 significant digits (main.cu:291-334).

Stored: the picture text matrix as uint8 (its values are the integers 0..200), the weights and
centroids as float32 (every 6-significant-digit decimal is the shortest repr of its float32
neighbour, so "%g" of the stored value gives the text back; checked here), the SHA-256 of
in.txt (the lattice test reproduces it byte for byte) and the configuration keys the solve
reads.  No reference code is read or run.

    python tests/golden/make_ustc_fixture.py [/root/reference]
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RUN = "ustc_centerL_tri_standard_sp_parallel_light_rl"


def main(ref="/root/reference"):
    pic = np.loadtxt(os.path.join(ref, "data", "pic_data", "ustc_centerL.txt"))
    assert np.all(pic == np.round(pic)) and pic.min() >= 0 and pic.max() <= 255
    w = np.loadtxt(os.path.join(ref, "data", "weights", RUN, "0", "final_proj_point_weights.txt"))
    out = np.loadtxt(os.path.join(ref, "data", "point_data", RUN, "0", "out.txt"))
    with open(os.path.join(ref, "data", "point_data", RUN, "in.txt"), "rb") as f:
        in_bytes = f.read()
    for a in (w, out):
        back = a.astype(np.float32).astype(np.float64).ravel()
        assert all("%g" % x == "%g" % y for x, y in zip(a.ravel(), back))
    with open(os.path.join(ref, "configs", "_py_sp_settings.json")) as f:
        cfg = json.load(f)
    sp, lp = cfg["sample_point_para"], cfg["len_para"]
    settings = {
        "point_num": sp["point_num"], "type": sp["type"], "sample_domain": sp["sample_domain"],
        "total_iter": sp["total_iter"], "convergence_tolerance": sp["convergence_tolerance"],
        "history_size": sp["history_size"], "initial_step": sp["initial_step"],
        "if_aa_op_initial": sp["if_aa_op_initial"], "aa_history_size": sp["aa_history_size"],
        "aa_fixed_para": sp["aa_fixed_para"], "proj_domain": lp["proj_domain"],
        "z_of_len": lp["z_of_len"], "z_of_receiver": lp["z_of_receiver"],
        "refraction_index": lp["refraction_index"], "light_type": cfg["light_para"]["light_type"],
        "point_light_coordinate": cfg["light_para"]["point_light_coordinate"],
        "picture": cfg["picture_para"]["name"], "test_num": cfg["test_num"],
        "data_name": sp["data_name"],
    }
    np.savez_compressed(
        os.path.join(HERE, "ustc_centerL_t0.npz"),
        picture=pic.astype(np.uint8), weights=w.astype(np.float32),
        centroids=out.astype(np.float32), in_sha256=hashlib.sha256(in_bytes).hexdigest(),
        in_lines=in_bytes.count(b"\n"), settings=json.dumps(settings))


if __name__ == "__main__":
    main(*sys.argv[1:])
