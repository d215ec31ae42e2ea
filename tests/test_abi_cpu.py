"""C ABI and CLI checks that need no GPU: exported symbols, struct layouts, argument errors."""
import ctypes
import json
import os
import re
import subprocess

import pytest

from caustic_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "caustic_ot.h")
CLI = os.path.join(ROOT, "caustic_amd", "bin", "caustic_gpu")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cot_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = _declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), f"{n} missing from libcaustic_ot.so"
        assert n in L.SIGNATURES, f"{n} has no ctypes signature"


def test_struct_layouts_match_the_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "caustic_ot.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\", "
        "sizeof(cot_problem), "
        "sizeof(cot_options), sizeof(cot_solve_opts), sizeof(cot_report), sizeof(cot_stats), "
        "offsetof(cot_solve_opts, newton_sweeps), sizeof(cot_reflection), "
        "offsetof(cot_reflection, if_aa_op_initial), offsetof(cot_options, full_pass_ring), "
        "offsetof(cot_solve_opts, newton_coarse), sizeof(cot_ladder_opts), "
        "offsetof(cot_ladder_opts, seed));"
        "return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [ctypes.sizeof(L.cot_problem), ctypes.sizeof(L.cot_options),
            ctypes.sizeof(L.cot_solve_opts), ctypes.sizeof(L.cot_report),
            ctypes.sizeof(L.cot_stats), L.cot_solve_opts.newton_sweeps.offset,
            ctypes.sizeof(L.cot_reflection), L.cot_reflection.if_aa_op_initial.offset,
            L.cot_options.full_pass_ring.offset, L.cot_solve_opts.newton_coarse.offset,
            ctypes.sizeof(L.cot_ladder_opts), L.cot_ladder_opts.seed.offset]
    assert got == want


def test_vertex_slots_match_the_header():
    m = re.search(r"#define COT_MAX_VERTS (\d+)", open(HEADER).read())
    assert m and int(m.group(1)) == L.COT_MAX_VERTS


def test_create_rejects_bad_arguments_without_touching_the_gpu():
    lib = L.load()
    ctx = ctypes.c_void_p()
    rc = lib.cot_create(None, 0, None, ctypes.byref(ctx))
    assert rc == L.COT_EINVAL
    assert b"no problems" in lib.cot_last_error()
    arr = (L.cot_problem * 1)()
    arr[0].n_sites = 0
    rc = lib.cot_create(arr, 1, None, ctypes.byref(ctx))
    assert rc == L.COT_EINVAL
    assert lib.cot_solve(None, None, None, None) == L.COT_EINVAL


def test_default_options():
    lib = L.load()
    o = L.cot_options()
    lib.cot_default_options(ctypes.byref(o))
    assert o.precision == L.COT_PREC_FP64 and o.ring_init == 2 and o.verify_tol == 1e-4
    s = L.cot_solve_opts()
    lib.cot_default_solve_opts(ctypes.byref(s))
    assert s.history == 10 and s.h0_threshold == 1e-5 and s.device_loop == 1


def _layout(tmp_path, settings=None, drop=()):
    """The reference's directory layout around Optimal_Transport_rl/build."""
    root = tmp_path / "caustic"
    build = root / "Optimal_Transport_rl" / "build"
    build.mkdir(parents=True)
    xi = "demo"
    sj = root / "Optimal_Transport_rl" / "set_json" / xi
    sj.mkdir(parents=True)
    cfg = settings or {
        "test_num": "t", "picture_para": {"name": "pic"},
        "sample_point_para": {"data_name": "d", "total_iter": 10, "convergence_tolerance": 0.01,
                              "history_size": 7, "aa_history_size": 7, "aa_fixed_para": -2,
                              "if_aa_op_initial": True, "initial_step": 1,
                              "sample_domain": [0, 4, 0, 4]},
        "len_para": {"z_of_len": 0, "z_of_receiver": 10, "proj_domain": [0, 4, 0, 4],
                     "refraction_index": 1.5},
        "light_para": {"light_type": "parallel_light", "point_light_coordinate": [0, 0, -1]},
    }
    (sj / "_py_sp_settings.json").write_text(json.dumps(cfg))
    pic = root / "data" / "pic_data"
    pic.mkdir(parents=True)
    if "pic" not in drop:
        (pic / "pic.txt").write_text("1 2\n3 4\n")
    seed = root / "Optimal_Transport_rl" / "seed_point" / xi
    (seed / "0").mkdir(parents=True)
    pts = "\n".join(f"{x} {y}" for x in (1, 3) for y in (1, 3))
    for f in ("points.txt", "proj_xy.txt"):
        if f not in drop:
            (seed / f).write_text(pts + "\n")
    for f in ("left_idx.txt", "right_idx.txt", "bottom_idx.txt", "top_idx.txt"):
        (seed / f).write_text("0\n")
    if "inside" not in drop:
        (seed / "inside_light_direct.txt").write_text("0 0 1\n" * 4)
    if "height" not in drop:
        (seed / "0" / "height.txt").write_text("0\n" * 4)
    return build, xi


needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")


@needs_cli
def test_cli_usage_and_missing_settings(tmp_path):
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([CLI, "0", "nothere"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 255 and "cannot read" in r.stderr


@needs_cli
def test_cli_settings_errors(tmp_path):
    build, xi = _layout(tmp_path, settings={"test_num": "x"})
    r = subprocess.run([CLI, "0", xi], cwd=build, capture_output=True, text=True)
    assert r.returncode == 2 and "missing key 'sample_point_para'" in r.stderr


@needs_cli
@pytest.mark.parametrize("drop,msg", [("pic", "picture"), ("points.txt", "points.txt"),
                                      ("inside", "inside_light_direct"), ("height", "height")])
def test_cli_missing_inputs(tmp_path, drop, msg):
    build, xi = _layout(tmp_path, drop=(drop,))
    r = subprocess.run([CLI, "0", xi], cwd=build, capture_output=True, text=True)
    assert r.returncode == 1 and msg in r.stderr


@needs_cli
def test_cli_rejects_unknown_light(tmp_path):
    """An unknown light_type is an error (the reference prints it and goes on with
    uninitialised source masses, main.cu:232-234); a point light on the lens plane is a
    settings error.  Both are caught before any GPU work."""
    build, xi = _layout(tmp_path)
    cfgp = build.parent / "set_json" / xi / "_py_sp_settings.json"
    cfg = json.loads(cfgp.read_text())
    cfg["light_para"]["light_type"] = "area_light"
    cfgp.write_text(json.dumps(cfg))
    r = subprocess.run([CLI, "0", xi], cwd=build, capture_output=True, text=True)
    assert r.returncode == 3 and "unknown light_type" in r.stderr
    cfg["light_para"]["light_type"] = "point_light"
    cfg["light_para"]["point_light_coordinate"] = [0.5, 0.5, cfg["len_para"]["z_of_len"]]
    cfgp.write_text(json.dumps(cfg))
    r = subprocess.run([CLI, "0", xi], cwd=build, capture_output=True, text=True)
    assert r.returncode == 2 and "point_light_coordinate" in r.stderr


def test_irra_render_rejects_bad_arguments_without_touching_the_gpu():
    lib = L.load()
    dom = (ctypes.c_double * 4)(0.0, 1.0, 0.0, 1.0)
    img = ctypes.c_void_p(0)
    assert lib.cot_irra_render(None, None, -1, None, 0, 1.0, dom, 4, 4, img, None) == L.COT_EINVAL
    assert lib.cot_irra_render(None, None, 0, None, 0, 1.0, dom, 0, 4, img, None) == L.COT_EINVAL
    bad = (ctypes.c_double * 4)(1.0, 1.0, 0.0, 1.0)
    assert lib.cot_irra_render(None, None, 0, None, 0, 1.0, bad, 4, 4, ctypes.c_void_p(8),
                               None) == L.COT_EINVAL
    assert b"proj_domain" in lib.cot_last_error()
