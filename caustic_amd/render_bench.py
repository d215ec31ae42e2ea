"""Timing of the stage-2/3 renderers at the reference's own sizes.

The lens mesh of the reference's configuration: the triangular lattice of 160,978 points
(initial_xy_point for point_num 160,000 on the 40 x 40 sample domain, utils/data_utils.py:374-410),
its Delaunay triangles (~321k faces), mapped onto the 160-unit receiver by the similarity map
(utils/data_utils.py:413-435) plus a smooth distortion, a 640 x 640 picture.  Times, with HIP
events on the current stream over `--iters` calls after one untimed call:

  render        irra_render.render (render.cu:553-591)
  render_grad   irra_render.render_grad without and with the Sobel branch (render.cu:593-728)
  splat         splat.render_gaussian_splat_two_pass_backward (utils/basic_function.py:674-772)
                for one ray hit per lattice point, 5 x 5 windows

Run ``python -m caustic_amd.render_bench`` on a GPU; prints one JSON line.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from . import workloads as W

SAMPLE_DOMAIN = (0.0, 40.0, 0.0, 40.0)
PROJ_DOMAIN = (0.0, 160.0, -60.0, 100.0)


def lens_mesh(point_num: int = 160000, seed: int = 0):
    """Lattice points, their Delaunay faces and receiver positions (distorted similarity map)."""
    from scipy.spatial import Delaunay

    pts = W.triangular_lattice(point_num, SAMPLE_DOMAIN)
    faces = Delaunay(pts).simplices.astype(np.int32)
    xy = W.similarity_map(pts, SAMPLE_DOMAIN, PROJ_DOMAIN)
    rng = np.random.default_rng(seed)
    u = (pts - 20.0) / 20.0
    a = rng.uniform(1.0, 3.0, 4)
    xy = xy + 6.0 * np.stack([np.sin(a[0] * u[:, 1]) * np.cos(a[1] * u[:, 0]),
                              np.sin(a[2] * u[:, 0]) * np.cos(a[3] * u[:, 1])], axis=1)
    return pts, faces, xy


def main(argv=None) -> int:
    import torch

    from . import irra_render, splat

    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--res", type=int, default=640)
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("no GPU visible")
        return 1
    dev = torch.device("cuda:0")
    pts, faces, xy = lens_mesh()
    P, F = xy.shape[0], faces.shape[0]
    a = xy[faces]
    area = 0.5 * np.abs((a[:, 1, 0] - a[:, 0, 0]) * (a[:, 2, 1] - a[:, 0, 1])
                        - (a[:, 2, 0] - a[:, 0, 0]) * (a[:, 1, 1] - a[:, 0, 1]))
    tri = torch.tensor(area / area.sum(), dtype=torch.float32, device=dev)
    fc = torch.tensor(faces, dtype=torch.int32, device=dev)
    pxy = torch.tensor(xy, dtype=torch.float32, device=dev)
    res = args.res
    gray = float(res * res * 100)
    pic = torch.rand((res, res), device=dev) * 200.0
    dom = torch.tensor(PROJ_DOMAIN, dtype=torch.float64)
    size = torch.tensor([res, res])

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(args.iters):
            fn()
        e.record()
        torch.cuda.synchronize()
        return s.elapsed_time(e) / args.iters

    out = {"points": int(P), "faces": int(F), "picture": [res, res], "iters": args.iters}
    out["render_ms"] = timed(lambda: irra_render.render(tri, fc, pxy, gray, dom, size))
    for sob in (False, True):
        out[f"render_grad{'_sobel' if sob else ''}_ms"] = timed(
            lambda: irra_render.render_grad(tri, fc, pxy, gray, pic, sob, dom, size))
    vp = pxy.to(torch.float64).requires_grad_(True)
    energy = torch.rand(P, dtype=torch.float64, device=dev) + 0.5
    target = torch.rand((res, res), dtype=torch.float64, device=dev)

    def do_splat():
        img, loss = splat.render_gaussian_splat_two_pass_backward(
            None, gray, vp, energy, (res, res), PROJ_DOMAIN, target, 1.0)
        vp.grad = None
        return loss

    out["splat_two_pass_ms"] = timed(do_splat)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
