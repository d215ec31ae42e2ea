"""The irra_render extension's forward pass on the MI355X (render_exe_program/render.cu).

``render(compute_tri_area, faces, proj_xy, total_gray_value, proj_domain, pic_size)`` mirrors
the reference's ``irra_render.render`` (render.cu:553-591; caller utils/optimization.py:729):
the same tensors in, ``[render_result]`` out, an H x W float32 image (H = pic_size[0] pixels
along x, W = pic_size[1] along y) in the reference's output layout (row H - 1 - ix, column iy).
Face f spreads compute_tri_area[f] * total_gray_value uniformly over its triangle; the pixel
pieces are exact polygon clips (cot_irra_render in libcaustic_ot.so).  The gradient pass
(``render_grad``) is not built.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


def render(compute_tri_area, faces, proj_xy, total_gray_value, proj_domain, pic_size):
    import torch

    dev = compute_tri_area.device
    if dev.type != "cuda":
        raise ValueError("irra_render.render takes CUDA tensors")
    tri = compute_tri_area.detach().reshape(-1).to(torch.float32).contiguous()
    fc = faces.detach().reshape(-1, 3).to(device=dev, dtype=torch.int32).contiguous()
    xy = proj_xy.detach().reshape(-1, 2).to(device=dev, dtype=torch.float32).contiguous()
    if fc.shape[0] != tri.shape[0]:
        raise ValueError(f"{tri.shape[0]} face energies for {fc.shape[0]} faces")
    gray = float(torch.as_tensor(total_gray_value).reshape(-1)[0].item())
    dom = np.ascontiguousarray(torch.as_tensor(proj_domain).detach().cpu().reshape(4)
                               .to(torch.float64).numpy())
    ps = torch.as_tensor(pic_size).detach().cpu().reshape(2)
    H, W = int(ps[0].item()), int(ps[1].item())
    out = torch.empty((H, W), dtype=torch.float32, device=dev)
    lib = L.load()
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        rc = lib.cot_irra_render(ctypes.c_void_p(tri.data_ptr()), ctypes.c_void_p(fc.data_ptr()),
                                 int(fc.shape[0]), ctypes.c_void_p(xy.data_ptr()),
                                 int(xy.shape[0]), gray,
                                 dom.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), H, W,
                                 ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
    L.check(rc)
    return [out]
