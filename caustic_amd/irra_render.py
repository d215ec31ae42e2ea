"""The irra_render extension on the MI355X (render_exe_program/render.cu).

``render(compute_tri_area, faces, proj_xy, total_gray_value, proj_domain, pic_size)`` mirrors
the reference's ``irra_render.render`` (render.cu:553-591; caller utils/optimization.py:729):
the same tensors in, ``[render_result]`` out, an H x W float32 image (H = pic_size[0] pixels
along x, W = pic_size[1] along y) in the reference's output layout (row H - 1 - ix, column iy).
Face f spreads compute_tri_area[f] * total_gray_value uniformly over its triangle; the pixel
pieces are exact polygon clips (cot_irra_render in libcaustic_ot.so).

``render_grad(compute_tri_area, faces, proj_xy, total_gray_value, real_picture, if_use_sobel,
proj_domain, pic_size)`` mirrors ``irra_render.render_grad`` (render.cu:593-728; caller
utils/optimization.py:281) with its nine outputs (cot_irra_render_grad); the two Sobel
outputs are zeros when ``if_use_sobel`` is false, as in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


def _mesh(compute_tri_area, faces, proj_xy, total_gray_value, proj_domain, pic_size):
    import torch

    dev = compute_tri_area.device
    if dev.type != "cuda":
        raise ValueError("irra_render takes CUDA tensors")
    tri = compute_tri_area.detach().reshape(-1).to(torch.float32).contiguous()
    fc = faces.detach().reshape(-1, 3).to(device=dev, dtype=torch.int32).contiguous()
    xy = proj_xy.detach().reshape(-1, 2).to(device=dev, dtype=torch.float32).contiguous()
    if fc.shape[0] != tri.shape[0]:
        raise ValueError(f"{tri.shape[0]} face energies for {fc.shape[0]} faces")
    gray = float(torch.as_tensor(total_gray_value).reshape(-1)[0].item())
    dom = np.ascontiguousarray(torch.as_tensor(proj_domain).detach().cpu().reshape(4)
                               .to(torch.float64).numpy())
    ps = torch.as_tensor(pic_size).detach().cpu().reshape(2)
    return dev, tri, fc, xy, gray, dom, int(ps[0].item()), int(ps[1].item())


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def render_grad(compute_tri_area, faces, proj_xy, total_gray_value, real_picture, if_use_sobel,
                proj_domain, pic_size, no_alloc=False):
    """[render_result, render_diff, render_proj_xy_grad, render_tri_area_grad, inte_diff,
    proj_xy_inte_grad, inte_result, sobel_proj_xy_grad, sobel_tri_area_grad], as the
    reference returns them (render.cu:716-727).  no_alloc: the library allocates no device
    memory for the call (cot_irra_render_grad_noalloc; Sobel images up to
    COT_NOALLOC_MAX_PIXELS pixels); the outputs are still allocated here."""
    import torch

    sobel = bool(torch.as_tensor(if_use_sobel).reshape(-1)[0].item())
    dev, tri, fc, xy, gray, dom, H, W = _mesh(compute_tri_area, faces, proj_xy,
                                              total_gray_value, proj_domain, pic_size)
    real = real_picture.detach().reshape(-1).to(device=dev, dtype=torch.float32).contiguous()
    if real.numel() != H * W:
        raise ValueError(f"real_picture holds {real.numel()} pixels, not {H} x {W}")
    F, P = fc.shape[0], xy.shape[0]
    f32 = dict(dtype=torch.float32, device=dev)
    out = [torch.empty((H, W), **f32), torch.empty((H, W), **f32), torch.empty((P, 2), **f32),
           torch.empty(F, **f32), torch.empty(F, **f32), torch.empty((P, 2), **f32),
           torch.empty(F, **f32), torch.zeros((P, 2), **f32), torch.zeros(F, **f32)]
    lib = L.load()
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        fn = lib.cot_irra_render_grad_noalloc if no_alloc else lib.cot_irra_render_grad
        rc = fn(
            _p(tri), _p(fc), int(F), _p(xy), int(P), gray, _p(real),
            dom.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), H, W, _p(out[0]), _p(out[1]),
            _p(out[2]), _p(out[3]), _p(out[4]), _p(out[5]), _p(out[6]), int(sobel), _p(out[7]),
            _p(out[8]), ctypes.c_void_p(stream))
    L.check(rc)
    return out


def render(compute_tri_area, faces, proj_xy, total_gray_value, proj_domain, pic_size):
    import torch

    dev, tri, fc, xy, gray, dom, H, W = _mesh(compute_tri_area, faces, proj_xy,
                                              total_gray_value, proj_domain, pic_size)
    out = torch.empty((H, W), dtype=torch.float32, device=dev)
    lib = L.load()
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        rc = lib.cot_irra_render(_p(tri), _p(fc), int(fc.shape[0]), _p(xy), int(xy.shape[0]),
                                 gray, dom.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), H, W,
                                 _p(out), ctypes.c_void_p(stream))
    L.check(rc)
    return [out]
