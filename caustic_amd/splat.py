"""Gaussian-splat receiver rendering on the MI355X (utils/basic_function.py:674-772).

``render_gaussian_splat_two_pass_backward`` mirrors the reference function of that name (caller
utils/optimization.py:887-893): the same arguments, the same (rendered_image, loss) result, and
the same effect on autograd -- the gradient of the RMS image loss reaches ``valid_points``
through ``torch.autograd.backward``.  Both passes run in HIP (cot_splat in libcaustic_ot.so):
the splat itself and the points' gradient; the masks, the loss and its image gradient stay in
torch, as in the reference.  The kernels compute in FP32.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


def _splat(points, energy, res_y, res_x, dom, sigma, radius, out, image_grad=None, scale=1.0):
    import torch

    lib = L.load()
    dev = out.device
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        rc = lib.cot_splat(ctypes.c_void_p(points.data_ptr()), ctypes.c_void_p(energy.data_ptr()),
                           int(points.shape[0]), res_y, res_x,
                           dom.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(sigma),
                           int(radius),
                           None if image_grad is None else ctypes.c_void_p(image_grad.data_ptr()),
                           float(scale), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
    L.check(rc)


def render_gaussian_splat_two_pass_backward(loader, total_gray_value, valid_points, ray_energy,
                                            pic_size, proj_domain, target_image, target_rms,
                                            sigma=0.5, radius=2, chunk_size=300000):
    import torch
    import torch.nn.functional as F

    device, dtype = valid_points.device, valid_points.dtype
    res_y, res_x = int(pic_size[0]), int(pic_size[1])
    x_min, x_max, y_min, y_max = [float(v) for v in proj_domain]
    dom = np.array([x_min, x_max, y_min, y_max], dtype=np.float64)
    with torch.no_grad():
        legal = torch.isfinite(valid_points).all(dim=1) & torch.isfinite(ray_energy)
        inside = legal & (valid_points[:, 0] >= x_min) & (valid_points[:, 0] < x_max) & \
            (valid_points[:, 1] >= y_min) & (valid_points[:, 1] < y_max)
        idx = torch.nonzero(inside, as_tuple=False).squeeze(1)
        total_energy = ray_energy[legal].sum()
        pts = valid_points.detach()[idx].to(torch.float32).contiguous()
        en = ray_energy.detach()[idx].to(torch.float32).contiguous()
    gray = torch.as_tensor(total_gray_value, device=device, dtype=dtype).detach()
    rms = torch.as_tensor(target_rms, device=device, dtype=dtype).detach().clamp_min(1e-12)
    norm = gray / total_energy
    raw = torch.zeros((res_y, res_x), device=device, dtype=torch.float32)
    _splat(pts, en, res_y, res_x, dom, sigma, radius, raw)
    rendered = raw.to(dtype) * norm
    leaf = rendered.detach().requires_grad_(True)
    loss = torch.sqrt(F.mse_loss(leaf, target_image.detach(), reduction="mean")) / rms
    image_grad = torch.autograd.grad(loss, leaf)[0].detach().to(torch.float32).contiguous()
    if idx.numel() > 0 and valid_points.requires_grad:
        g = torch.empty((idx.numel(), 2), device=device, dtype=torch.float32)
        _splat(pts, en, res_y, res_x, dom, sigma, radius, g, image_grad, float(norm))
        grad_all = torch.zeros_like(valid_points)
        grad_all[idx] = g.to(dtype)
        torch.autograd.backward(valid_points, grad_all)
    return rendered.detach(), loss.detach()
