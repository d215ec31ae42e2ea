"""CPU oracle of the warm-start network (test infrastructure only).

The network maps the target density, box-downsampled to S x S (S = 128), to the dual weights of an
S x S site grid.  It is a four-level U-Net, NHWC layout:

    level 0 (S):    e0a 3x3 1->32, e0b 3x3 32->32                      -> skip0
    level 1 (S/2):  avg-pool 2x2, e1a 3x3 32->64, e1b 64->64          -> skip1
    level 2 (S/4):  avg-pool, e2a 64->128, e2b 128->128               -> skip2
    level 3 (S/8):  avg-pool, e3a 128->192, e3b 192->192
    decoder:        nearest 2x up, concat(up, skip) -> two 3x3 convs, per level
                    d2a 320->128, d2b 128->128, d1a 192->64, d1b 64->64, d0a 96->32, d0b 32->32
    output:         1x1 conv 32->1 (no activation)

Every conv has a bias; every conv but the output is followed by ReLU.  Zero padding.  The GPU
kernels run each conv as an implicit GEMM on BF16 MFMA with FP32 accumulation: both operands
(activations or gradients, and weights) are rounded to BF16 as they are loaded, products summed
in FP32.  This oracle does the same rounding (round to nearest even) and sums in float64, so the
two differ only by the FP32 summation order.  ``bf16=False`` gives the plain FP32 network
(no rounding), which finite differences and torch's CPU autograd pin.

There is no counterpart in the reference (it has no learned warm start); the network is this
build's own.
"""
from __future__ import annotations

import numpy as np

# (name, k, c_in, c_out, relu)
LAYERS = [
    ("e0a", 3, 1, 32, True), ("e0b", 3, 32, 32, True),
    ("e1a", 3, 32, 64, True), ("e1b", 3, 64, 64, True),
    ("e2a", 3, 64, 128, True), ("e2b", 3, 128, 128, True),
    ("e3a", 3, 128, 192, True), ("e3b", 3, 192, 192, True),
    ("d2a", 3, 320, 128, True), ("d2b", 3, 128, 128, True),
    ("d1a", 3, 192, 64, True), ("d1b", 3, 64, 64, True),
    ("d0a", 3, 96, 32, True), ("d0b", 3, 32, 32, True),
    ("out", 1, 32, 1, False),
]


def n_params():
    return sum(k * k * ci * co + co for _, k, ci, co, _ in LAYERS)


def bf16_round(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    # NaN/inf pass through unchanged
    bad = ~np.isfinite(a)
    if bad.any():
        out = out.copy()
        out[bad] = a[bad]
    return out


def init_params(seed=0):
    """He-normal weights (fan-in), zero biases; weights as [k*k*c_in, c_out] with the row index
    (ky * k + kx) * c_in + ci."""
    rng = np.random.default_rng(seed)
    P = {}
    for name, k, ci, co, relu in LAYERS:
        fan = k * k * ci
        std = np.sqrt(2.0 / fan) if relu else np.sqrt(1.0 / fan)
        P[name + ".w"] = (rng.standard_normal((fan, co)) * std).astype(np.float32)
        P[name + ".b"] = np.zeros(co, dtype=np.float32)
    return P


def _im2col(x, k):
    """x: [B, H, W, C] -> [B*H*W, k*k*C] (zero padding, row index (ky*k+kx)*C + c)."""
    B, H, W, C = x.shape
    if k == 1:
        return x.reshape(B * H * W, C)
    p = k // 2
    xp = np.zeros((B, H + 2 * p, W + 2 * p, C), dtype=x.dtype)
    xp[:, p:p + H, p:p + W, :] = x
    cols = np.empty((B, H, W, k * k, C), dtype=x.dtype)
    for ky in range(k):
        for kx in range(k):
            cols[:, :, :, ky * k + kx, :] = xp[:, ky:ky + H, kx:kx + W, :]
    return cols.reshape(B * H * W, k * k * C)


def _col2im(cols, shape, k):
    """Adjoint of _im2col."""
    B, H, W, C = shape
    if k == 1:
        return cols.reshape(B, H, W, C)
    p = k // 2
    c5 = cols.reshape(B, H, W, k * k, C)
    xp = np.zeros((B, H + 2 * p, W + 2 * p, C), dtype=cols.dtype)
    for ky in range(k):
        for kx in range(k):
            xp[:, ky:ky + H, kx:kx + W, :] += c5[:, :, :, ky * k + kx, :]
    return xp[:, p:p + H, p:p + W, :]


# Accumulation type of the BF16 products: float64 (exact sums of the rounded operands) or
# float32 (another summation order of FP32 sums, to measure the spread that order alone causes).
ACC = np.float64


def _mm(a, b, bf16):
    """Product with BF16-rounded operands (float32 result), or, bf16=False, the plain product in
    the operands' precision (float64 operands stay float64, for finite differences)."""
    if bf16:
        a, b = bf16_round(a), bf16_round(b)
        return (a.astype(ACC) @ b.astype(ACC)).astype(np.float32)
    dt = np.result_type(a.dtype, b.dtype)
    return (a.astype(np.float64) @ b.astype(np.float64)).astype(dt)


def conv_forward(x, w, b, k, relu, bf16=True):
    """One conv layer: returns the post-activation output [B, H, W, c_out]."""
    B, H, W, _ = x.shape
    y = _mm(_im2col(x, k), w, bf16) + b[None, :]
    if relu:
        y = np.maximum(y, 0.0)
    return y.reshape(B, H, W, -1).astype(y.dtype)


def conv_backward(x, w, y, dy, k, relu, bf16=True):
    """Gradients of one conv layer given its input x, output y (post-activation) and dL/dy.
    Returns (dx, dw, db).  The pre-activation gradient dz = dy * (y > 0) is formed in FP32; the
    three products round their operands to BF16 like the GPU's MFMA operands."""
    B, H, W, co = y.shape
    dz = dy if not relu else dy * (y > 0)
    dz2 = dz.reshape(B * H * W, co)
    cols = _im2col(x, k)
    dw = _mm(cols.T, dz2, bf16)
    db = dz2.astype(np.float64).sum(axis=0).astype(dw.dtype)
    dcols = _mm(dz2, w.T, bf16)
    dx = _col2im(dcols, x.shape, k).astype(dcols.dtype)
    return dx, dw, db


def pool(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4)).astype(x.dtype)


def pool_backward(dy):
    return (np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) * 0.25).astype(dy.dtype)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_backward(dy):
    B, H, W, C = dy.shape
    return dy.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4)).astype(dy.dtype)


def forward(P, x, bf16=True, keep=False):
    """x: [B, S, S, 1] -> [B, S, S, 1].  keep=True also returns the activations for backward."""
    acts = {"in": x}

    def cv(name, inp):
        _, k, _, _, relu = next(L for L in LAYERS if L[0] == name)
        out = conv_forward(inp, P[name + ".w"], P[name + ".b"], k, relu, bf16)
        acts[name] = out
        return out

    s0 = cv("e0b", cv("e0a", x))
    s1 = cv("e1b", cv("e1a", pool(s0)))
    s2 = cv("e2b", cv("e2a", pool(s1)))
    s3 = cv("e3b", cv("e3a", pool(s2)))
    acts["p0"], acts["p1"], acts["p2"] = pool(s0), pool(s1), pool(s2)
    c2 = np.concatenate([up(s3), s2], axis=3)
    acts["c2"] = c2
    t2 = cv("d2b", cv("d2a", c2))
    c1 = np.concatenate([up(t2), s1], axis=3)
    acts["c1"] = c1
    t1 = cv("d1b", cv("d1a", c1))
    c0 = np.concatenate([up(t1), s0], axis=3)
    acts["c0"] = c0
    t0 = cv("d0b", cv("d0a", c0))
    out = cv("out", t0)
    return (out, acts) if keep else out


# input of each layer in the activation dictionary
_INPUT = {"e0a": "in", "e0b": "e0a", "e1a": "p0", "e1b": "e1a", "e2a": "p1", "e2b": "e2a",
          "e3a": "p2", "e3b": "e3a", "d2a": "c2", "d2b": "d2a", "d1a": "c1", "d1b": "d1a",
          "d0a": "c0", "d0b": "d0a", "out": "d0b"}


def backward(P, acts, dout, bf16=True):
    """Gradients of every parameter given dL/d(out)."""
    G = {}
    spec = {L[0]: L for L in LAYERS}

    def bk(name, dy):
        _, k, _, _, relu = spec[name]
        dx, dw, db = conv_backward(acts[_INPUT[name]], P[name + ".w"], acts[name], dy, k, relu,
                                   bf16)
        G[name + ".w"], G[name + ".b"] = dw, db
        return dx

    d = bk("out", dout)
    d = bk("d0a", bk("d0b", d))
    C0 = acts["d1b"].shape[3]
    d_up1, d_s0 = d[..., :C0], d[..., C0:]
    d = bk("d1a", bk("d1b", up_backward(d_up1)))
    C1 = acts["d2b"].shape[3]
    d_up2, d_s1 = d[..., :C1], d[..., C1:]
    d = bk("d2a", bk("d2b", up_backward(d_up2)))
    C2 = acts["e3b"].shape[3]
    d_up3, d_s2 = d[..., :C2], d[..., C2:]
    d = bk("e3a", bk("e3b", up_backward(d_up3)))
    d_s2 = d_s2 + pool_backward(d)
    d = bk("e2a", bk("e2b", d_s2))
    d_s1 = d_s1 + pool_backward(d)
    d = bk("e1a", bk("e1b", d_s1))
    d_s0 = d_s0 + pool_backward(d)
    bk("e0a", bk("e0b", d_s0))
    return G


def field_loss(out, target, grad_weight=0.0):
    """Loss of a predicted field against a zero-mean target, the prediction's per-sample mean
    removed (the gauge): with e = out - mean(out) - target over the S x S image,
        L = (sum e^2 + grad_weight * sum over neighbouring pixel pairs (e_a - e_b)^2) / (B S^2).
    Returns (loss, dL/dout)."""
    B, S = out.shape[0], out.shape[1]
    o = out.reshape(B, S, S).astype(np.float64)
    t = target.reshape(B, S, S).astype(np.float64)
    e = o - o.mean(axis=(1, 2), keepdims=True) - t
    n = S * S
    dx = e[:, :, 1:] - e[:, :, :-1]
    dy = e[:, 1:, :] - e[:, :-1, :]
    loss = float(((e * e).sum() + grad_weight * ((dx * dx).sum() + (dy * dy).sum())) / (B * n))
    lap = np.zeros_like(e)
    lap[:, :, 1:] += dx
    lap[:, :, :-1] -= dx
    lap[:, 1:, :] += dy
    lap[:, :-1, :] -= dy
    de = (2.0 * e + 2.0 * grad_weight * lap) / (B * n)
    dout = de - de.mean(axis=(1, 2), keepdims=True)
    return loss, dout.reshape(out.shape).astype(out.dtype)


def mse_loss(out, target):
    """field_loss without the difference term."""
    return field_loss(out, target, 0.0)
