"""Learned warm start of the OT dual solve.

A U-Net on the GPU (libcaustic_wsn.so, include/caustic_wsn.h) maps the target density,
box-downsampled to S x S (S = 128), to the dual weights of an S x S jittered site grid.  The
solver (caustic_amd.ot) then starts L-BFGS from those weights instead of zeros; convergence is
still decided by the solve, and the zero start is kept whenever the predicted weights give a
larger mass residual than zeros do.

    net = WarmStartNet.load(DEFAULT_MODEL)
    w0 = learned_start(net, problem)            # any n_side x n_side jittered grid
    w, rep = OTBatch([problem], "mixed").solve(w0)

Weights are predicted as a field: w(s) ~ |s|^2 - 2 psi(s) with psi the Brenier potential of the
continuous problem, independent of the number of sites, so the S x S prediction is bilinearly
resampled for other grid sizes (256 x 256 for the N = 65,536 workload).

Training data comes from converged solves of shapes targets (disks and rectangles at 100:1
contrast), produced by the batched solver with a coarse-to-fine schedule: each problem is first
solved on coarser grids with box-downsampled targets and the weights interpolated upward, which
gives the same converged weights (the dual solution is unique up to a constant) in far fewer
fine-grid iterations than a zero start.
"""
from __future__ import annotations

import ctypes
import os
import time
from typing import Sequence

import numpy as np

from . import workloads as W

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libcaustic_wsn.so")
DEFAULT_MODEL = os.path.join(_PKG, "warmstart_unet128.npz")
S_DEFAULT = 128
# the learned start's low-pass width (grid nodes) and scale (learned_start)
LEARNED_SIGMA = 4.0
LEARNED_ALPHA = 0.5

_fp = ctypes.POINTER(ctypes.c_float)
_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m caustic_amd.build` "
                          "(the warm-start network has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "wsn_last_error": (ctypes.c_char_p, []),
        "wsn_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
        "wsn_destroy": (ctypes.c_int, [ctypes.c_void_p]),
        "wsn_param_count": (ctypes.c_longlong, [ctypes.c_void_p]),
        "wsn_set_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
        "wsn_get_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
        "wsn_forward": (ctypes.c_int, [ctypes.c_void_p, _fp, ctypes.c_int, _fp]),
        "wsn_gradients": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp, ctypes.c_int, _fp, _fp]),
        "wsn_set_dataset": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp, ctypes.c_int]),
        "wsn_set_loss": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
        "wsn_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                     ctypes.c_int, ctypes.c_float, _fp]),
        "wsn_train_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_longlong),
                                           ctypes.POINTER(ctypes.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"warm-start network error {rc}: {_load().wsn_last_error().decode()}")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data_as(_fp)


class WarmStartNet:
    """The U-Net on the current GPU.  Parameters are a flat float32 array (see caustic_wsn.h)."""

    def __init__(self, S: int = S_DEFAULT, max_batch: int = 32, seed: int | None = 0):
        lib = _load()
        h = ctypes.c_void_p()
        _check(lib.wsn_create(S, max_batch, ctypes.byref(h)))
        self._h = h
        self._lib = lib
        self.S = S
        self.max_batch = max_batch
        self.n_params = int(lib.wsn_param_count(h))
        # network output = target / t_scale (the training set's standard deviation)
        self.t_scale = 1.0
        if seed is not None:
            self.set_params(init_params(seed))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wsn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = _f32(flat)
        if flat.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {flat.size}")
        _check(self._lib.wsn_set_params(self._h, _p(flat)))

    def get_params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        _check(self._lib.wsn_get_params(self._h, _p(out)))
        return out

    def forward(self, x):
        """x: [B, S, S] float -> [B, S, S] (any B; split into batches of max_batch)."""
        x = _f32(x).reshape(-1, self.S, self.S)
        y = np.zeros_like(x)
        for b0 in range(0, x.shape[0], self.max_batch):
            xb = np.ascontiguousarray(x[b0:b0 + self.max_batch])
            yb = np.zeros_like(xb)
            _check(self._lib.wsn_forward(self._h, _p(xb), xb.shape[0], _p(yb)))
            y[b0:b0 + xb.shape[0]] = yb
        return y

    def gradients(self, x, t):
        x = _f32(x).reshape(-1, self.S, self.S)
        t = _f32(t).reshape(-1, self.S, self.S)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = ctypes.c_float()
        _check(self._lib.wsn_gradients(self._h, _p(x), _p(t), x.shape[0], _p(g),
                                       ctypes.byref(loss)))
        return g, float(loss.value)

    def set_loss(self, grad_weight: float):
        """Weight of the neighbouring-difference term of the training loss (0: plain MSE)."""
        _check(self._lib.wsn_set_loss(self._h, ctypes.c_float(grad_weight)))

    def set_dataset(self, x, t):
        x = _f32(x).reshape(-1, self.S, self.S)
        t = _f32(t).reshape(-1, self.S, self.S)
        _check(self._lib.wsn_set_dataset(self._h, _p(x), _p(t), x.shape[0]))
        self.n_data = x.shape[0]

    def train(self, order, batch: int, lr: float):
        order = np.ascontiguousarray(order, dtype=np.int32)
        steps = order.size // batch
        losses = np.zeros(max(steps, 1), dtype=np.float32)
        _check(self._lib.wsn_train(self._h, order.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   order.size, batch, ctypes.c_float(lr), _p(losses)))
        return losses[:steps]

    def train_stats(self):
        s, n, f = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
        _check(self._lib.wsn_train_stats(self._h, ctypes.byref(s), ctypes.byref(n),
                                         ctypes.byref(f)))
        return {"seconds": s.value, "samples": n.value, "flops_per_sample": f.value}

    def save(self, path, **meta):
        np.savez(path, params=self.get_params(), S=self.S, t_scale=self.t_scale, **meta)

    @classmethod
    def load(cls, path, max_batch: int = 8):
        with np.load(path, allow_pickle=False) as z:
            S = int(z["S"])
            net = cls(S=S, max_batch=max_batch, seed=None)
            net.set_params(z["params"])
            net.t_scale = float(z["t_scale"])
        return net


# ---------------------------------------------------------------------------------------------
# parameters (flat layout of caustic_wsn.h; same layer table as the CPU oracle of the network)
# ---------------------------------------------------------------------------------------------

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


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, flat."""
    rng = np.random.default_rng(seed)
    parts = []
    for _, k, ci, co, relu in LAYERS:
        fan = k * k * ci
        std = np.sqrt(2.0 / fan) if relu else np.sqrt(1.0 / fan)
        parts.append((rng.standard_normal((fan, co)) * std).astype(np.float32).ravel())
        parts.append(np.zeros(co, dtype=np.float32))
    return np.concatenate(parts)


def flat_to_dict(flat) -> dict:
    out, off = {}, 0
    for name, k, ci, co, _ in LAYERS:
        n = k * k * ci * co
        out[name + ".w"] = np.asarray(flat[off:off + n], dtype=np.float32).reshape(k * k * ci, co)
        off += n
        out[name + ".b"] = np.asarray(flat[off:off + co], dtype=np.float32)
        off += co
    return out


def dict_to_flat(d) -> np.ndarray:
    parts = []
    for name, *_ in LAYERS:
        parts.append(np.asarray(d[name + ".w"], dtype=np.float32).ravel())
        parts.append(np.asarray(d[name + ".b"], dtype=np.float32).ravel())
    return np.concatenate(parts)


# ---------------------------------------------------------------------------------------------
# features and targets
# ---------------------------------------------------------------------------------------------

def features(rho: np.ndarray, S: int = S_DEFAULT) -> np.ndarray:
    """Network input: log of the target box-downsampled to S x S, normalised to mean 1."""
    d = W.box_downsample(np.asarray(rho, dtype=np.float64), S)
    d = d / d.mean()
    return np.log(np.maximum(d, 1e-6)).astype(np.float32)


def grid_nodes(S: int, frame=W.UNIT_DOMAIN) -> np.ndarray:
    """Cell centres of an S x S grid of the frame (x0, x1, y0, y1), in site order (row-major,
    y first)."""
    k = np.arange(S * S)
    x0, x1, y0, y1 = frame
    return np.stack([x0 + (x1 - x0) * (k % S + 0.5) / S, y0 + (y1 - y0) * (k // S + 0.5) / S],
                    axis=1)


def to_frame(xy: np.ndarray, frame=W.UNIT_DOMAIN) -> np.ndarray:
    """Coordinates in the frame's unit square."""
    x0, x1, y0, y1 = frame
    return np.stack([(xy[:, 0] - x0) / (x1 - x0), (xy[:, 1] - y0) / (y1 - y0)], axis=1)


def field_from_solution(sites: np.ndarray, w: np.ndarray, cen: np.ndarray, S: int,
                        frame=W.UNIT_DOMAIN) -> np.ndarray:
    """The weight field U (w_i = -U(s_i)) sampled at the S x S grid nodes of the frame (the
    region the sites sample, Problem.frame), for a jittered grid whose site k sits near node k.  With psi the Brenier potential, w(s) = |s|^2 - 2 psi(s), so
    U(s) = 2 psi(s) - |s|^2 is smooth with gradient -2 (s - c(s)), c(s) the centroid of the cell
    of s; hence U(node_k) = -w_k + 2 (s_k - c_k).(s_k - node_k) to first order.  The mean is
    removed (the gauge)."""
    nodes = grid_nodes(S, frame)
    u = -np.asarray(w, dtype=np.float64) + 2.0 * np.sum((sites - cen) * (sites - nodes), axis=1)
    return (u - u.mean()).reshape(S, S)


def weights_from_field(U: np.ndarray, sites: np.ndarray, frame=W.UNIT_DOMAIN) -> np.ndarray:
    """w_i = -U(s_i), bilinear in the field given on the frame's grid nodes, mean removed."""
    w = -resample_field(np.asarray(U, dtype=np.float64), to_frame(sites, frame))
    return w - w.mean()


def field_to_target(U: np.ndarray) -> np.ndarray:
    """Network target: the field in units of the grid spacing squared."""
    S = U.shape[0]
    return (np.asarray(U) * S * S).astype(np.float32)


def target_to_field(t: np.ndarray) -> np.ndarray:
    S = t.shape[-1]
    U = np.asarray(t, dtype=np.float64).reshape(S, S) / (S * S)
    return U - U.mean()


def resample_field(f: np.ndarray, xy: np.ndarray) -> np.ndarray:
    """Bilinear interpolation of a field given at the S x S cell centres of the unit square
    (f[iy, ix] at ((ix + 1/2) / S, (iy + 1/2) / S)) at points xy, clamped at the border."""
    S = f.shape[0]
    u = np.clip(xy[:, 0] * S - 0.5, 0.0, S - 1.0)
    v = np.clip(xy[:, 1] * S - 0.5, 0.0, S - 1.0)
    i0 = np.minimum(np.floor(u).astype(np.int64), S - 2)
    j0 = np.minimum(np.floor(v).astype(np.int64), S - 2)
    a, b = u - i0, v - j0
    return ((1 - a) * (1 - b) * f[j0, i0] + a * (1 - b) * f[j0, i0 + 1]
            + (1 - a) * b * f[j0 + 1, i0] + a * b * f[j0 + 1, i0 + 1])


def predict_fields(net: WarmStartNet, problems) -> list:
    """The network's weight fields (S x S, see field_from_solution) for the problems' targets."""
    x = np.stack([features(p.rho, net.S) for p in problems])
    y = net.forward(x)
    return [target_to_field(yk * net.t_scale) for yk in y]


def smooth_field(U: np.ndarray, sigma: float) -> np.ndarray:
    """The field low-passed by a Gaussian of sigma grid nodes (edges extended), mean removed."""
    from scipy.ndimage import gaussian_filter

    f = gaussian_filter(np.asarray(U, dtype=np.float64), sigma, mode="nearest")
    return f - f.mean()


def learned_start(net: WarmStartNet, problems, sigma: float = LEARNED_SIGMA,
                  alpha: float = LEARNED_ALPHA) -> list:
    """Start weights from the network for problems on jittered grids of the unit square: the
    predicted field low-passed (smooth_field) and scaled by alpha, sampled at the sites.

    The prediction is right at large scales (on a held-out target it takes the dual objective
    from -1.1e-5 at zero weights to -1.7e-2) but its errors at the scale of a few grid nodes
    move bisectors by more than a cell: taken as it is, it empties most cells (12,890 of 16,384 on a held-out
    target), and evaluations with that many empty cells cost the solver far more than the zero
    start.  Low-passed over 4 nodes and halved, it leaves ~70-100 empty cells and starts
    L-BFGS closer to the solution than zeros do."""
    single = not isinstance(problems, (list, tuple))
    probs = [problems] if single else list(problems)
    out = [alpha * weights_from_field(smooth_field(U, sigma), p.sites)
           for U, p in zip(predict_fields(net, probs), probs)]
    return out[0] if single else out


def predict_weights(net: WarmStartNet, problems) -> list:
    """Predicted dual weights for problems on jittered square grids in the unit square (any
    side): the network's field sampled at the sites."""
    single = not isinstance(problems, (list, tuple))
    probs = [problems] if single else list(problems)
    out = [weights_from_field(U, p.sites) for U, p in zip(predict_fields(net, probs), probs)]
    return out[0] if single else out


# ---------------------------------------------------------------------------------------------
# training data: coarse-to-fine batched solves
# ---------------------------------------------------------------------------------------------

def coarse_problem(p: W.Problem, n_side: int, seed: int) -> W.Problem:
    """The problem on an n_side x n_side jittered grid of the fine problem's site frame (the
    region its sites sample), with the target box-downsampled by the largest factor that divides
    the image side and keeps at least the fine problem's pixels per site and 32 x 32 pixels."""
    fine_side = int(round(np.sqrt(p.n)))
    side = p.rho.shape[0]
    f = max(1, min(fine_side // max(n_side, 1), side // 32))
    while side % f:
        f -= 1
    rho = W.box_downsample(p.rho, side // f) if f > 1 else p.rho
    x0, x1, y0, y1 = p.frame
    sites = W.jittered_grid(n_side, W.rng_for(seed)) * [x1 - x0, y1 - y0] + [x0, y0]
    cp = W.make_problem(sites, rho / rho.sum(), p.tol, p.name + f"_c{n_side}", seed)
    cp.proj_domain = tuple(p.proj_domain)
    cp.site_box = p.site_box
    return cp


def _solve_level(problems, init, tol, precision, max_iter, tuning=None, verbose=0,
                 newton_from=None, newton_coarse=256):
    from .ot import OTBatch

    with OTBatch(list(problems), precision=precision, tuning=tuning) as b:
        w, rep = b.solve(None if init is None else np.concatenate(init), history=10, tol=tol,
                         stop="maxrel", h0_threshold=0.0, max_iter=max_iter,
                         raise_on_fail=False, verbose=verbose, newton_from=newton_from,
                         newton_coarse=newton_coarse)
        cen, _ = b.centroids(w, weighted=True)
        ws = [x.copy() for x in b.split(w)]
        cs = [c.copy() for c in b.split(cen.reshape(-1, 2))]
    return ws, cs, rep


def coarse_problems(problems, levels):
    """The coarse problems of every level, as multiscale_solve builds them: a list per level."""
    return [[coarse_problem(p, n_side, 7919 * (li + 1) + k) for k, p in enumerate(problems)]
            for li, n_side in enumerate(levels)]


def multiscale_solve(problems: Sequence[W.Problem], levels=(32, 64), precision="mixed",
                     coarse_tol=1e-3, tol=1e-5, max_iter=20000, fields=None, log=print,
                     tunings=None, verbose=0, newton_from=None, newton_coarse=256):
    """Solve every problem to max-rel `tol`, one batch per level: coarse grids first (to
    `coarse_tol`), each level's weights carried to the next level's sites through the weight
    field (field_from_solution / weights_from_field).  fields: optional per-problem fields that
    start the first level (a learned start); levels=() with fields starts the final level from
    them.  tunings: optional launch shapes (caustic_amd.autotune) keyed by level side or
    "final".  verbose > 0: the solver prints a progress line every 2 s.  newton_from: every
    level runs L-BFGS to this max-rel, then damped Newton steps (OTBatch.solve), with
    newton_coarse patches per problem in the preconditioner's coarse term.  Returns
    (weights, centroids, reports) of the final level."""
    tunings = tunings or {}
    probs = list(problems)
    for li, n_side in enumerate(levels):
        cps = [coarse_problem(p, n_side, 7919 * (li + 1) + k) for k, p in enumerate(probs)]
        init = None if fields is None else [weights_from_field(U, cp.sites, cp.frame)
                                            for U, cp in zip(fields, cps)]
        t0 = time.perf_counter()
        ws, cs, rep = _solve_level(cps, init, coarse_tol, precision, max_iter,
                                   tunings.get(n_side), verbose, newton_from, newton_coarse)
        fields = [field_from_solution(cp.sites, wk, ck, n_side, cp.frame)
                  for cp, wk, ck in zip(cps, ws, cs)]
        log(f"[multiscale] level {n_side}: {len(cps)} problems, "
            f"iters max {max(r['iterations'] for r in rep)} "
            f"mean {np.mean([r['iterations'] for r in rep]):.0f}, "
            f"{time.perf_counter() - t0:.1f}s")
    init = None if fields is None else [weights_from_field(U, p.sites, p.frame)
                                        for U, p in zip(fields, probs)]
    t0 = time.perf_counter()
    ws, cs, rep = _solve_level(probs, init, tol, precision, max_iter, tunings.get("final"),
                               verbose, newton_from, newton_coarse)
    log(f"[multiscale] final: {len(probs)} problems, "
        f"iters max {max(r['iterations'] for r in rep)} "
        f"mean {np.mean([r['iterations'] for r in rep]):.0f}, converged "
        f"{sum(r['converged'] for r in rep)}, {time.perf_counter() - t0:.1f}s")
    return ws, cs, rep


# ---------------------------------------------------------------------------------------------
# training pipeline: python -m caustic_amd.warmstart --out caustic_amd/warmstart_unet128.npz
# ---------------------------------------------------------------------------------------------

def dihedral(a: np.ndarray, k: int) -> np.ndarray:
    """The k-th of the 8 symmetries of the square applied to the last two axes."""
    if k & 4:
        a = np.swapaxes(a, -1, -2)
    if k & 1:
        a = a[..., ::-1]
    if k & 2:
        a = a[..., ::-1, :]
    return np.ascontiguousarray(a)


def generate(n: int, seed0: int, batch: int = 64, budget_s: float = 1e9, S: int = S_DEFAULT,
             res: int = 512, log=print):
    """Converged (max-rel 1e-5, mixed precision) weights of n shapes problems on S x S grids."""
    xs, ts, seeds = [], [], []
    t0 = time.perf_counter()
    for b0 in range(0, n, batch):
        if time.perf_counter() - t0 > budget_s:
            log(f"[generate] time budget reached after {len(xs)} problems")
            break
        probs = [W.warm_start_sample(seed0 + k, n_side=S, res=res)
                 for k in range(b0, min(b0 + batch, n))]
        ws, cs, rep = multiscale_solve(probs, levels=(S // 4, S // 2), log=log)
        for p, w, c, r in zip(probs, ws, cs, rep):
            if not r["converged"]:
                continue
            xs.append(features(p.rho, S))
            ts.append(field_to_target(field_from_solution(p.sites, w, c, S)))
            seeds.append(p.seed)
        log(f"[generate] {len(xs)} / {n} problems, {time.perf_counter() - t0:.0f}s")
    return np.stack(xs), np.stack(ts), np.array(seeds)


def train(net: WarmStartNet, x: np.ndarray, t: np.ndarray, epochs: int, batch: int = 16,
          lr: float = 1e-3, lr_min: float = 1e-5, augment: bool = True, budget_s: float = 1e9,
          grad_weight: float = 100.0, seed: int = 0, log=print):
    """Adam with a cosine learning-rate schedule on the field loss (values plus grad_weight times
    neighbouring differences: the differences decide the cells); the 8 symmetries of the square
    as augmentation (the training set is stored with all of them).  Returns the per-epoch mean
    losses."""
    rng = np.random.default_rng(seed)
    net.set_loss(grad_weight)
    if augment:
        x = np.concatenate([dihedral(x, k) for k in range(8)])
        t = np.concatenate([dihedral(t, k) for k in range(8)])
    net.t_scale = float(t.std())
    net.set_dataset(x, t / net.t_scale)
    n = x.shape[0]
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        lr_e = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * ep / max(epochs - 1, 1)))
        losses = net.train(rng.permutation(n), batch, lr_e)
        hist.append(float(np.mean(losses)))
        log(f"[train] epoch {ep + 1}/{epochs} lr {lr_e:.2e} loss {hist[-1]:.4e} "
            f"({time.perf_counter() - t0:.0f}s)")
        if time.perf_counter() - t0 > budget_s:
            log("[train] time budget reached")
            break
    return hist


def residual(problems, ws, precision="mixed", batch=None):
    """|mass - nu|_2 of each problem at the given weights (the reference's residual,
    ot_op.cuh:468-476), on the GPU; batch: an OTBatch of these problems to reuse."""
    from .ot import OTBatch

    own = batch is None
    b = OTBatch(list(problems), precision=precision) if own else batch
    try:
        ev = b.evaluate(np.concatenate(ws), allow_unverified=True)
        ms = b.split(ev["mass"])
    finally:
        if own:
            b.close()
    if not ev["verified"]:
        if len(problems) == 1:  # cells too wild to resolve: no better than anything
            return [float("inf")]
        # the batch's check reports the batch: find the problems whose cells verify, one by one
        return [residual([p], [w], precision)[0] for p, w in zip(problems, ws)]
    return [float(np.linalg.norm(m - p.nu)) for m, p in zip(ms, problems)]


def choose_start(problems, w_pred, precision="mixed", batch=None):
    """The predicted weights where their mass residual |mass - nu|_2 is below the zero start's,
    else zeros (the cold start is the fallback).  Returns (starts, residuals of the predictions,
    residuals of zeros)."""
    r_pred = residual(problems, w_pred, precision, batch)
    r_zero = residual(problems, [np.zeros(p.n) for p in problems], precision, batch)
    return [w if a < b else np.zeros_like(w) for w, a, b in zip(w_pred, r_pred, r_zero)], \
        r_pred, r_zero


def evaluate_warm_start(net: WarmStartNet, seeds, S: int = S_DEFAULT, res: int = 512,
                        precision="mixed", newton_from: float = 1.0, log=print):
    """Held-out problems solved to max-rel 1e-5 one at a time, each start in a fresh context
    (both pay the same set-up): from zeros (the cold start) and from the network
    (learned_start; zeros instead when its |mass - nu|_2 is the larger, choose_start), with
    L-BFGS until max-rel newton_from and damped Newton steps after.  Iterations and wall-clock
    of each; the learned start's time includes the network, the low-pass and the residual
    comparison."""
    from .ot import OTBatch

    rows = []
    kw = dict(history=10, tol=1e-5, stop="maxrel", h0_threshold=0.0, max_iter=20000,
              raise_on_fail=False, newton_from=newton_from)
    for sd in seeds:
        p = W.warm_start_sample(int(sd), n_side=S, res=res)
        with OTBatch([p], precision=precision) as b:
            t0 = time.perf_counter()
            wc, rc = b.solve(None, **kw)
            t_cold = time.perf_counter() - t0
        with OTBatch([p], precision=precision) as b:
            t0 = time.perf_counter()
            (w0,), (rp,), (rz,) = choose_start([p], [learned_start(net, p)], precision, b)
            ww, rw = b.solve(w0, **kw)
            t_warm = time.perf_counter() - t0
        rows.append({"seed": int(sd), "cold_iters": rc[0]["iterations"], "cold_s": t_cold,
                     "warm_iters": rw[0]["iterations"], "warm_s": t_warm,
                     "pred_residual": rp, "zero_residual": rz, "learned_used": bool(rp < rz),
                     "converged": bool(rc[0]["converged"] and rw[0]["converged"]),
                     "w_cold": wc, "w_warm": ww})
        log(f"[eval] seed {sd}: cold {rows[-1]['cold_iters']} it {t_cold:.2f}s, learned "
            f"{rows[-1]['warm_iters']} it {t_warm:.2f}s (start residual {rp:.3g} vs {rz:.3g})")
    return rows


def evaluate_warm_start_batched(net: WarmStartNet, seeds, batch: int = 64,
                                S: int = S_DEFAULT, res: int = 512, precision="mixed",
                                newton_from: float = 2.0, log=print):
    """Held-out problems solved to max-rel 1e-5 in batches (one context per batch and start):
    from zeros, from the network (learned_start, zeros per problem where its |mass - nu|_2 is
    the larger: choose_start) and, for reference, by the library's coarse-to-fine ladder
    (cot_solve_ladder, no learning).  Per problem: the final stage's iterations of each start;
    per batch: wall-clock of each (the learned start's includes the network, the low-pass and
    the residual comparison; every problem of a batch finishes with the batch)."""
    from .ot import OTBatch

    kw = dict(history=10, tol=1e-5, stop="maxrel", h0_threshold=0.0, max_iter=20000,
              raise_on_fail=False, newton_from=newton_from)
    rows, batches = [], []
    seeds = [int(v) for v in seeds]
    for b0 in range(0, len(seeds), batch):
        sd = seeds[b0:b0 + batch]
        probs = [W.warm_start_sample(k, n_side=S, res=res) for k in sd]
        with OTBatch(probs, precision=precision) as b:
            t0 = time.perf_counter()
            _, rc = b.solve(None, **kw)
            t_zero = time.perf_counter() - t0
        with OTBatch(probs, precision=precision) as b:
            t0 = time.perf_counter()
            w0, rp, rz = choose_start(probs, learned_start(net, probs), precision, b)
            _, rw = b.solve(np.concatenate(w0), **kw)
            t_learn = time.perf_counter() - t0
        with OTBatch(probs, precision=precision) as b:
            t0 = time.perf_counter()
            _, rl = b.solve_ladder(tol=1e-5, raise_on_fail=False)
            t_ladder = time.perf_counter() - t0
        for k, seed in enumerate(sd):
            rows.append({"seed": seed, "zero_iters": rc[k]["iterations"],
                         "learned_iters": rw[k]["iterations"],
                         "ladder_final_iters": rl[k]["iterations"],
                         "learned_used": bool(rp[k] < rz[k]),
                         "converged": bool(rc[k]["converged"] and rw[k]["converged"]
                                           and rl[k]["converged"])})
        batches.append({"problems": len(sd), "zero_s": t_zero, "learned_s": t_learn,
                        "ladder_s": t_ladder})
        log(f"[eval batch] {len(sd)} problems: zero {t_zero:.2f}s, learned {t_learn:.2f}s, "
            f"ladder {t_ladder:.2f}s")
    return rows, batches


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="train the warm-start network on the GPU")
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--samples", type=int, default=2048)
    ap.add_argument("--gen-budget", type=float, default=480.0, help="seconds for data generation")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--train-budget", type=float, default=300.0)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--grad-weight", type=float, default=100.0)
    ap.add_argument("--eval", type=int, default=16, help="held-out problems to evaluate")
    ap.add_argument("--log", default=None, help="JSON summary path")
    ap.add_argument("--eval-only", default=None, help="evaluate this saved model, no training")
    ap.add_argument("--throughput", type=int, default=0,
                    help="train on this many synthetic S x S pairs (no solver, nothing saved) and "
                         "report the training kernels' TFLOP/s")
    args = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    if args.throughput:
        # smooth random inputs and targets of the production shape: the training kernels'
        # speed does not depend on the values
        rng = np.random.default_rng(0)
        c = (np.arange(S_DEFAULT) + 0.5) / S_DEFAULT
        X, Y = np.meshgrid(c, c, indexing="xy")
        xs = np.stack([np.sin(2 * np.pi * (a * X + b * Y)) for a, b in
                       rng.uniform(0.5, 3.0, (args.throughput, 2))]).astype(np.float32)
        ts = (xs ** 2 - (xs ** 2).mean(axis=(1, 2), keepdims=True)).astype(np.float32)
        net = WarmStartNet(S=S_DEFAULT, max_batch=args.batch, seed=0)
        train(net, xs, ts, args.epochs, args.batch, args.lr, augment=False,
              grad_weight=args.grad_weight, log=log)
        st = net.train_stats()
        log(f"[throughput] {st['samples']} samples in {st['seconds']:.2f}s: "
            f"{st['samples'] * st['flops_per_sample'] / st['seconds'] / 1e12:.1f} TFLOP/s "
            f"({st['flops_per_sample'] / 1e9:.2f} GFLOP per sample, forward + backward)")
        net.close()
        return 0
    if args.eval_only:
        rows = evaluate_warm_start(WarmStartNet.load(args.eval_only),
                                   range(900000, 900000 + args.eval), log=log)
        rows = [{k: v for k, v in r.items() if not k.startswith("w_")} for r in rows]
        if args.log:
            with open(args.log, "w") as f:
                json.dump({"eval": rows}, f, indent=1)
        return 0
    x, t, seeds = generate(args.samples, seed0=100000, budget_s=args.gen_budget, log=log)
    net = WarmStartNet(S=S_DEFAULT, max_batch=args.batch, seed=0)
    hist = train(net, x, t, args.epochs, args.batch, args.lr, budget_s=args.train_budget,
                 grad_weight=args.grad_weight, log=log)
    net.save(args.out, train_seeds=seeds, loss_history=np.array(hist),
             grad_weight=args.grad_weight)
    st = net.train_stats()
    log(f"[train] {st['samples']} samples in {st['seconds']:.1f}s: "
        f"{st['samples'] * st['flops_per_sample'] / st['seconds'] / 1e12:.1f} TFLOP/s")
    rows = evaluate_warm_start(WarmStartNet.load(args.out),
                               range(900000, 900000 + args.eval), log=log)
    rows = [{k: v for k, v in r.items() if not k.startswith("w_")} for r in rows]
    summary = {"train_problems": int(len(seeds)), "augmented_pairs": int(8 * len(seeds)),
               "loss_history": hist, "train_stats": st, "eval": rows,
               "median_iter_ratio": float(np.median([r["cold_iters"] / max(r["warm_iters"], 1)
                                                     for r in rows])),
               "median_time_ratio": float(np.median([r["cold_s"] / r["warm_s"] for r in rows]))}
    log(f"[eval] median cold/warm iterations {summary['median_iter_ratio']:.2f}, "
        f"time {summary['median_time_ratio']:.2f}")
    if args.log:
        with open(args.log, "w") as f:
            json.dump(summary, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
