"""ctypes wrapper of tests/hostemu/libhostemu.so (the kernel code built for the host)."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# COT_HOSTEMU_LIB names another build of cot_hostemu.cpp (the sanitizer build of the CPU tests)
_PATH = os.environ.get("COT_HOSTEMU_LIB") or os.path.join(_HERE, "libhostemu.so")
_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run([os.path.join(_HERE, "build.sh")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            build()
        _lib = ctypes.CDLL(_PATH)
        _lib.emu_create.restype = ctypes.c_void_p
        _lib.emu_create.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        _lib.emu_destroy.argtypes = [ctypes.c_void_p]
        _lib.emu_xcd_slot.argtypes = [ctypes.c_int, ctypes.c_int]
        _lib.emu_prefix_mismatches.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib.emu_state.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 5
        _lib.emu_coarse_inverse.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib.emu_set_adjacency.argtypes = [ctypes.c_void_p] * 3
        _lib.emu_eval.restype = ctypes.c_int
        _lib.emu_eval.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, ctypes.c_void_p,
                                  ctypes.c_void_p]
    return _lib


def _P(a):
    return a.ctypes.data_as(_dp)


class Session:
    """One problem with the library's per-site state (bucket rings, previous adjacency, FP32
    frames) carried from one evaluation to the next."""

    def __init__(self, p, fp32=False, ring_init=2, max_repairs=12, verify_tol=1e-4):
        self.n = p.n
        self._keep = [np.ascontiguousarray(p.sites, dtype=np.float64),
                      np.ascontiguousarray(p.nu, dtype=np.float64),
                      np.ascontiguousarray(p.rho, dtype=np.float64),
                      np.ascontiguousarray(p.proj_domain, dtype=np.float64)]
        s, nu, rho, proj = self._keep
        self.h = lib().emu_create(p.n, _P(s), _P(nu), _P(rho), rho.shape[1], rho.shape[0],
                                  _P(proj), int(fp32), ring_init, max_repairs, verify_tol)

    def evaluate(self, w):
        n = self.n
        w = np.ascontiguousarray(w, dtype=np.float64)
        out = {k: np.zeros(n) for k in ("mass", "fcell")}
        out["centroid"] = np.zeros((n, 2))
        out["geo"] = np.zeros((n, 2))
        ex = np.zeros(n, dtype=np.uint8)
        stats = np.zeros(6, dtype=np.int32)
        lib().emu_eval(self.h, _P(w), _P(out["mass"]), _P(out["fcell"]), _P(out["centroid"]),
                       _P(out["geo"]), ex.ctypes.data_as(ctypes.c_void_p),
                       stats.ctypes.data_as(ctypes.c_void_p))
        out["exist"] = ex.astype(bool)
        out["f"] = float(out["fcell"].sum())
        (out["fails"], out["repairs"], out["overflow"], out["ring_max"],
         out["rebuilt"]) = (int(x) for x in stats[:5])
        out["covered"] = bool(stats[5])
        return out

    def set_adjacency(self, neighbours):
        """The next evaluation's candidates: neighbours[i] lists site i's (caller indices), as
        if a previous evaluation's cells had had these edges."""
        K = int(lib().emu_max_verts())
        counts = np.array([len(v) for v in neighbours], dtype=np.int32)
        labels = np.full((self.n, K), -1, dtype=np.int32)
        for i, v in enumerate(neighbours):
            labels[i, :len(v)] = v
        lib().emu_set_adjacency(self.h, counts.ctypes.data_as(ctypes.c_void_p),
                                labels.ctypes.data_as(ctypes.c_void_p))

    def state(self):
        """Per-site state after the last evaluation: conformity flags, edge counts, edge labels
        (caller indices; sides below zero), spans, bucket rings."""
        n, K = self.n, int(lib().emu_max_verts())
        redo = np.zeros(n, dtype=np.uint8)
        ne = np.zeros(n, dtype=np.int32)
        lab = np.zeros((n, K), dtype=np.int32)
        span = np.zeros((n, K, 2), dtype=np.float32)
        ring = np.zeros(n, dtype=np.int32)
        v = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        lib().emu_state(self.h, v(redo), v(ne), v(lab), v(span), v(ring))
        return {"redo": redo.astype(bool), "nedge": ne, "elab": lab, "espan": span, "ring": ring}

    def close(self):
        if self.h:
            lib().emu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def evaluate(p, w, fp32=False, ring_init=2, max_repairs=12, verify_tol=1e-4):
    """One evaluation in a fresh session (bucket-ring candidates)."""
    s = Session(p, fp32, ring_init, max_repairs, verify_tol)
    try:
        return s.evaluate(w)
    finally:
        s.close()


def xcd_slot(w, G):
    """The kernels' workgroup renumbering (cot_internal.h xcd_slot)."""
    return int(lib().emu_xcd_slot(int(w), int(G)))


def prefix_mismatches(rho, fp32):
    """Pixels where the density groups' prefix lookup differs in any bit from the sequential
    chunk-local prefix sums (or chunk totals from the chunk sums)."""
    r = np.ascontiguousarray(rho, dtype=np.float64)
    return int(lib().emu_prefix_mismatches(_P(r), r.shape[1], r.shape[0], int(fp32)))


def coarse_inverse(a):
    """The library's coarse-matrix inverse (cot_layout.h) on a copy of the n x n array a."""
    m = np.ascontiguousarray(a, dtype=np.float64).copy()
    lib().emu_coarse_inverse(m.ctypes.data, m.shape[0])
    return m
