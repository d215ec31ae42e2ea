"""ctypes wrapper of tests/hostemu/libhostemu.so (the kernel code built for the host)."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "libhostemu.so")
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
        _lib.emu_evaluate.restype = ctypes.c_int
    return _lib


def evaluate(p, w, fp32=False, ring_init=2, max_repairs=12, verify_tol=1e-4, rings=None):
    """Evaluate problem p (workloads.Problem) at w with the host build of the cell kernel.

    rings: optional int32 array carried between calls (the library keeps per-site rings)."""
    n = p.n
    sites = np.ascontiguousarray(p.sites, dtype=np.float64)
    nu = np.ascontiguousarray(p.nu, dtype=np.float64)
    rho = np.ascontiguousarray(p.rho, dtype=np.float64)
    proj = np.ascontiguousarray(p.proj_domain, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = {k: np.zeros(n) for k in ("mass", "fcell")}
    out["centroid"] = np.zeros((n, 2))
    out["geo"] = np.zeros((n, 2))
    ex = np.zeros(n, dtype=np.uint8)
    stats = np.zeros(4, dtype=np.int32)
    ring = rings if rings is not None else np.zeros(n, dtype=np.int32)
    P = lambda a: a.ctypes.data_as(_dp)
    lib().emu_evaluate(ctypes.c_int(n), P(sites), P(nu), P(rho), ctypes.c_int(rho.shape[1]),
                       ctypes.c_int(rho.shape[0]), P(proj), P(w), ctypes.c_int(int(fp32)),
                       ctypes.c_int(ring_init), ctypes.c_int(max_repairs),
                       ctypes.c_double(verify_tol), P(out["mass"]), P(out["fcell"]),
                       P(out["centroid"]), P(out["geo"]), ex.ctypes.data_as(ctypes.c_void_p),
                       stats.ctypes.data_as(ctypes.c_void_p), ring.ctypes.data_as(ctypes.c_void_p))
    out["exist"] = ex.astype(bool)
    out["f"] = float(out["fcell"].sum())
    out["fails"], out["repairs"], out["overflow"], out["ring_max"] = (int(x) for x in stats)
    out["ring"] = ring
    return out
