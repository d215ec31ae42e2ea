"""ctypes binding of libcaustic_ot.so (include/caustic_ot.h).

There is no CPU fallback: if the HIP library is missing this raises, so a caller can never
silently run something other than the GPU path.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libcaustic_ot.so")

COT_OK, COT_EINVAL, COT_ENOMEM, COT_EHIP, COT_ENOTCONV, COT_ECOMPLEX, COT_EUNVERIFIED = range(7)
COT_PREC_FP64, COT_PREC_MIXED = 0, 1
COT_STOP_REFERENCE, COT_STOP_MAXREL, COT_STOP_HANDOFF = 0, 1, 2
COT_MAX_VERTS = 48  # vertex slots per cell in cot_cell_polygons (include/caustic_ot.h)
COT_NOALLOC_MAX_PIXELS = 1 << 20  # Sobel image limit of cot_irra_render_grad_noalloc

STATUS_NAMES = {
    COT_OK: "ok", COT_EINVAL: "invalid argument", COT_ENOMEM: "out of device memory",
    COT_EHIP: "HIP error", COT_ENOTCONV: "not converged", COT_ECOMPLEX: "cell too complex",
    COT_EUNVERIFIED: "conformity check failed",
}

_dp = ctypes.POINTER(ctypes.c_double)
_up = ctypes.POINTER(ctypes.c_ubyte)


class cot_problem(ctypes.Structure):
    _fields_ = [("n_sites", ctypes.c_int), ("sites", _dp), ("nu", _dp), ("rho", _dp),
                ("pm", ctypes.c_int), ("pn", ctypes.c_int), ("proj_domain", ctypes.c_double * 4)]


class cot_options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("sites_per_bucket", ctypes.c_double),
                ("ring_init", ctypes.c_int), ("ring_max", ctypes.c_int),
                ("verify_tol", ctypes.c_double), ("max_repairs", ctypes.c_int),
                ("int_lanes", ctypes.c_int), ("repair_blocks", ctypes.c_int),
                ("full_pass_ring", ctypes.c_int), ("repair_mode", ctypes.c_int),
                ("pass_mode", ctypes.c_int)]


class cot_solve_opts(ctypes.Structure):
    _fields_ = [("history", ctypes.c_int), ("max_iter", ctypes.c_int), ("tol", ctypes.c_double),
                ("initial_step", ctypes.c_double), ("stop_rule", ctypes.c_int),
                ("h0_threshold", ctypes.c_double), ("verbose", ctypes.c_int),
                ("device_loop", ctypes.c_int), ("accept_rule", ctypes.c_int),
                ("repair_rounds", ctypes.c_int), ("newton_sweeps", ctypes.c_int),
                ("newton_coarse", ctypes.c_int), ("small_cg", ctypes.c_int)]


class cot_report(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("evaluations", ctypes.c_int),
                ("converged", ctypes.c_int), ("status", ctypes.c_int), ("f", ctypes.c_double),
                ("grad_norm", ctypes.c_double), ("max_rel_err", ctypes.c_double),
                ("seconds", ctypes.c_double)]


class cot_reflection(ctypes.Structure):
    _fields_ = [("points", _dp), ("inside_dir", _dp), ("height", _dp), ("clip_area", _dp),
                ("z_of_len", ctypes.c_double), ("z_of_receiver", ctypes.c_double),
                ("refraction_index", ctypes.c_double), ("area_scalar", ctypes.c_double),
                ("aa_history_size", ctypes.c_int), ("aa_fixed_para", ctypes.c_double),
                ("if_aa_op_initial", ctypes.c_int), ("incident_dir", _dp)]


COT_LADDER_MAX = 8


class cot_ladder_opts(ctypes.Structure):
    _fields_ = [("n_levels", ctypes.c_int), ("sides", ctypes.c_int * COT_LADDER_MAX),
                ("floors", ctypes.c_double * COT_LADDER_MAX), ("n_final_floors", ctypes.c_int),
                ("final_floors", ctypes.c_double * 4), ("coarse_tol", ctypes.c_double),
                ("handoff", ctypes.c_double), ("seed", ctypes.c_ulonglong), ("frames", _dp),
                ("level_options", ctypes.POINTER(cot_options))]


# name -> (restype, argtypes); every symbol include/caustic_ot.h declares
SIGNATURES = {
    "cot_default_options": (None, [ctypes.POINTER(cot_options)]),
    "cot_default_solve_opts": (None, [ctypes.POINTER(cot_solve_opts)]),
    "cot_create": (ctypes.c_int, [ctypes.POINTER(cot_problem), ctypes.c_int,
                                  ctypes.POINTER(cot_options), ctypes.POINTER(ctypes.c_void_p)]),
    "cot_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "cot_last_error": (ctypes.c_char_p, []),
    "cot_total_sites": (ctypes.c_int, [ctypes.c_void_p]),
    "cot_evaluate": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp, _up]),
    "cot_solve": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(cot_solve_opts),
                                 ctypes.POINTER(cot_report)]),
    "cot_solve_reflection": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(cot_solve_opts),
                                            ctypes.POINTER(cot_reflection),
                                            ctypes.POINTER(cot_report)]),
    "cot_solve_newton": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(cot_solve_opts),
                                        ctypes.POINTER(cot_report)]),
    "cot_default_ladder_opts": (None, [ctypes.POINTER(cot_problem), ctypes.c_int,
                                       ctypes.POINTER(cot_ladder_opts)]),
    "cot_solve_ladder": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(cot_problem), _dp,
                                        ctypes.POINTER(cot_solve_opts),
                                        ctypes.POINTER(cot_ladder_opts),
                                        ctypes.POINTER(cot_report)]),
    "cot_hessian_product": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "cot_render": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "cot_cell_polygons": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_int),
                                         _dp]),
    "cot_centroids": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int, _dp, _up]),
    "cot_cell_areas": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "cot_evaluate_dev": (ctypes.c_int, [ctypes.c_void_p] * 6 + [ctypes.c_void_p]),
    "cot_solve_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.POINTER(cot_solve_opts), ctypes.POINTER(cot_report),
                                     ctypes.c_void_p]),
    "cot_centroids_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "cot_cell_solid_angles": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, ctypes.c_double, _dp]),
    "cot_irra_render": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_float, _dp,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p]),
    "cot_splat": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, _dp, ctypes.c_float, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]),
    "cot_irra_render_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
                                            ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int]
                             + [ctypes.c_void_p] * 7 + [ctypes.c_int] + [ctypes.c_void_p] * 3),
    "cot_irra_render_grad_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                    ctypes.c_float, ctypes.c_void_p, _dp,
                                                    ctypes.c_int, ctypes.c_int]
                                     + [ctypes.c_void_p] * 7 + [ctypes.c_int]
                                     + [ctypes.c_void_p] * 3),
    "cot_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "cot_reset_stats": (ctypes.c_int, [ctypes.c_void_p]),
}


class cot_stats(ctypes.Structure):
    _fields_ = [("cell_seconds", ctypes.c_double), ("cell_launches", ctypes.c_longlong),
                ("repair_seconds", ctypes.c_double), ("repair_launches", ctypes.c_longlong),
                ("verify_seconds", ctypes.c_double), ("verify_launches", ctypes.c_longlong),
                ("evaluations", ctypes.c_longlong), ("repair_rounds", ctypes.c_longlong),
                ("integ_seconds", ctypes.c_double), ("integ_launches", ctypes.c_longlong),
                ("newton_steps", ctypes.c_longlong), ("cg_iterations", ctypes.c_longlong),
                ("coarse_seconds", ctypes.c_double), ("cg_restarts", ctypes.c_longlong),
                ("handoff_degenerate", ctypes.c_longlong), ("handoff_overflow", ctypes.c_longlong),
                ("check_failures", ctypes.c_longlong),
                ("repair_check_failures", ctypes.c_longlong),
                ("repaired_cells", ctypes.c_longlong),
                ("reappeared_cells", ctypes.c_longlong)]

_lib = None


class CausticLibError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def load():
    """Load libcaustic_ot.so; raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m caustic_amd.build` "
                          "(the OT solver has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code, allow=()):
    if code != COT_OK and code not in allow:
        raise CausticLibError(code, load().cot_last_error().decode())
    return code
