"""ctypes binding of libhclpp_hip.so (the C-ABI in include/hclpp_hip.h).

The library must have been built in-tree (``python -m heraclespp_amd.build``);
there is no fallback of any kind: a missing or unloadable library raises.
HCLPP_LIB in the environment names another build of the same C-ABI to load
(libhclpp_hip_ieee.so: IEEE division in every kernel).
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HCLPP_LIB") or os.path.join(HERE, "libhclpp_hip.so")

HCLPP_OK = 0
HCLPP_ERR_ARG = 1
HCLPP_ERR_HIP = 2
HCLPP_ERR_STATE = 3
HCLPP_ERR_UNSUPPORTED = 4
HCLPP_ERR_NEGATIVE_DT = 5  # clock status: the next output time is not after the current time

GEOM = {"Cartesian": 0, "Spherical": 1}
LIMITER = {"Constant": 0, "VanLeer": 1, "Minmod": 2, "VanAlbada": 3}
RIEMANN = {"HLL": 0, "HLLC": 1, "Low Mach": 2}
GRAVITY = {"Uniform": 0, "Point_mass": 1, "Internal_mass": 2}
BC = {"Periodic": 0, "NullGradient": 1, "Reflexive": 2}
PRECISION = {"fp64": 0, "fp32": 1}

# every symbol declared in include/hclpp_hip.h
EXPORTS = [
    "hclpp_last_error", "hclpp_ctx_create", "hclpp_ctx_destroy", "hclpp_set_stream", "hclpp_range_with_ghosts",
    "hclpp_limiter_from_name", "hclpp_riemann_from_name", "hclpp_bc_from_name",
    "hclpp_conv_prim_to_cons", "hclpp_conv_cons_to_prim", "hclpp_face_reconstruction", "hclpp_hancock_extrapolation",
    "hclpp_face_prim_to_cons",
    "hclpp_hydro_reconstruction", "hclpp_godunov", "hclpp_time_step", "hclpp_min_internal_energy", "hclpp_integrate",
    "hclpp_fill_ghosts", "hclpp_clock_start", "hclpp_clock_read", "hclpp_step", "hclpp_step_fx",
    "hclpp_clock_start_f32", "hclpp_step_f32", "hclpp_step_fx_f32",
    "hclpp_cast_f64_to_f32", "hclpp_cast_f32_to_f64", "hclpp_step_dt", "hclpp_clock_record_dt", "hclpp_step_adjoint",
    "hclpp_steps_dt", "hclpp_steps_adjoint", "hclpp_pressure_fix", "hclpp_internal_mass_gravity", "hclpp_gravity_table",
    "hclpp_profile_begin", "hclpp_profile_end", "hclpp_fused_plan", "hclpp_time_step_noalloc",
    "hclpp_clock_set_outputs", "hclpp_clock_resume", "hclpp_clock_resume_f32", "hclpp_temperature",
    "hclpp_set_pressure_fix",
    "hclpp_set_fast_arithmetic",
]


class Config(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("geom", ctypes.c_int),
        ("limiter", ctypes.c_int),
        ("riemann", ctypes.c_int),
        ("gravity", ctypes.c_int),
        ("nfx", ctypes.c_int),
        ("ng", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("bc", ctypes.c_int * 6),
        ("bc_order", ctypes.c_int * 6),
        ("gamma", ctypes.c_double),
        ("mmw", ctypes.c_double),
        ("g", ctypes.c_double * 3),
        ("central_mass", ctypes.c_double),
        ("x", ctypes.POINTER(ctypes.c_double) * 3),
        ("precision", ctypes.c_int),
    ]


class Range(ctypes.Structure):
    _fields_ = [("begin", ctypes.c_int * 3), ("end", ctypes.c_int * 3)]


class Clock(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("t_end", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("min_eint", ctypes.c_double),
        ("mass", ctypes.c_double),
        ("momentum", ctypes.c_double * 3),
        ("energy", ctypes.c_double),
        ("iter", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("done", ctypes.c_int),
        ("status", ctypes.c_int),
        ("iter_output_id", ctypes.c_int),
        ("time_output_id", ctypes.c_int),
        ("output_due", ctypes.c_int),
        ("exiting", ctypes.c_int),
    ]


class HclppError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python -m heraclespp_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int
    d = ctypes.c_double
    rp = ctypes.POINTER(Range)
    L.hclpp_last_error.restype = ctypes.c_char_p
    L.hclpp_last_error.argtypes = []
    sig = {
        "hclpp_ctx_create": [ctypes.POINTER(Config), ctypes.POINTER(vp)],
        "hclpp_ctx_destroy": [vp],
        "hclpp_set_stream": [vp, vp],
        "hclpp_range_with_ghosts": [vp, ip, rp],
        "hclpp_conv_prim_to_cons": [vp, rp, dp, dp, dp, dp, dp],
        "hclpp_conv_cons_to_prim": [vp, rp, dp, dp, dp, dp, dp],
        "hclpp_face_reconstruction": [vp, ip, rp, dp, dp],
        "hclpp_hancock_extrapolation": [vp, rp, d, dp, dp, dp, dp, dp, dp],
        "hclpp_hydro_reconstruction": [vp, ip, rp, d] + [dp] * 10,
        "hclpp_face_prim_to_cons": [vp, rp] + [dp] * 5,
        "hclpp_godunov": [vp, ip, rp, d] + [dp] * 12,
        "hclpp_time_step": [vp, rp, dp, dp, dp, ctypes.POINTER(d)],
        "hclpp_time_step_noalloc": [vp, rp, dp, dp, dp, dp],
        "hclpp_min_internal_energy": [vp, rp, dp, dp, dp, ctypes.POINTER(d)],
        "hclpp_integrate": [vp, rp, dp, ctypes.POINTER(d)],
        "hclpp_fill_ghosts": [vp, dp, dp, dp, dp],
        "hclpp_clock_start": [vp, d, ip, d, d, ip, dp, dp, dp],
        "hclpp_clock_read": [vp, ctypes.POINTER(Clock)],
        "hclpp_clock_set_outputs": [vp, ip, d, d, ip, ip],
        "hclpp_set_pressure_fix": [vp, ip, d],
        "hclpp_set_fast_arithmetic": [vp, ip],
        "hclpp_clock_resume": [vp, dp, dp, dp],
        "hclpp_clock_resume_f32": [vp, dp, dp, dp],
        "hclpp_temperature": [vp, rp, dp, dp, dp],
        "hclpp_step": [vp, dp, dp, dp, dp, dp, dp],
        "hclpp_step_fx": [vp, dp, dp, dp, dp, dp, dp, dp, dp],
        "hclpp_clock_start_f32": [vp, d, ip, d, d, ip, dp, dp, dp],
        "hclpp_step_f32": [vp, dp, dp, dp, dp, dp, dp],
        "hclpp_step_fx_f32": [vp, dp, dp, dp, dp, dp, dp, dp, dp],
        "hclpp_fused_plan": [vp, ctypes.POINTER(ip), ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ip)],
        "hclpp_cast_f64_to_f32": [vp, dp, dp, ctypes.c_longlong],
        "hclpp_cast_f32_to_f64": [vp, dp, dp, ctypes.c_longlong],
        "hclpp_step_dt": [vp, ctypes.c_double, dp, dp, dp, dp, dp, dp],
        "hclpp_clock_record_dt": [vp, dp, ip],
        "hclpp_step_adjoint": [vp, ctypes.c_double] + [dp] * 10,
        "hclpp_steps_dt": [vp, ip, ctypes.POINTER(ctypes.c_double)] + [ctypes.POINTER(ctypes.c_void_p)] * 3,
        "hclpp_steps_adjoint": [vp, ip, ctypes.POINTER(ctypes.c_double)] + [ctypes.POINTER(ctypes.c_void_p)] * 6 + [dp],
        "hclpp_pressure_fix": [vp, rp, d, d] + [dp] * 9,
        "hclpp_internal_mass_gravity": [vp, dp],
        "hclpp_gravity_table": [vp, ctypes.POINTER(ctypes.c_double)],
        "hclpp_profile_begin": [vp, ip],
        "hclpp_profile_end": [vp, ctypes.POINTER(d), ctypes.POINTER(ctypes.c_int)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    for name in ("hclpp_limiter_from_name", "hclpp_riemann_from_name", "hclpp_bc_from_name"):
        f = getattr(L, name)
        f.argtypes = [ctypes.c_char_p]
        f.restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != HCLPP_OK:
        raise HclppError(rc, lib().hclpp_last_error().decode())
