"""ctypes binding of libcup.so (the C-ABI declared in include/cup.h).

The library is built in-tree by ``cubismup_3d_amd.build.build()`` (or
``__graft_entry__.build()``) and is the only compute path: there is no CPU fallback, and
``lib()`` raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcup.so")

CUP_OK = 0
CUP_E_INVALID = -1
CUP_E_DEVICE = -2
CUP_E_STATE = -3
CUP_E_NOMEM = -4
CUP_E_IO = -5
CUP_NOALLOC_MAX_PARTS = 32768

CUP_CHI, CUP_U, CUP_V, CUP_W, CUP_P, CUP_TMPU, CUP_TMPV, CUP_TMPW = range(8)
FIELD_IDS = {"chi": 0, "u": 1, "v": 2, "w": 3, "p": 4, "tmpU": 5, "tmpV": 6, "tmpW": 7,
             "adjU": 8, "adjV": 9, "adjW": 10}
CUP_PLAIN, CUP_FUSED = 0, 1
CUP_BC_PERIODIC, CUP_BC_OPEN, CUP_BC_WALL, CUP_BC_FREESPACE = 0, 1, 2, 3
BC_IDS = {"periodic": 0, "dirichlet": 1, "open": 1, "absorbing": 1, "wall": 2, "freespace": 3, "unbounded": 3}
CUP_CENTRAL_EULER, CUP_UPWIND3_EULER, CUP_UPWIND3_RK2 = 0, 1, 2


class GridDesc(C.Structure):
    _fields_ = [("bpd", C.c_int * 3), ("maxextent", C.c_double), ("precision", C.c_int),
                ("decomposition", C.c_int), ("device", C.c_int), ("bc", C.c_int * 3)]


class StepParams(C.Structure):
    _fields_ = [("CFL", C.c_double), ("nu", C.c_double), ("uinf", C.c_double * 3),
                ("rampup", C.c_int), ("scheme", C.c_int), ("spectral_forcing", C.c_int),
                ("tke_target", C.c_double), ("injection_rate", C.c_double),
                ("fixed_dt", C.c_double), ("has_body", C.c_int), ("lambda_", C.c_double),
                ("implicit_penalization", C.c_int), ("body_vel", C.c_double * 3),
                ("body_angvel", C.c_double * 3)]


# Every symbol include/cup.h declares, with its ctypes signature.
_d, _i, _p, _u64 = C.c_double, C.c_int, C.c_void_p, C.c_uint64
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
SIGNATURES = {
    "cup_create": (_i, [C.POINTER(GridDesc), C.POINTER(C.c_void_p)]),
    "cup_destroy": (_i, [_p]),
    "cup_last_error": (C.c_char_p, [_p]),
    "cup_get_geometry": (_i, [_p, _ip, _dp, _dp]),
    "cup_synchronize": (_i, [_p]),
    "cup_set_field": (_i, [_p, _i, _dp]),
    "cup_get_field": (_i, [_p, _i, _dp]),
    "cup_field_device_ptr": (_i, [_p, _i, C.POINTER(C.c_void_p)]),
    "cup_set_field_device": (_i, [_p, _i, _p]),
    "cup_get_field_device": (_i, [_p, _i, _p]),
    "cup_lab_shift": (_i, [_p, _i, _i, _i, _i]),
    "cup_max_velocity": (_i, [_p, _dp, _dp]),
    "cup_calc_max_timestep": (_i, [_p, _d, _d, _dp, _i, _i, _dp, _dp]),
    "cup_advect_diffuse": (_i, [_p, _d, _d, _dp, _i]),
    "cup_pressure_rhs": (_i, [_p, _d]),
    "cup_poisson_periodic": (_i, [_p]),
    "cup_poisson_mixed": (_i, [_p]),
    "cup_poisson_unbounded": (_i, [_p]),
    "cup_pressure_project": (_i, [_p, _d]),
    "cup_chi_from_sphere": (_i, [_p, _dp, _d, _dp]),
    "cup_chi_from_hemisphere": (_i, [_p, _dp, _d, _dp]),
    "cup_hit_analysis": (_i, [_p, _d, _dp, _dp, _i]),
    "cup_chi_from_sphere_mollified": (_i, [_p, _dp, _d, _d, _dp]),
    "cup_penalize": (_i, [_p, _d, _d, _i, _dp, _dp, _dp]),
    "cup_surface_forces": (_i, [_p, _d, _dp, _dp, _dp]),
    "cup_obstacle_momenta": (_i, [_p, _d, _d, _i, _dp]),
    "cup_body_count": (_i, [_p, C.POINTER(C.c_int)]),
    "cup_add_sphere": (_i, [_p, _dp, _d, _i, _dp, C.POINTER(C.c_int)]),
    "cup_move_body": (_i, [_p, _i, _dp, _dp]),
    "cup_body_velocity": (_i, [_p, _i, _dp, _dp]),
    "cup_last_body_force_body": (_i, [_p, _i, _dp]),
    "cup_body_obstacle_blocks": (_i, [_p, _i, C.POINTER(C.c_int), _i, C.POINTER(C.c_int)]),
    "cup_body_set_udef": (_i, [_p, _i, _dp, _i]),
    "cup_body_get_udef": (_i, [_p, _i, _dp, _i]),
    "cup_body_remove_udef_momenta": (_i, [_p, _i, _dp]),
    "cup_penalize_body": (_i, [_p, _i, _d, _d, _i, _dp, _dp, _dp]),
    "cup_surface_forces_body": (_i, [_p, _i, _d, _dp, _dp, _dp]),
    "cup_obstacle_momenta_body": (_i, [_p, _i, _d, _d, _i, _dp]),
    "cup_add_uniform_velocity": (_i, [_p, _dp]),
    "cup_ic_hit_spectrum": (_i, [_p, _dp, _dp, _i, _d, _u64]),
    "cup_diagnostics": (_i, [_p, _d, _dp]),
    "cup_diagnostics_noalloc": (_i, [C.POINTER(GridDesc), C.POINTER(C.c_void_p), _d, _p, _p]),
    "cup_check_nan": (_i, [_p, C.POINTER(C.c_longlong)]),
    "cup_ic_taylor_green": (_i, [_p, _d]),
    "cup_ic_hit": (_i, [_p, _d, _d, _u64]),
    "cup_spectral_forcing": (_i, [_p, _d, _d, _d, _d, _d, _dp]),
    "cup_run_steps": (_i, [_p, C.POINTER(StepParams), _i]),
    "cup_step_state": (_i, [_p, _ip, _dp, _dp, _dp, _dp]),
    "cup_set_step_state": (_i, [_p, _i, _d]),
    "cup_adjoint_step": (_i, [_p, _d, _d, _dp, _i]),
    "cup_adjoint_step_forcing": (_i, [_p, _d, _d, _dp, _i, _d, _d]),
    "cup_adjoint_step_body": (_i, [_p, _d, _d, _dp, _d, _i, _dp, _d, _dp, _dp]),
    "cup_adjoint_step_body_full": (_i, [_p, _d, _d, _dp, _i, _d, _d, _d, _i, _dp, _dp, _d, _dp, _dp, _dp]),
    "cup_adjoint_step_bodies": (_i, [_p, _d, _d, _dp, _i, _d, _d, _d, _i, _i, _dp, _dp, _dp, _dp, _dp]),
    "cup_last_body_force": (_i, [_p, _dp]),
    "cup_stream": (_i, [_p, C.POINTER(C.c_void_p)]),
    "cup_stage_timing_begin": (_i, [_p, _i, _i]),
    "cup_stage_timing_end": (_i, [_p, _dp, _ip]),
    "cup_sim_create": (_i, [_i, C.POINTER(C.c_char_p), C.POINTER(C.c_void_p)]),
    "cup_sim_set_cells": (_i, [_p, _i, _i, _i]),
    "cup_sim_init": (_i, [_p]),
    "cup_sim_destroy": (_i, [_p]),
    "cup_sim_last_error": (C.c_char_p, [_p]),
    "cup_sim_add_sphere": (_i, [_p, _d, _dp, _dp, _i, _i, _i, _d, _d, _i]),
    "cup_sim_run": (_i, [_p]),
    "cup_sim_calc_max_timestep": (_i, [_p, _dp]),
    "cup_sim_timestep": (_i, [_p, _d, _ip]),
    "cup_sim_advance": (_i, [_p, _i, _ip]),
    "cup_sim_context": (_i, [_p, C.POINTER(C.c_void_p)]),
    "cup_sim_get": (_i, [_p, C.c_char_p, _dp]),
    "cup_sim_set": (_i, [_p, C.c_char_p, _d]),
    "cup_sim_obstacle": (_i, [_p, _i, _dp]),
    "cup_sim_obstacle_state": (_i, [_p, _i, _dp]),
    "cup_sim_diagnostics": (_i, [_p, _dp]),
    "cup_sim_flush_logs": (_i, [_p]),
}

_lib = None


class CupError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cup error {code}: {msg}")
        self.code = code


def lib():
    """Load libcup.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                          "(make -C cubismup_3d_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, ctx=None, sim=None):
    if rc == CUP_OK:
        return
    L = lib()
    if ctx is not None:
        msg = L.cup_last_error(ctx)
    elif sim is not None:
        msg = L.cup_sim_last_error(sim)
    else:
        msg = b""
    msg = msg.decode() if msg else ""
    if rc == CUP_E_INVALID:
        raise ValueError(msg or "invalid argument")
    if rc == CUP_E_IO:
        raise OSError(msg or "file error")
    raise CupError(rc, msg)


def dptr(a):
    return a.ctypes.data_as(_dp)


def vec3(x):
    return (C.c_double * 3)(*[float(v) for v in x])
