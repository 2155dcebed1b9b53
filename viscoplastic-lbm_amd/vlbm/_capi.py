"""ctypes binding of libvlbm.so (include/vlbm.h).

The library is the only compute path: importing this module fails loudly when
the built shared object is missing, and nothing falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("VLBM_LIB", _PKG / "lib" / "libvlbm.so"))

VLBM_OK = 0
ERRORS = {1: "invalid argument", 2: "device error", 3: "allocation failed", 4: "I/O error", 5: "unsupported"}

FP64, FP32, FP16, BF16 = 8, 4, 2, 3
ROUND_STOCHASTIC, ROUND_NEAREST = 0, 1
LAYOUT_AUTO, LAYOUT_POPULATIONS, LAYOUT_MOMENTS = 0, 1, 2
NEWTONIAN, BINGHAM, POWERLAW = 0, 1, 2


class VlbmDesc(C.Structure):
    """Mirror of vlbm_desc (include/vlbm.h)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("q", C.c_int), ("precision", C.c_int), ("rheology", C.c_int), ("builder", C.c_int),
        ("tau", C.c_double), ("rho0", C.c_double),
        ("fx", C.c_double), ("fy", C.c_double), ("fz", C.c_double),
        ("eta_p", C.c_double), ("s_y", C.c_double),
        ("n_index", C.c_double), ("k_consistency", C.c_double), ("gamma_0", C.c_double),
        ("omega_clamp", C.c_int),
        ("tau_min", C.c_double), ("tau_max", C.c_double),
        ("ux_bc", C.c_double * 8), ("uy_bc", C.c_double * 8),
        ("uz_bc", C.c_double * 8), ("rho_bc", C.c_double * 8),
        ("duct_outer_radius", C.c_double), ("duct_inner_radius", C.c_double),
        ("outer_rotation", C.c_double), ("outer_velocity", C.c_double),
        ("inner_rotation", C.c_double), ("inner_velocity", C.c_double),
        ("zouhe_textbook", C.c_int),
        ("collision", C.c_int),
        ("les_const", C.c_double),
        ("layout", C.c_int),
        ("rounding", C.c_int),
        ("les_visc_turb", C.c_int),
    ]


class VlbmIbmDesc(C.Structure):
    """Mirror of vlbm_ibm_desc (include/vlbm.h)."""

    _fields_ = [
        ("max_iterations", C.c_int),
        ("periodic_x", C.c_int), ("periodic_y", C.c_int), ("periodic_z", C.c_int),
        ("thickness", C.c_double), ("alpha", C.c_double), ("fluid_density", C.c_double),
        ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double),
        ("pp_friction", C.c_double), ("pw_friction", C.c_double),
        ("pp_rest", C.c_double), ("pw_rest", C.c_double),
        ("young", C.c_double), ("poisson", C.c_double),
        ("wall_young", C.c_double), ("wall_poisson", C.c_double),
        ("ref_float_helpers", C.c_int),
        ("mesh_scale", C.c_double), ("breugem", C.c_double), ("coulomb", C.c_int),
        ("max_spheres", C.c_int), ("max_nodes", C.c_int),
    ]


MESH_POLAR, MESH_FIBONACCI = 0, 1
IBM_RECORD = 48
IBM_HISTORY = 13


class VlbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vlbm error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"libvlbm.so not found at {LIB_PATH}; build it with "
            "`python viscoplastic-lbm_amd/build.py` (there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    dp = P(C.c_double)
    sig = {
        "vlbm_desc_defaults": (None, [P(VlbmDesc)]),
        "vlbm_create": (C.c_int, [P(VlbmDesc), P(C.c_void_p)]),
        "vlbm_destroy": (None, [C.c_void_p]),
        "vlbm_last_error": (C.c_char_p, []),
        "vlbm_build_map": (C.c_int, [P(VlbmDesc), P(C.c_uint32)]),
        "vlbm_bc_nodes": (C.c_int, [C.c_void_p, P(C.c_uint64), C.c_size_t, P(C.c_size_t)]),
        "vlbm_init": (C.c_int, [C.c_void_p, dp, dp, dp, dp]),
        "vlbm_load_pops": (C.c_int, [C.c_void_p, dp]),
        "vlbm_get_pops": (C.c_int, [C.c_void_p, dp]),
        "vlbm_get_pops_noalloc": (C.c_int, [C.c_void_p, dp]),
        "vlbm_set_map": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
        "vlbm_visc_turb": (C.c_int, [C.c_void_p, dp]),
        "vlbm_set_omega": (C.c_int, [C.c_void_p, dp]),
        "vlbm_step": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "vlbm_compute_macroscopics": (C.c_int, [C.c_void_p]),
        "vlbm_macroscopics": (C.c_int, [C.c_void_p, dp, dp, dp, dp, dp]),
        "vlbm_reduce": (C.c_int, [C.c_void_p, dp, dp]),
        "vlbm_save_macr_bin": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_uint]),
        "vlbm_save_info": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_uint, C.c_double, C.c_double]),
        "vlbm_steps_done": (C.c_int, [C.c_void_p, P(C.c_longlong)]),
        "vlbm_save_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]),
        "vlbm_load_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, P(C.c_int)]),
        "vlbm_time_steps": (C.c_int, [C.c_void_p, C.c_int, dp]),
        "vlbm_sync": (C.c_int, [C.c_void_p]),
        "vlbm_device_bytes": (C.c_int, [C.c_void_p, P(C.c_size_t)]),
        "vlbm_layout": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "vlbm_copy_probe": (C.c_int, [C.c_size_t, C.c_int, dp]),
        "vlbm_ibm_desc_defaults": (None, [P(VlbmIbmDesc)]),
        "vlbm_ibm_enable": (C.c_int, [C.c_void_p, P(VlbmIbmDesc)]),
        "vlbm_ibm_add_sphere": (C.c_int, [C.c_void_p, dp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                          P(C.c_int)]),
        "vlbm_ibm_add_particle": (C.c_int, [C.c_void_p, dp, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp,
                                            P(C.c_int)]),
        "vlbm_ibm_counts": (C.c_int, [C.c_void_p, P(C.c_int), P(C.c_int)]),
        "vlbm_ibm_get_particles": (C.c_int, [C.c_void_p, dp]),
        "vlbm_ibm_set_particles": (C.c_int, [C.c_void_p, dp]),
        "vlbm_ibm_get_collisions": (C.c_int, [C.c_void_p, P(C.c_int), dp, P(C.c_int)]),
        "vlbm_ibm_get_nodes": (C.c_int, [C.c_void_p, dp, dp, dp]),
        "vlbm_ibm_get_force": (C.c_int, [C.c_void_p, dp]),
        "vlbm_ibm_cells": (C.c_int, [C.c_void_p, P(C.c_int), P(C.c_int)]),
        "vlbm_ibm_stage": (C.c_int, [C.c_void_p]),
        "vlbm_ibm_status": (C.c_int, [C.c_void_p]),
        "vlbm_ibm_time_steps": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(code: int) -> None:
    if code != VLBM_OK:
        raise VlbmError(code, lib.vlbm_last_error().decode())


def default_desc() -> VlbmDesc:
    d = VlbmDesc()
    lib.vlbm_desc_defaults(C.byref(d))
    return d


def default_ibm_desc(**kw) -> VlbmIbmDesc:
    d = VlbmIbmDesc()
    lib.vlbm_ibm_desc_defaults(C.byref(d))
    for k, v in kw.items():
        if not hasattr(d, k):
            raise KeyError(f"unknown IBM descriptor field {k!r}")
        setattr(d, k, v)
    return d
