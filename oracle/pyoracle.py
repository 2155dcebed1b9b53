"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


class OrDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("q", C.c_int),
        ("rheo", C.c_int), ("builder", C.c_int),
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
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        dp = P(C.c_double)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [P(OrDesc), C.c_int]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_build_map.argtypes = [P(OrDesc), P(C.c_uint32)]
        L.or_bc_count.restype = C.c_size_t
        L.or_bc_count.argtypes = [C.c_void_p]
        L.or_bc_list.argtypes = [C.c_void_p, P(C.c_uint64)]
        L.or_init.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_get_pops.argtypes = [C.c_void_p, dp]
        L.or_set_pops.argtypes = [C.c_void_p, dp]
        L.or_step.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.or_time_steps.restype = C.c_double
        L.or_time_steps.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.or_max_threads.restype = C.c_int
        L.or_get_macr.argtypes = [C.c_void_p, dp, dp, dp, dp, dp]
        L.or_set_omega.argtypes = [C.c_void_p, dp]
        L.or_collide_node.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp]
        L.or_apply_bc_raw.argtypes = [C.c_void_p, dp, dp]
        _lib = L
    return _lib


def desc_from(vdesc) -> OrDesc:
    """Oracle descriptor with the same simulation as a product vlbm_desc."""
    o = OrDesc()
    for name, _ in OrDesc._fields_:
        if name == "rheo":
            o.rheo = vdesc.rheology
        elif name in ("ux_bc", "uy_bc", "uz_bc", "rho_bc"):
            src, dst = getattr(vdesc, name), getattr(o, name)
            for k in range(8):
                dst[k] = src[k]
        else:
            setattr(o, name, getattr(vdesc, name))
    return o


def _p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def build_map(odesc: OrDesc) -> np.ndarray:
    out = np.zeros((odesc.nz, odesc.ny, odesc.nx), dtype=np.uint32)
    lib().or_build_map(C.byref(odesc), out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


class Oracle:
    """Reference-layout lattice stepped on the CPU (FP64 by default)."""

    def __init__(self, odesc: OrDesc, prec: int = 8):
        self.desc = odesc
        self.h = lib().or_create(C.byref(odesc), prec)
        if not self.h:
            raise ValueError("invalid oracle descriptor")
        self.shape = (odesc.nz, odesc.ny, odesc.nx)
        self.q = odesc.q

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self, rho=None, ux=None, uy=None, uz=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (rho, ux, uy, uz)]
        lib().or_init(self.h, *[_p(a) for a in arrs])

    def step(self, n=1, save_last=False, threads=0):
        lib().or_step(self.h, int(n), 1 if save_last else 0, int(threads))

    def time_steps(self, n, threads=0) -> float:
        return lib().or_time_steps(self.h, int(n), int(threads))

    def get_pops(self) -> np.ndarray:
        nz, ny, nx = self.shape
        out = np.zeros((self.q, nz + 1, ny, nx))
        lib().or_get_pops(self.h, _p(out))
        return out

    def set_pops(self, pop):
        pop = np.ascontiguousarray(pop, dtype=np.float64)
        lib().or_set_pops(self.h, _p(pop))

    def set_omega(self, omega):
        lib().or_set_omega(self.h, _p(np.ascontiguousarray(omega, dtype=np.float64)))

    def macroscopics(self) -> dict:
        out = {k: np.zeros(self.shape) for k in ("rho", "ux", "uy", "uz", "omega")}
        lib().or_get_macr(self.h, *[_p(out[k]) for k in ("rho", "ux", "uy", "uz", "omega")])
        return out

    def bc_list(self) -> np.ndarray:
        n = lib().or_bc_count(self.h)
        out = np.zeros(n, dtype=np.uint64)
        if n:
            lib().or_bc_list(self.h, out.ctypes.data_as(C.POINTER(C.c_uint64)))
        return out

    def collide_node(self, f, fx=0.0, fy=0.0, fz=0.0, omega_old=0.0):
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros(self.q)
        m = np.zeros(5)
        lib().or_collide_node(self.h, _p(f), fx, fy, fz, omega_old, _p(out), _p(m))
        return out, m

    def apply_bc_raw(self, post_stream, post_col):
        ps = np.ascontiguousarray(post_stream, dtype=np.float64).copy()
        pc = np.ascontiguousarray(post_col, dtype=np.float64)
        lib().or_apply_bc_raw(self.h, _p(ps), _p(pc))
        return ps


def max_threads() -> int:
    return lib().or_max_threads()
