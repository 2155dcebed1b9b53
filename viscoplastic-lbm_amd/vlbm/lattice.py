"""One lattice resident on the GPU, driven through the C-ABI (include/vlbm.h).

Mirrors the reference's per-simulation objects (Populations, Macroscopics,
NodeTypeMap; src/CUDA/structs/) and the operations main.cu applies to them.
Arrays cross the boundary as numpy float64, shaped [NZ][NY][NX] for scalar
fields and [Q][NZ+1][NY][NX] for populations (the reference's idxPop layout).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _capi
from ._capi import check, lib

_dp = C.POINTER(C.c_double)


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a) -> Optional[np.ndarray]:
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.float64)


def build_map(desc: _capi.VlbmDesc) -> np.ndarray:
    """Node-type words [NZ][NY][NX] (host only; gpuBuildBoundaryConditions)."""
    out = np.zeros((desc.nz, desc.ny, desc.nx), dtype=np.uint32)
    check(lib.vlbm_build_map(C.byref(desc), out.ctypes.data_as(C.POINTER(C.c_uint32))))
    return out


class Lattice:
    def __init__(self, desc: _capi.VlbmDesc):
        self.desc = desc
        h = C.c_void_p()
        check(lib.vlbm_create(C.byref(desc), C.byref(h)))
        self._h = h
        self.shape = (desc.nz, desc.ny, desc.nx)
        self.q = desc.q

    @classmethod
    def from_config(cls, cfg) -> "Lattice":
        lat = cls(cfg.to_desc())
        if cfg.use_ibm:
            lat.ibm_enable(cfg.to_ibm_desc())
            for sp in cfg.spheres or []:
                lat.ibm_add_sphere(sp["center"], vel=sp.get("vel", (0, 0, 0)), w=sp.get("w", (0, 0, 0)),
                                   radius=sp["radius"], density=sp.get("density", 2.0),
                                   movable=sp.get("movable", True),
                                   mesh=_capi.MESH_FIBONACCI if sp.get("mesh") == "fibonacci" else _capi.MESH_POLAR,
                                   n_fibonacci=int(sp.get("nodes", 0)))
        return lat

    # -- lifetime -------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.vlbm_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ----------------------------------------------------------------
    def init(self, rho=None, ux=None, uy=None, uz=None) -> None:
        arrs = [_f64(a) for a in (rho, ux, uy, uz)]
        for a in arrs:
            if a is not None and a.size != np.prod(self.shape):
                raise ValueError("initial fields must have NX*NY*NZ entries")
        check(lib.vlbm_init(self._h, *[_ptr(a) for a in arrs]))

    def load_pops(self, pop: np.ndarray) -> None:
        pop = _f64(pop)
        if pop.size != self.q * (self.shape[0] + 1) * self.shape[1] * self.shape[2]:
            raise ValueError("populations must have Q*(NZ+1)*NY*NX entries")
        check(lib.vlbm_load_pops(self._h, _ptr(pop)))

    def get_pops(self) -> np.ndarray:
        nz, ny, nx = self.shape
        out = np.zeros((self.q, nz + 1, ny, nx), dtype=np.float64)
        check(lib.vlbm_get_pops(self._h, _ptr(out)))
        return out

    def visc_turb(self) -> np.ndarray:
        """LES eddy viscosity tau_t / 3 of the last step ([NZ][NY][NX])."""
        out = np.zeros(self.shape, dtype=np.float64)
        check(lib.vlbm_visc_turb(self._h, _ptr(out)))
        return out

    def set_map(self, words: np.ndarray) -> None:
        """Replaces the node map with words written outside the library
        ([NZ][NY][NX] uint32 in the layout of build_map)."""
        w = np.ascontiguousarray(words, dtype=np.uint32)
        if w.size != self.shape[0] * self.shape[1] * self.shape[2]:
            raise ValueError("the node map must have NX*NY*NZ words")
        check(lib.vlbm_set_map(self._h, w.ctypes.data_as(C.POINTER(C.c_uint32))))

    def set_omega(self, omega: np.ndarray) -> None:
        check(lib.vlbm_set_omega(self._h, _ptr(_f64(omega))))

    # -- stepping -------------------------------------------------------------
    def step(self, n: int = 1, save_last: bool = False) -> None:
        check(lib.vlbm_step(self._h, int(n), 1 if save_last else 0))

    def time_steps(self, n: int) -> float:
        ms = C.c_double()
        check(lib.vlbm_time_steps(self._h, int(n), C.byref(ms)))
        return ms.value

    def sync(self) -> None:
        check(lib.vlbm_sync(self._h))

    @property
    def steps_done(self) -> int:
        s = C.c_longlong()
        check(lib.vlbm_steps_done(self._h, C.byref(s)))
        return s.value

    @property
    def device_bytes(self) -> int:
        b = C.c_size_t()
        check(lib.vlbm_device_bytes(self._h, C.byref(b)))
        return b.value

    @property
    def layout(self) -> str:
        """State representation in use: "populations" or "moments"."""
        v = C.c_int()
        check(lib.vlbm_layout(self._h, C.byref(v)))
        return "moments" if v.value == _capi.LAYOUT_MOMENTS else "populations"

    # -- macroscopics -----------------------------------------------------------
    def compute_macroscopics(self) -> None:
        check(lib.vlbm_compute_macroscopics(self._h))

    def macroscopics(self, omega: bool = False) -> dict:
        out = {k: np.zeros(self.shape) for k in ("rho", "ux", "uy", "uz")}
        om = np.zeros(self.shape) if omega else None
        check(lib.vlbm_macroscopics(self._h, _ptr(out["rho"]), _ptr(out["ux"]), _ptr(out["uy"]),
                                    _ptr(out["uz"]), _ptr(om)))
        if omega:
            out["omega"] = om
        return out

    def reduce(self) -> Tuple[float, float]:
        res, mr = C.c_double(), C.c_double()
        check(lib.vlbm_reduce(self._h, C.byref(res), C.byref(mr)))
        return res.value, mr.value

    def bc_nodes(self) -> np.ndarray:
        cnt = C.c_size_t()
        check(lib.vlbm_bc_nodes(self._h, None, 0, C.byref(cnt)))
        out = np.zeros(cnt.value, dtype=np.uint64)
        if cnt.value:
            check(lib.vlbm_bc_nodes(self._h, out.ctypes.data_as(C.POINTER(C.c_uint64)), cnt.value, C.byref(cnt)))
        return out

    # -- output -----------------------------------------------------------------
    def save_macr_bin(self, path: str, sim_id: str, step: int) -> None:
        check(lib.vlbm_save_macr_bin(self._h, path.encode(), sim_id.encode(), int(step)))

    def save_checkpoint(self, path: str, sim_id: str, step: Optional[int] = None) -> None:
        """PATH/ID/checkpoint/ in the reference's layout; step defaults to the
        steps done, which makes a loaded run continue exactly."""
        s = self.steps_done if step is None else int(step)
        check(lib.vlbm_save_checkpoint(self._h, path.encode(), sim_id.encode(), s))

    def load_checkpoint(self, path: str, sim_id: str) -> int:
        """Loads PATH/ID/checkpoint/; returns the stored step."""
        s = C.c_int(0)
        check(lib.vlbm_load_checkpoint(self._h, path.encode(), sim_id.encode(), C.byref(s)))
        return s.value

    def save_info(self, path: str, sim_id: str, total_steps: int, mlups: float, seconds: float) -> None:
        check(lib.vlbm_save_info(self._h, path.encode(), sim_id.encode(), int(total_steps),
                                 float(mlups), float(seconds)))

    # -- immersed boundary stage (src/CUDA/IBM/) --------------------------------
    def ibm_enable(self, idesc: "_capi.VlbmIbmDesc") -> None:
        check(lib.vlbm_ibm_enable(self._h, C.byref(idesc)))

    def ibm_add_sphere(self, pos, vel=(0, 0, 0), w=(0, 0, 0), radius=1.0, density=2.0, movable=True,
                       mesh=_capi.MESH_POLAR, n_fibonacci=0) -> int:
        state = _f64(np.concatenate([pos, vel, w]))
        sid = C.c_int(-1)
        check(lib.vlbm_ibm_add_sphere(self._h, _ptr(state), float(radius), float(density), int(bool(movable)),
                                      int(mesh), int(n_fibonacci), C.byref(sid)))
        return sid.value

    def ibm_add_particle(self, pos, vel, w, radius, density, movable, xyz, area) -> int:
        state = _f64(np.concatenate([pos, vel, w]))
        xyz = _f64(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        area = _f64(area)
        sid = C.c_int(-1)
        check(lib.vlbm_ibm_add_particle(self._h, _ptr(state), float(radius), float(density), int(bool(movable)),
                                        len(area), _ptr(xyz), _ptr(area), C.byref(sid)))
        return sid.value

    def ibm_counts(self) -> Tuple[int, int]:
        ns, nn = C.c_int(0), C.c_int(0)
        check(lib.vlbm_ibm_counts(self._h, C.byref(ns), C.byref(nn)))
        return ns.value, nn.value

    def ibm_particles(self) -> np.ndarray:
        ns, _ = self.ibm_counts()
        out = np.zeros((ns, _capi.IBM_RECORD))
        if ns:
            check(lib.vlbm_ibm_get_particles(self._h, _ptr(out)))
        return out

    def ibm_set_particles(self, rec: np.ndarray) -> None:
        check(lib.vlbm_ibm_set_particles(self._h, _ptr(_f64(rec))))

    def ibm_collision_state(self):
        ns, _ = self.ibm_counts()
        partner = np.zeros((ns, _capi.IBM_HISTORY), dtype=np.int32)
        last = np.zeros((ns, _capi.IBM_HISTORY), dtype=np.int32)
        tang = np.zeros((ns, _capi.IBM_HISTORY, 3))
        ip = C.POINTER(C.c_int)
        check(lib.vlbm_ibm_get_collisions(self._h, partner.ctypes.data_as(ip), _ptr(tang), last.ctypes.data_as(ip)))
        return partner, tang, last

    def ibm_nodes(self):
        _, nn = self.ibm_counts()
        pos, f, df = np.zeros((3, nn)), np.zeros((3, nn)), np.zeros((3, nn))
        check(lib.vlbm_ibm_get_nodes(self._h, _ptr(pos), _ptr(f), _ptr(df)))
        return pos, f, df

    def ibm_force(self) -> np.ndarray:
        out = np.zeros((3,) + self.shape)
        check(lib.vlbm_ibm_get_force(self._h, _ptr(out)))
        return out

    def ibm_cells(self):
        ns, _ = self.ibm_counts()
        dims = np.zeros(3, dtype=np.int32)
        cells = np.zeros(max(ns, 1), dtype=np.int32)
        ip = C.POINTER(C.c_int)
        check(lib.vlbm_ibm_cells(self._h, dims.ctypes.data_as(ip), cells.ctypes.data_as(ip)))
        return dims, cells[:ns]

    def ibm_stage(self) -> None:
        check(lib.vlbm_ibm_stage(self._h))

    def ibm_status(self) -> None:
        check(lib.vlbm_ibm_status(self._h))

    def ibm_time_steps(self, n: int) -> Tuple[float, float]:
        a, b = C.c_double(0), C.c_double(0)
        check(lib.vlbm_ibm_time_steps(self._h, int(n), C.byref(a), C.byref(b)))
        return a.value, b.value
