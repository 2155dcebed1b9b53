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
        ("collision", C.c_int), ("les_const", C.c_double), ("les_visc_turb", C.c_int),
    ]


class OrIbmDesc(C.Structure):
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
    ]


# record of one particle returned by Oracle.ibm_particles(): name -> slice
PARTICLE_FIELDS = {}
_off = 0
for _name, _n in (("pos", 3), ("vel", 3), ("w", 3), ("f", 3), ("M", 3), ("I", 6), ("pos_old", 3),
                  ("vel_old", 3), ("w_old", 3), ("w_avg", 3), ("w_pos", 3), ("f_old", 3), ("M_old", 3),
                  ("dP", 3), ("dL", 3)):
    PARTICLE_FIELDS[_name] = slice(_off, _off + _n)
    _off += _n
PARTICLE_RECORD = _off  # 48
MAX_COLLISIONS = 13


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
        L.or_set_pops2.argtypes = [C.c_void_p, dp, dp]
        L.or_step.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.or_time_steps.restype = C.c_double
        L.or_time_steps.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.or_max_threads.restype = C.c_int
        L.or_get_macr.argtypes = [C.c_void_p, dp, dp, dp, dp, dp]
        L.or_set_omega.argtypes = [C.c_void_p, dp]
        L.or_get_visc_turb.argtypes = [C.c_void_p, dp]
        L.or_collide_node.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp]
        L.or_apply_bc_raw.argtypes = [C.c_void_p, dp, dp]
        L.or_set_map.argtypes = [C.c_void_p, P(C.c_uint32)]
        L.or_ibb_q.restype = C.c_double
        L.or_ibb_q.argtypes = [C.c_int] + [C.c_double] * 5
        ip = P(C.c_int)
        L.or_sphere_polar.restype = C.c_int
        L.or_sphere_polar.argtypes = [C.c_double, dp, C.c_double, C.c_double, C.c_int, dp, dp, C.c_int]
        L.or_ibm_enable.restype = C.c_int
        L.or_ibm_enable.argtypes = [C.c_void_p, P(OrIbmDesc)]
        L.or_ibm_add_particle.restype = C.c_int
        L.or_ibm_add_particle.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp]
        L.or_ibm_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.or_ibm_stage.argtypes = [C.c_void_p, C.c_int]
        L.or_ibm_collisions.argtypes = [C.c_void_p, C.c_int]
        L.or_ibm_update_velocity.argtypes = [C.c_void_p]
        L.or_ibm_count.restype = C.c_int
        L.or_ibm_count.argtypes = [C.c_void_p, ip]
        L.or_ibm_get_particles.argtypes = [C.c_void_p, dp]
        L.or_ibm_get_axes.argtypes = [C.c_void_p, dp]
        L.or_ibm_move.argtypes = [C.c_void_p]
        L.or_ibm_set_particles.argtypes = [C.c_void_p, dp]
        L.or_ibm_get_collisions.argtypes = [C.c_void_p, ip, dp, ip]
        L.or_ibm_get_nodes.argtypes = [C.c_void_p, dp, dp, dp]
        L.or_ibm_get_force.argtypes = [C.c_void_p, dp]
        L.or_ibm_set_force.argtypes = [C.c_void_p, dp]
        L.or_ibm_interp_spread.argtypes = [C.c_void_p, dp]
        L.or_ibm_update_macr.argtypes = [C.c_void_p]
        L.or_ibm_set_macr.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_ibm_cells.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, dp, C.c_int, ip, ip]
        L.or_stencil4.restype = C.c_double
        L.or_stencil4.argtypes = [C.c_double]
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

    def set_pops(self, pop, pop_aux=None):
        """Both buffers take pop, or pop and pop_aux (a checkpoint load)."""
        pop = np.ascontiguousarray(pop, dtype=np.float64)
        if pop_aux is None:
            lib().or_set_pops(self.h, _p(pop))
        else:
            aux = np.ascontiguousarray(pop_aux, dtype=np.float64)
            assert aux.shape == pop.shape
            lib().or_set_pops2(self.h, _p(pop), _p(aux))

    def visc_turb(self) -> np.ndarray:
        out = np.zeros(self.shape)
        lib().or_get_visc_turb(self.h, _p(out))
        return out

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

    # ---- immersed boundary stage (FP64 lattices) ----
    def ibm_enable(self, idesc: "OrIbmDesc"):
        if lib().or_ibm_enable(self.h, C.byref(idesc)) != 0:
            raise ValueError("IBM needs an FP64 oracle lattice")

    def ibm_add_particle(self, pos, vel, w, radius, density, movable, xyz, S) -> int:
        state = np.ascontiguousarray(np.concatenate([pos, vel, w]), dtype=np.float64)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        S = np.ascontiguousarray(S, dtype=np.float64)
        return lib().or_ibm_add_particle(self.h, _p(state), float(radius), float(density), int(bool(movable)),
                                         len(S), _p(xyz), _p(S))

    def ibm_run(self, n, first_step, threads=0):
        lib().or_ibm_run(self.h, int(n), int(first_step), int(threads))

    def ibm_stage(self, step):
        lib().or_ibm_stage(self.h, int(step))

    def ibm_collisions(self, step):
        lib().or_ibm_collisions(self.h, int(step))

    def ibm_move(self):
        lib().or_ibm_move(self.h)

    def ibm_update_velocity(self):
        lib().or_ibm_update_velocity(self.h)

    def ibm_counts(self):
        n = C.c_int(0)
        np_ = lib().or_ibm_count(self.h, C.byref(n))
        return np_, n.value

    def ibm_particles(self) -> np.ndarray:
        np_, _ = self.ibm_counts()
        out = np.zeros((np_, PARTICLE_RECORD))
        if np_:
            lib().or_ibm_get_particles(self.h, _p(out))
        return out

    def ibm_axes(self) -> np.ndarray:
        """Semi-axis points (semiAxis, semiAxis2, semiAxis3) per particle, (n, 3, 3)."""
        np_, _ = self.ibm_counts()
        out = np.zeros((np_, 3, 3))
        if np_:
            lib().or_ibm_get_axes(self.h, _p(out))
        return out

    def ibm_set_particles(self, rec):
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        lib().or_ibm_set_particles(self.h, _p(rec))

    def ibm_collision_state(self):
        np_, _ = self.ibm_counts()
        partner = np.zeros((np_, MAX_COLLISIONS), dtype=np.int32)
        last = np.zeros((np_, MAX_COLLISIONS), dtype=np.int32)
        tang = np.zeros((np_, MAX_COLLISIONS, 3))
        if np_:
            lib().or_ibm_get_collisions(self.h, partner.ctypes.data_as(C.POINTER(C.c_int)), _p(tang),
                                        last.ctypes.data_as(C.POINTER(C.c_int)))
        return partner, tang, last

    def ibm_nodes(self):
        _, n = self.ibm_counts()
        pos, f, df = np.zeros((3, n)), np.zeros((3, n)), np.zeros((3, n))
        if n:
            lib().or_ibm_get_nodes(self.h, _p(pos), _p(f), _p(df))
        return pos, f, df

    def ibm_force(self) -> np.ndarray:
        out = np.zeros((3,) + self.shape)
        lib().or_ibm_get_force(self.h, _p(out))
        return out

    def ibm_set_force(self, f):
        lib().or_ibm_set_force(self.h, _p(np.ascontiguousarray(f, dtype=np.float64)))

    def ibm_interp_spread(self) -> np.ndarray:
        out = np.zeros((6,) + self.shape)
        lib().or_ibm_interp_spread(self.h, _p(out))
        return out

    def ibm_update_macr(self):
        lib().or_ibm_update_macr(self.h)

    def ibm_set_macr(self, rho, ux, uy, uz):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, ux, uy, uz)]
        lib().or_ibm_set_macr(self.h, *[_p(a) for a in arrs])

    def set_map(self, words):
        """Test hook: replaces the node map ([NZ][NY][NX] words) and the BC list."""
        w = np.ascontiguousarray(words, dtype=np.uint32)
        lib().or_set_map(self.h, w.ctypes.data_as(C.POINTER(C.c_uint32)))

    def apply_bc_raw(self, post_stream, post_col):
        ps = np.ascontiguousarray(post_stream, dtype=np.float64).copy()
        pc = np.ascontiguousarray(post_col, dtype=np.float64)
        lib().or_apply_bc_raw(self.h, _p(ps), _p(pc))
        return ps


def ibb_q(kind, x1, y1, x2, y2, r) -> float:
    """Interpolated bounce-back wall distance of a link (kind 0: along x,
    1: along y, 2: diagonal slope +1, 3: slope -1), as the oracle's BC uses it."""
    return lib().or_ibb_q(int(kind), float(x1), float(y1), float(x2), float(y2), float(r))


def max_threads() -> int:
    return lib().or_max_threads()


def ibm_desc(**kw) -> OrIbmDesc:
    """Oracle IBM descriptor; defaults are the reference's (src/CUDA/IBM/ibmVar.h,
    src/CUDA/IBM/ibmBoundaryCondition.h)."""
    d = OrIbmDesc(max_iterations=1, periodic_x=1, periodic_y=0, periodic_z=1, thickness=1.0, alpha=0.5,
                  fluid_density=1.0, gx=0.0, gy=0.0, gz=0.0, pp_friction=0.1, pw_friction=0.1, pp_rest=0.9,
                  pw_rest=0.9, young=0.5, poisson=0.3, wall_young=100.0, wall_poisson=0.3, ref_float_helpers=1)
    for k, v in kw.items():
        setattr(d, k, v)
    return d


def ibm_desc_from(vdesc) -> OrIbmDesc:
    """Oracle IBM descriptor with the same settings as a product vlbm_ibm_desc."""
    d = OrIbmDesc()
    for name, _ in OrIbmDesc._fields_:
        setattr(d, name, getattr(vdesc, name))
    return d


def sphere_polar(diameter, center, mesh_scale=1.0, breugem=0.0, coulomb=0):
    """Layered polar sphere mesh of the reference: (xyz [n,3], S [n])."""
    c = np.ascontiguousarray(center, dtype=np.float64)
    n = lib().or_sphere_polar(float(diameter), _p(c), float(mesh_scale), float(breugem), int(coulomb),
                              None, None, 0)
    xyz, S = np.zeros((n, 3)), np.zeros(n)
    lib().or_sphere_polar(float(diameter), _p(c), float(mesh_scale), float(breugem), int(coulomb),
                          _p(xyz), _p(S), n)
    return xyz, S


def ibm_cells(shape_xyz, r_max, pos):
    """(dims[3], cell[n]) of the uniform contact cell list for positions pos[n, 3]."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    dims = np.zeros(3, dtype=np.int32)
    cells = np.zeros(max(len(pos), 1), dtype=np.int32)
    ip = C.POINTER(C.c_int)
    lib().or_ibm_cells(int(shape_xyz[0]), int(shape_xyz[1]), int(shape_xyz[2]), float(r_max), _p(pos), len(pos),
                       dims.ctypes.data_as(ip), cells.ctypes.data_as(ip))
    return dims, cells[:len(pos)]


def stencil4(x: float) -> float:
    return lib().or_stencil4(float(x))
