"""Simulation definitions from pyvisc-style YAML.

Accepts the key set of the reference's ``SchemeVarH`` (pyvisc/schemes.py:49-74,
rendered into var.h by pyvisc/templates/var.h.j2:17-111) with the same types,
literals and defaults, read from a YAML list of maps of which only the first
entry is used (pyvisc/conf_handler.py:13-22). The reference keeps the
non-Newtonian parameters, boundary values and the boundary builder outside the
YAML (src/CUDA/NNF/nnf.h:10-24, src/CUDA/var.h:114-160, a copied builder
file); here they are optional extra keys with the reference defaults:

    builder: parallelPlatesBounceBack_Y   # name of the reference builder file
    rho_0: 1.0                            # RHO_0
    u_max: 0.1                            # U_MAX (default UX/UY/UZ_BC tables)
    eta_p: 0.3                            # Bingham ETA_P (default (tau-0.5)/3)
    yield_stress: 1.0e-4                  # Bingham S_Y
    n_index: 0.5                          # power-law N_INDEX
    k_consistency: ...                    # default RHO_0*(tau-0.5)/3
    gamma_0: 0.0                          # GAMMA_0
    tau_clamp: [0.51, 2.0]                # apparent tau clamp (extension)
    ux_bc / uy_bc / uz_bc / rho_bc: [..]  # up to 8 values each
    duct_outer_radius, duct_inner_radius, outer_rotation, outer_velocity,
    inner_rotation, inner_velocity        # duct builders
    zouhe_textbook: false                 # textbook Zou-He instead of the reference lines
    collision: regularized                # or bgk (collisionSchemes/recursiveCollisionForceNoReg.txt)
    les_const: 0.0                        # LES_MODEL Implicit_const, Newtonian only (lbm.cu:226-242)
    les_visc_turb: false                  # LES_EXPORT_VISC_TURBULENT: keep and save tau_t/3 per node
    layout: auto                          # state in HBM: auto | populations | moments
    rounding: stochastic                  # 16-bit storage (half, bfloat16): stochastic | nearest
    checkpoint_save: 0                    # CHECKPOINT_SAVE interval (var.h:69-71)
    load_checkpoint: null                 # folder to resume from: LOAD_CHECKPOINT with
                                          # SIMULATION_FOLDER_LOAD_CHECKPOINT (var.h:77-84)
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional

import yaml

from . import _capi

BUILDERS = {
    "taylorGreenVortex": 0,
    "parallelPlatesBounceBack_Y": 1,
    "parallelPlatesBounceBack_X": 2,
    "parallelPlatesBounceBack_Z": 3,
    "squareDuctBounceBack": 4,
    "closedBoxBounceBack": 5,
    "parallelPlatesZouHe": 6,
    "lidDrivenBounceBack": 7,
    "couetteZouHe": 8,
    "inflowFreeSlip": 9,
    "parallelPlatesBounceBackFreeSlip": 10,
    "squareDuctZouHe": 11,
    "circularDuctInterpBounceBack": 12,
    "annularDuctInterpBounceBack": 13,
    "parallelPlatesRheometer": 14,
}

PRECISIONS = {"double": _capi.FP64, "single": _capi.FP32, "half": _capi.FP16, "bfloat16": _capi.BF16}
ROUNDINGS = {"stochastic": 0, "nearest": 1}
MODELS = {None: _capi.NEWTONIAN, "BINGHAM": _capi.BINGHAM, "POWERLAW": _capi.POWERLAW}
COLLISIONS = {"regularized": 0, "bgk": 1}
LAYOUTS = {"auto": 0, "populations": 1, "moments": 2}


@dataclass
class SimConfig:
    # pyvisc SchemeVarH keys (required ones first)
    precision: str
    NX: int
    NY: int
    NZ: int
    tau: float
    vel_set: str
    use_ibm: bool
    non_newtonian_model: Optional[str]
    sim_id: int
    path_save: str
    steps: int
    macr_save: int
    data_report: int
    FX: float = 0.0
    FY: float = 0.0
    FZ: float = 0.0
    data_stop: bool = False
    data_save: bool = False
    pop_save: bool = False
    ini_step: int = 0
    n_gpus: int = 1
    resid_max: float = 1e-5
    # keys the reference sets outside the YAML
    builder: str = "closedBoxBounceBack"
    rho_0: float = 1.0
    u_max: float = 0.1
    eta_p: Optional[float] = None
    yield_stress: float = 1e-4
    n_index: float = 0.5
    k_consistency: Optional[float] = None
    gamma_0: float = 0.0
    tau_clamp: Optional[List[float]] = None
    ux_bc: Optional[List[float]] = None
    uy_bc: Optional[List[float]] = None
    uz_bc: Optional[List[float]] = None
    rho_bc: Optional[List[float]] = None
    duct_outer_radius: Optional[float] = None
    duct_inner_radius: Optional[float] = None
    outer_rotation: float = 0.0
    outer_velocity: float = 0.0
    inner_rotation: float = 0.0
    inner_velocity: float = 0.0
    zouhe_textbook: bool = False
    collision: str = "regularized"
    les_const: float = 0.0
    les_visc_turb: bool = False
    layout: str = "auto"
    rounding: str = "stochastic"
    checkpoint_save: int = 0
    load_checkpoint: Optional[str] = None
    # immersed boundary (use_ibm): vlbm_ibm_desc fields (ibmVar.h /
    # ibmBoundaryCondition.h constants) and the spheres, each a map with
    # center, radius and optional density, vel, w, movable, mesh
    # ("polar" | "fibonacci"), nodes (Fibonacci node count)
    ibm: Optional[Dict[str, Any]] = None
    spheres: Optional[List[Dict[str, Any]]] = None

    def __post_init__(self):
        _validate(self)

    @property
    def id_sim(self) -> str:
        """ID_SIM as rendered by var.h.j2: sim_id zero-padded to 3 digits."""
        return "%03d" % self.sim_id

    def to_ibm_desc(self) -> _capi.VlbmIbmDesc:
        """IBM descriptor: library defaults, then the ibm map; capacities sized
        for the listed spheres."""
        d = _capi.default_ibm_desc(**(self.ibm or {}))
        spheres = self.spheres or []
        if not self.ibm or "max_spheres" not in self.ibm:
            d.max_spheres = max(d.max_spheres, len(spheres))
        return d

    def to_desc(self) -> _capi.VlbmDesc:
        d = _capi.default_desc()
        d.nx, d.ny, d.nz = self.NX, self.NY, self.NZ
        d.q = 19 if self.vel_set == "D3Q19" else 27
        d.precision = PRECISIONS[self.precision]
        d.rheology = MODELS[self.non_newtonian_model]
        d.builder = BUILDERS[self.builder]
        d.tau = self.tau
        d.rho0 = self.rho_0
        d.fx, d.fy, d.fz = self.FX, self.FY, self.FZ
        d.eta_p = self.eta_p if self.eta_p is not None else (self.tau - 0.5) / 3.0
        d.s_y = self.yield_stress
        d.n_index = self.n_index
        d.k_consistency = (self.k_consistency if self.k_consistency is not None
                           else self.rho_0 * (self.tau - 0.5) / 3.0)
        d.gamma_0 = self.gamma_0
        if self.tau_clamp is not None:
            d.omega_clamp = 1
            d.tau_min, d.tau_max = float(self.tau_clamp[0]), float(self.tau_clamp[1])
        um = self.u_max
        tables = {
            "ux_bc": [0, um, 0, 0, 0, 0, 0, 0],
            "uy_bc": [0, um / 2, -um / 2, 0, 0, 0, 0, 0],
            "uz_bc": [0, um, -um, 0, 0, 0, 0, 0],
            "rho_bc": [self.rho_0, 1, 1, 1, 1, 1, 1, 1],
        }
        for name, default in tables.items():
            vals = getattr(self, name)
            vals = list(vals) if vals is not None else default
            vals = (vals + default[len(vals):])[:8]
            arr = getattr(d, name)
            for k in range(8):
                arr[k] = float(vals[k])
        outer = self.duct_outer_radius if self.duct_outer_radius is not None else self.NY / 2.0 - 1.5
        d.duct_outer_radius = outer
        d.duct_inner_radius = self.duct_inner_radius if self.duct_inner_radius is not None else outer / 4.0
        d.outer_rotation, d.outer_velocity = self.outer_rotation, self.outer_velocity
        d.inner_rotation, d.inner_velocity = self.inner_rotation, self.inner_velocity
        d.zouhe_textbook = 1 if self.zouhe_textbook else 0
        d.collision = COLLISIONS[self.collision]
        d.les_const = self.les_const
        d.les_visc_turb = 1 if self.les_visc_turb else 0
        d.layout = LAYOUTS[self.layout]
        d.rounding = ROUNDINGS[self.rounding]
        return d

    def replace(self, **kw) -> "SimConfig":
        return dataclasses.replace(self, **kw)


def _validate(c: SimConfig) -> None:
    def need(cond: bool, msg: str):
        if not cond:
            raise ValueError(msg)

    # numbers are coerced from strings as pydantic does (YAML 1.1 reads 1e-5 as a string)
    for name in ("NX", "NY", "NZ", "sim_id", "steps", "macr_save", "data_report", "ini_step", "n_gpus",
                 "checkpoint_save"):
        v = getattr(c, name)
        if isinstance(v, str):
            try:
                v = int(v)
            except ValueError:
                pass
        need(isinstance(v, int) and not isinstance(v, bool), f"{name} must be an integer")
        setattr(c, name, v)
    floats = ("tau", "FX", "FY", "FZ", "resid_max", "rho_0", "u_max", "yield_stress", "n_index", "gamma_0",
              "outer_rotation", "outer_velocity", "inner_rotation", "inner_velocity", "les_const")
    opt_floats = ("eta_p", "k_consistency", "duct_outer_radius", "duct_inner_radius")
    for name in floats + opt_floats:
        v = getattr(c, name)
        if v is None and name in opt_floats:
            continue
        if isinstance(v, str):
            try:
                v = float(v)
            except ValueError:
                pass
        need(isinstance(v, (int, float)) and not isinstance(v, bool), f"{name} must be a number")
        setattr(c, name, float(v))
    for name in ("use_ibm", "data_stop", "data_save", "pop_save", "zouhe_textbook", "les_visc_turb"):
        need(isinstance(getattr(c, name), bool), f"{name} must be a boolean")
    need(c.checkpoint_save >= 0, "checkpoint_save must be >= 0")
    need(c.collision in COLLISIONS, "collision must be 'regularized' or 'bgk'")
    need(c.layout in LAYOUTS, "layout must be 'auto', 'populations' or 'moments'")
    need(c.rounding in ROUNDINGS, "rounding must be 'stochastic' or 'nearest'")
    need(c.les_const >= 0, "les_const must be >= 0")
    need(c.les_const == 0 or c.non_newtonian_model is None, "LES applies to Newtonian fluids only")
    need(not c.les_visc_turb or c.les_const > 0, "les_visc_turb needs LES (les_const > 0)")
    need(c.load_checkpoint is None or isinstance(c.load_checkpoint, str), "load_checkpoint must be a folder name")
    need(c.precision in PRECISIONS, "precision must be 'single' or 'double' (or 'half' / 'bfloat16')")
    need(c.vel_set in ("D3Q19", "D3Q27"), "vel_set must be 'D3Q19' or 'D3Q27'")
    need(c.non_newtonian_model in MODELS, "non_newtonian_model must be null, BINGHAM or POWERLAW")
    need(c.builder in BUILDERS, f"unknown builder {c.builder!r}")
    need(isinstance(c.path_save, str), "path_save must be a string")
    need(c.n_gpus == 1, "this library runs one lattice per GPU: n_gpus must be 1")
    if c.ibm is not None:
        need(isinstance(c.ibm, dict), "ibm must be a map of IBM descriptor fields")
        known = {f for f, _ in _capi.VlbmIbmDesc._fields_}
        bad = set(c.ibm) - known
        need(not bad, f"unknown ibm keys: {sorted(bad)}")
    for sp in c.spheres or []:
        need(isinstance(sp, dict) and "center" in sp and "radius" in sp,
             "each sphere needs center and radius")
        need(sp.get("mesh", "polar") in ("polar", "fibonacci"), "sphere mesh must be 'polar' or 'fibonacci'")
    need(c.NX >= 1 and c.NY >= 1 and c.NZ >= 1, "NX, NY, NZ must be >= 1")


def from_dict(cfg: Dict[str, Any]) -> SimConfig:
    names = {f.name for f in dataclasses.fields(SimConfig)}
    unknown = set(cfg) - names
    if unknown:
        raise ValueError(f"unknown keys: {sorted(unknown)}")
    return SimConfig(**cfg)


def read_configs(filename: str) -> SimConfig:
    """First entry of a YAML list of simulation maps (pyvisc/conf_handler.py:13-22)."""
    try:
        with open(filename, "r") as f:
            cfg_dicts = yaml.safe_load(f)
        return from_dict(dict(cfg_dicts[0]))
    except Exception as e:
        raise ValueError(f"Unable to load YAML from {filename}. Exception {e}") from e
