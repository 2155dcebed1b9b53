"""Particle-laden flow benchmark on one GPU: D3Q19 regularized Newtonian fluid
(tau 0.6) in a closed 256 x 256 x 512 box with bounce-back walls and gravity
g_z = -1e-5, 200 movable spheres of radius 8 with 804 Fibonacci nodes each,
density ratio 2, seeded uniform non-overlapping centres, Hertz-Mindlin
contacts (E = 1e3, Poisson 0.3, friction 0.1, restitution 0.9 for
sphere-sphere and sphere-wall), 3 IBM iterations per step.

Reports MLUPS of the coupled step, the IBM stage's share of the device time
(HIP events around the LBM kernel and the stage of every step) and the
trajectory deviation of the FP32-fluid run from the FP64 run of the same
case (particles are FP64 in both).

    python -m vlbm.ibm_bench [--steps 5000] [--warmup 20] [--spheres 200] [--no-fp64]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _capi
from .config import SimConfig
from .lattice import Lattice

SHAPE = (256, 256, 512)      # x, y, z
RADIUS = 8.0
NODES = 804


def centres(n, seed=2024, shape=SHAPE, r=RADIUS):
    """Uniform in [r+1, L-r-2] per axis, rejection of overlaps (distance < 2r)."""
    rng = np.random.default_rng(seed)
    lo = r + 1
    hi = np.array(shape, dtype=float) - r - 2
    out = []
    while len(out) < n:
        c = lo + rng.random(3) * (hi - lo)
        if all(np.linalg.norm(c - p) >= 2 * r for p in out):
            out.append(c)
    return np.array(out)


def make(precision, n_spheres, seed, layout="auto"):
    cfg = SimConfig(precision=precision, layout=layout, NX=SHAPE[0], NY=SHAPE[1], NZ=SHAPE[2], tau=0.6, vel_set="D3Q19",
                    use_ibm=True, non_newtonian_model=None, sim_id=0, path_save="bench", steps=5000,
                    macr_save=0, data_report=0, builder="closedBoxBounceBack")
    lat = Lattice(cfg.to_desc())
    lat.init()
    idesc = _capi.default_ibm_desc(max_iterations=3, periodic_x=0, periodic_y=0, periodic_z=0, gz=-1e-5,
                                   young=1e3, poisson=0.3, wall_young=1e3, wall_poisson=0.3, pp_friction=0.1,
                                   pw_friction=0.1, pp_rest=0.9, pw_rest=0.9, fluid_density=1.0,
                                   max_spheres=max(n_spheres, 1), max_nodes=max(n_spheres, 1) * NODES)
    lat.ibm_enable(idesc)
    for c in centres(n_spheres, seed):
        lat.ibm_add_sphere(c, radius=RADIUS, density=2.0, mesh=_capi.MESH_FIBONACCI, n_fibonacci=NODES)
    return lat


def run(precision, steps, warmup, n_spheres, seed, layout="auto"):
    print(f"[ibm_bench] {precision}: set-up of {n_spheres} spheres", file=sys.stderr, flush=True)
    lat = make(precision, n_spheres, seed, layout)
    used = lat.layout
    lat.step(warmup)
    lat.sync()
    print(f"[ibm_bench] {precision}: timing {steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ms_lbm, ms_ibm = lat.ibm_time_steps(steps)
    wall = time.perf_counter() - t0
    lat.ibm_status()
    rec = lat.ibm_particles()
    lat.close()
    return rec, wall, ms_lbm, ms_ibm, used


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--spheres", type=int, default=200)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--layout", default="auto", choices=["auto", "populations", "moments"],
                    help="state layout of the FP32 lattice (auto: the library's choice)")
    a = ap.parse_args()
    nodes = SHAPE[0] * SHAPE[1] * SHAPE[2]
    rec32, wall, ms_lbm, ms_ibm, used = run("single", a.steps, a.warmup, a.spheres, a.seed, a.layout)
    total = a.warmup + a.steps
    out = {
        "metric": "MLUPS, IBM particle-laden box", "value": round(nodes * a.steps / wall / 1e6, 1),
        "unit": "MLUPS", "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(wall * 1e3 / a.steps, 4),
        "ibm_share": round(ms_ibm / (ms_lbm + ms_ibm), 4),
        "ms_lbm_per_step": round(ms_lbm / a.steps, 4), "ms_ibm_per_step": round(ms_ibm / a.steps, 4),
        "dtype": "f32 fluid, f64 particles", "data": "synthetic",
        "config": {"workload": "IBM box 256x256x512 D3Q19 tau 0.6, 200 spheres r 8, 804 Fibonacci nodes, "
                               "3 iterations", "spheres": a.spheres, "seed": a.seed, "layout": used},
    }
    if not a.no_fp64:
        rec64, wall64, _, _, _ = run("double", a.steps, a.warmup, a.spheres, a.seed)
        d = np.linalg.norm(rec32[:, 0:3] - rec64[:, 0:3], axis=1)
        travel = np.linalg.norm(rec64[:, 0:3] - centres(a.spheres, a.seed), axis=1)
        out["trajectory_deviation"] = {
            "steps": total, "max_abs": float(d.max()), "mean_abs": float(d.mean()),
            "max_over_travel": float(d.max() / max(travel.max(), 1e-300)),
            "fp64_mlups": round(nodes * a.steps / wall64 / 1e6, 1)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
