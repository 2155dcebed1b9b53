"""Benchmark of the LBM time-step hot path on one MI355X.

Workload: D3Q19 regularized Bingham flow on a 256^3 lattice in FP32, velocity
Zou-He inlet u_z=0.01 at z=0, pressure Zou-He outlet rho=1 at z=255,
bounce-back walls on the x and y faces, tau=0.8, yield stress 1e-5, initial
rho=1, u=(0,0,0.01) + U(-1e-4, 1e-4) per component (seeded). One step = one
fused gather/BC/collide/store pass over the lattice.

The lattice takes the library's automatic state layout (for this workload
the moment coefficients: 10 per node instead of 19 populations); the other
layout is timed beside it. Prints one JSON line: MLUPS over the timed steps
(wall clock, inputs resident in HBM), the roofline of the step kernel that ran
(algorithmic bytes per node update of its layout: 2*10*4 = 80 B for the
coefficients, 2*19*4 = 152 B for populations, over the HIP-event time of the
same steps), the same figures for the other layout, a CPU baseline (the oracle's OpenMP
port of the reference per-node loop on the host cores, bounded sample), and
the serving sweep: the first N jobs of the seeded Poisson stream of D3Q19
FP32 Bingham channels (20 jobs/s; Nx U[32,256], Ny U[32,128], Nz U[16,128],
Bn U[0,5], tau U[0.55,1.5], F_x U[1e-6,1e-5]), each run to a velocity
residual < 1e-6 or 50,000 steps (checked every 1,000 steps), batched through
one persistent kernel and, for comparison, one at a time through the plain
single-lattice path: jobs/s, aggregate MLUPS, p50/p99 arrival-to-result
latency, converged count and the per-job error against the analytic plug
profile.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--sweep-jobs N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "viscoplastic-lbm_amd"))

METRIC = "MLUPS + %HBM roofline, D3Q19 256^3 Bingham; sweep jobs/s + p99 latency"
HBM_PEAK_GBS = 8000.0
N_SIDE = 256
Q = 19
STORAGE_BYTES = 4
# algorithmic bytes per node update (one read and one write of the state) and
# the HBM traffic summary of the layout's step kernel
LAYOUT_BYTES = {"populations": 2 * Q * STORAGE_BYTES, "moments": 2 * 10 * STORAGE_BYTES}
TRAFFIC_FILES = {"populations": "step_kernel_traffic.json", "moments": "step_tile_kernel_traffic.json"}


def workload_config(precision="single"):
    from vlbm.config import SimConfig
    return SimConfig(precision=precision, NX=N_SIDE, NY=N_SIDE, NZ=N_SIDE, tau=0.8, vel_set="D3Q19",
                     use_ibm=False, non_newtonian_model="BINGHAM", sim_id=0, path_save="bench",
                     steps=5000, macr_save=0, data_report=0, builder="squareDuctZouHe",
                     yield_stress=1e-5, uz_bc=[0, 0.01, 0, 0, 0, 0, 0, 0])


def initial_fields(seed=12345):
    shape = (N_SIDE, N_SIDE, N_SIDE)
    rng = np.random.default_rng(seed)
    rho = np.ones(shape)
    ux = rng.uniform(-1e-4, 1e-4, size=shape)
    uy = rng.uniform(-1e-4, 1e-4, size=shape)
    uz = 0.01 + rng.uniform(-1e-4, 1e-4, size=shape)
    return rho, ux, uy, uz


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(init, target_s=15.0, fp64_s=8.0):
    """The oracle (OpenMP restatement of the reference per-node path, FP32
    storage as the reference's SINGLE_PRECISION build) on the host cores,
    on a bounded number of steps of the same 256^3 workload; beside it a
    shorter sample of the FP64 (DOUBLE_PRECISION) build."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    threads = po.max_threads()
    out = {}
    for prec, budget in ((4, target_s), (8, fp64_s)):
        cfg = workload_config("single" if prec == 4 else "double")
        o = po.Oracle(po.desc_from(cfg.to_desc()), prec)
        o.init(*init)
        t1 = o.time_steps(1, threads)           # first touch + one step
        n = max(2, min(200, int(budget / max(t1, 1e-3))))
        t = o.time_steps(n, threads)
        o.close()
        out[prec] = (N_SIDE ** 3 * n / t / 1e6, n, t)
    mlups, n, t = out[4]
    m64, n64, t64 = out[8]
    return {"value": round(mlups, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{n} steps of the 256^3 D3Q19 FP32 Bingham workload in {t:.1f} s "
                      f"(oracle OpenMP loop, collapse(2) over z,y)",
            "fp64": {"value": round(m64, 2), "unit": "MLUPS",
                     "sample": f"{n64} steps of the same workload with FP64 storage in {t64:.1f} s"}}


SWEEP_RATE, SWEEP_SEED, SWEEP_RESID, SWEEP_MAX_STEPS, SWEEP_CHECK = 20.0, 7, 1e-6, 50000, 1000


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def serving_sweep(n_jobs):
    """The first n_jobs of the seeded 2,000-job stream, batched and then one at
    a time (same stop test and check cadence)."""
    from vlbm.serve import make_jobs, sequential, sweep
    jobs = make_jobs(n_jobs, SWEEP_RATE, SWEEP_SEED)
    bat = sweep(jobs, SWEEP_RESID, SWEEP_MAX_STEPS, SWEEP_CHECK, realtime=True, hbm_budget=64 << 30,
                max_jobs=256, max_resident_nodes=1 << 24)
    progress(f"sweep: {n_jobs} jobs batched in {bat.seconds:.1f} s")
    seq = sequential(jobs, SWEEP_RESID, SWEEP_MAX_STEPS, SWEEP_CHECK)
    progress(f"sweep: {n_jobs} jobs one at a time in {seq.seconds:.1f} s")
    b, q = bat.summary(), seq.summary()
    return {
        "workload": f"first {n_jobs} jobs of the seeded Poisson stream (rate {SWEEP_RATE:g}/s, seed {SWEEP_SEED}) "
                    f"of D3Q19 FP32 Bingham channels, Nx U[32,256], Ny U[32,128], Nz U[16,128], Bn U[0,5], "
                    f"tau U[0.55,1.5], F_x U[1e-6,1e-5]; each to residual < {SWEEP_RESID:g} or "
                    f"{SWEEP_MAX_STEPS} steps, checked every {SWEEP_CHECK} steps",
        "jobs": n_jobs,
        "jobs_per_s": b["jobs_per_s"], "mlups": b["mlups"], "p50_s": b["p50_s"], "p99_s": b["p99_s"],
        "converged": b["states"].get("done", 0), "capped": b["states"].get("max_steps", 0),
        "failed": b["states"].get("failed", 0),
        "profile_err_median": b["profile_err_median"], "profile_err_max": b["profile_err_max"],
        "batch_config": bat.config,
        "sequential": {"jobs_per_s": q["jobs_per_s"], "mlups": q["mlups"], "p50_s": q["p50_s"], "p99_s": q["p99_s"],
                       "converged": q["states"].get("done", 0), "profile_err_median": q["profile_err_median"],
                       "profile_err_max": q["profile_err_max"]},
        "speedup_jobs_per_s": round(b["jobs_per_s"] / q["jobs_per_s"], 3),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--settle", type=float, default=0.25, metavar="SECONDS",
                    help="step launches on a separate lattice before the warm-up (0: none)")
    ap.add_argument("--sweep-jobs", type=int, default=24,
                    help="jobs of the serving sweep (0: no sweep); the first N of the seeded stream")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="run only a plain device copy of BYTES (counter calibration) and print its time")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the library steps one "
              "lattice per GPU and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import ctypes as C
        from vlbm._capi import check, lib
        ms = C.c_double(0)
        check(lib.vlbm_copy_probe(args.copy_calibration, 10, C.byref(ms)))
        print(json.dumps({"copy_bytes": args.copy_calibration, "ms": ms.value,
                          "gbs": 2 * args.copy_calibration / (ms.value * 1e-3) / 1e9}), flush=True)
        return

    # torch is not imported: its wheel bundles its own HIP runtime, and two HIP
    # runtimes in one process do not share the device. All work runs on the
    # library's stream, so vlbm_sync() brackets the timed region.
    import vlbm

    cfg = workload_config("single")
    init = initial_fields()
    nodes = N_SIDE ** 3

    def timed(layout_cfg, settle_s):
        """W warm-up and K timed steps of a fresh lattice (after settle_s of
        the same kernel on a second lattice); the wall and HIP-event times."""
        lat = vlbm.Lattice(layout_cfg.to_desc())
        lat.init(*init)
        # Device settle: after the host-side set-up leaves the GPU idle, its
        # first few dozen step launches run 10-20 % slower than the steady
        # state (power management: 0.455 ms, rising to 0.52 ms by the tenth
        # launch, back to 0.45 ms by the 35th), so a short warm-up would time
        # that transient. The same step kernel runs on a second lattice of the
        # same shape for settle_s seconds first; the measured lattice still
        # gets exactly W warm-up and K timed steps from its initial state.
        # The second lattice is freed only after the timed steps: freeing it
        # idles the device for milliseconds, which brings the transient back.
        settled = 0
        other = None
        if settle_s > 0:
            other = vlbm.Lattice(layout_cfg.to_desc())
            other.init()
            t_settle = time.perf_counter()
            while time.perf_counter() - t_settle < settle_s:
                other.step(20)
                other.sync()
                settled += 20
        lat.step(args.warmup)
        lat.sync()
        t0 = time.perf_counter()
        event_ms = lat.time_steps(args.steps)      # HIP events on the library stream around the K steps
        lat.sync()
        wall = time.perf_counter() - t0
        if other is not None:
            other.close()
        layout = lat.layout
        # sanity of the simulated state after the timed steps
        lat.step(1, save_last=True)
        _, mean_rho = lat.reduce()
        lat.close()
        if not np.isfinite(mean_rho):
            print(f"non-finite state after the timed steps ({layout})", file=sys.stderr)
            sys.exit(1)
        kernel_ms = event_ms / args.steps
        achieved = LAYOUT_BYTES[layout] * nodes / (kernel_ms * 1e-3) / 1e9
        return {"layout": layout, "wall": wall, "kernel_ms": kernel_ms, "achieved": achieved,
                "mean_rho": mean_rho, "settled": settled}

    main_run = timed(cfg, args.settle)
    layout = main_run["layout"]
    other_layout = "populations" if layout == "moments" else "moments"
    other_cfg = workload_config("single")
    other_cfg.layout = other_layout
    other_run = timed(other_cfg, args.settle)

    wall, kernel_ms, achieved = main_run["wall"], main_run["kernel_ms"], main_run["achieved"]
    mlups = nodes * args.steps / wall / 1e6
    ms_per_step = wall * 1e3 / args.steps
    settled_steps, mean_rho = main_run["settled"], main_run["mean_rho"]
    traffic = None
    tfile = ROOT / "profiles" / TRAFFIC_FILES[layout]
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None

    progress(f"{args.steps} timed steps: {kernel_ms:.5f} ms per step launch")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(init)
        progress(f"cpu baseline: {cpu['value']} MLUPS")
    sweep_out = serving_sweep(args.sweep_jobs) if args.sweep_jobs > 0 else None

    line = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform velocity noise, no dataset)",
        "config": {
            "workload": "D3Q19 regularized Bingham, 256x256x256, FP32, Zou-He inlet/outlet z, bounce-back x/y walls",
            "lattice": [N_SIDE, N_SIDE, N_SIDE],
            "tau": 0.8, "yield_stress": 1e-5, "u_inlet": 0.01,
            "layout": layout,
            "storage": ("fp32 regularized-state coefficients (10 per node), two lattices, pull streaming through "
                        "an LDS plane ring" if layout == "moments" else
                        "fp32 deviations f_i - w_i, two lattices, pull streaming"),
            "device_settle": f"{settled_steps} step launches ({args.settle:g} s) on a separate 256^3 lattice "
                             "before the warm-up",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel_ms": round(kernel_ms, 5),
            "algorithmic_bytes_per_node": LAYOUT_BYTES[layout],
        },
        "other_layout": {
            "layout": other_layout,
            "mlups": round(nodes * args.steps / other_run["wall"] / 1e6, 1),
            "kernel_ms": round(other_run["kernel_ms"], 5),
            "algorithmic_bytes_per_node": LAYOUT_BYTES[other_layout],
            "roofline_frac": round(other_run["achieved"] / HBM_PEAK_GBS, 4),
        },
        "cpu_baseline": cpu,
        "sweep": sweep_out,
        "mean_rho": mean_rho,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
