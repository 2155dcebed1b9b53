"""Serving of many Bingham channel simulations on one GPU.

`Batch` wraps the batched engine of the C-ABI (vlbm_batch_*): jobs are
queued, admitted by the HBM budget as memory and slots free up, stepped
together by one persistent-grid kernel (captured step sequences), checked
with the reference's residual and retired with their macroscopics.

`sweep()` drives a seeded Poisson stream of jobs (arrival times on the wall
clock) through a Batch and reports jobs/s, aggregate MLUPS and the p50/p99
time from arrival to result; `sequential()` runs the same jobs one at a time
through the plain single-lattice path for comparison.

Bingham number of a job (the reference has none): Bn = tau_y H / (rho nu_p U)
with U the mean velocity of the analytic Bingham channel profile and H the
channel width; tau_y is found by bisection, so the yield half-width stays
inside the channel for every Bn (Bn = 0 is the Newtonian limit, which takes
omega = omega_p).

    python -m vlbm.serve [--jobs 200] [--rate 20] [--max-steps 5000] [--no-sequential]
"""
from __future__ import annotations

import argparse
import dataclasses
import ctypes as C
import json
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _capi
from ._capi import check, lib

JOB_STATES = {0: "pending", 1: "running", 2: "done", 3: "max_steps", 4: "failed"}


class VlbmBatchDesc(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_size_t), ("max_jobs", C.c_int), ("q", C.c_int), ("precision", C.c_int),
        ("check_every", C.c_int), ("steps_per_launch", C.c_int), ("block", C.c_int), ("grid", C.c_int),
        ("use_graph", C.c_int), ("tune_cache", C.c_char_p), ("max_resident_nodes", C.c_size_t),
    ]


class VlbmJobInfo(C.Structure):
    _fields_ = [
        ("state", C.c_int), ("steps", C.c_longlong), ("residual", C.c_double), ("mean_rho", C.c_double),
        ("t_submit", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double), ("bytes", C.c_size_t),
    ]


def _bind():
    P = C.POINTER
    dp = P(C.c_double)
    ip = P(C.c_int)
    sig = {
        "vlbm_batch_desc_defaults": (None, [P(VlbmBatchDesc)]),
        "vlbm_batch_create": (C.c_int, [P(VlbmBatchDesc), P(C.c_void_p)]),
        "vlbm_batch_destroy": (None, [C.c_void_p]),
        "vlbm_batch_job_bytes": (C.c_int, [C.c_void_p, P(_capi.VlbmDesc), P(C.c_size_t)]),
        "vlbm_batch_submit": (C.c_int, [C.c_void_p, P(_capi.VlbmDesc), C.c_double, C.c_longlong, ip]),
        "vlbm_batch_advance": (C.c_int, [C.c_void_p, ip, ip, ip]),
        "vlbm_batch_job_info": (C.c_int, [C.c_void_p, C.c_int, P(VlbmJobInfo)]),
        "vlbm_batch_job_result": (C.c_int, [C.c_void_p, C.c_int, dp, dp, dp, dp]),
        "vlbm_batch_release": (C.c_int, [C.c_void_p, C.c_int]),
        "vlbm_batch_config": (C.c_int, [C.c_void_p, ip, ip, ip]),
        "vlbm_batch_clock": (C.c_double, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


_bind()


def batch_desc(**kw) -> VlbmBatchDesc:
    d = VlbmBatchDesc()
    lib.vlbm_batch_desc_defaults(C.byref(d))
    for k, v in kw.items():
        if k == "tune_cache" and isinstance(v, str):
            v = v.encode()
        setattr(d, k, v)
    return d


class Batch:
    def __init__(self, desc: Optional[VlbmBatchDesc] = None, **kw):
        self.desc = desc or batch_desc(**kw)
        self._cache = self.desc.tune_cache   # keep the bytes alive
        h = C.c_void_p()
        check(lib.vlbm_batch_create(C.byref(self.desc), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.vlbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def job_bytes(self, d: _capi.VlbmDesc) -> int:
        out = C.c_size_t(0)
        check(lib.vlbm_batch_job_bytes(self._h, C.byref(d), C.byref(out)))
        return out.value

    def submit(self, d: _capi.VlbmDesc, resid_max: float, max_steps: int) -> int:
        jid = C.c_int(-1)
        check(lib.vlbm_batch_submit(self._h, C.byref(d), float(resid_max), int(max_steps), C.byref(jid)))
        return jid.value

    def advance(self):
        r, p, f = C.c_int(0), C.c_int(0), C.c_int(0)
        check(lib.vlbm_batch_advance(self._h, C.byref(r), C.byref(p), C.byref(f)))
        return r.value, p.value, f.value

    def info(self, jid: int) -> Dict:
        i = VlbmJobInfo()
        check(lib.vlbm_batch_job_info(self._h, int(jid), C.byref(i)))
        return {"state": JOB_STATES[i.state], "steps": i.steps, "residual": i.residual, "mean_rho": i.mean_rho,
                "t_submit": i.t_submit, "t_start": i.t_start, "t_end": i.t_end, "bytes": i.bytes}

    def result(self, jid: int, shape) -> Dict[str, np.ndarray]:
        out = {k: np.zeros(shape) for k in ("rho", "ux", "uy", "uz")}
        dp = C.POINTER(C.c_double)
        check(lib.vlbm_batch_job_result(self._h, int(jid), *[out[k].ctypes.data_as(dp) for k in ("rho", "ux", "uy", "uz")]))
        return out

    def release(self, jid: int):
        check(lib.vlbm_batch_release(self._h, int(jid)))

    def config(self):
        k, b, g = C.c_int(0), C.c_int(0), C.c_int(0)
        check(lib.vlbm_batch_config(self._h, C.byref(k), C.byref(b), C.byref(g)))
        return {"steps_per_launch": k.value, "block": b.value, "grid": g.value}

    def clock(self) -> float:
        return lib.vlbm_batch_clock(self._h)


# ---------------------------------------------------------------------------
# Bingham channel jobs
# ---------------------------------------------------------------------------
def bingham_profile(s, h, F, nu_p, tau_y):
    """Analytic Bingham channel velocity at distance s from the centre line
    (half-width h, body force F, plastic viscosity nu_p, rho = 1)."""
    y0 = min(tau_y / F, h) if F > 0 else h
    s = np.abs(s)
    u = (F / (2 * nu_p)) * (h * h - s * s) - (tau_y / nu_p) * (h - s)
    u_plug = (F / (2 * nu_p)) * (h - y0) ** 2
    return np.where(s >= y0, u, u_plug)


def mean_velocity(h, F, nu_p, tau_y, n=2001):
    s = np.linspace(-h, h, n)
    return float(np.trapezoid(bingham_profile(s, h, F, nu_p, tau_y), s) / (2 * h))


def tau_y_for_bn(bn, h, F, nu_p):
    """tau_y with tau_y * 2h / (nu_p * U(tau_y)) = bn, by bisection on [0, F h)."""
    if bn <= 0:
        return 0.0
    H = 2 * h
    lo, hi = 0.0, F * h * (1 - 1e-9)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        U = mean_velocity(h, F, nu_p, mid)
        val = mid * H / (nu_p * U) if U > 0 else math.inf
        if val < bn:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


@dataclass
class JobSpec:
    nx: int
    ny: int
    nz: int
    bn: float
    tau: float
    fx: float
    tau_y: float
    arrival: float

    def desc(self, precision=_capi.FP32) -> _capi.VlbmDesc:
        d = _capi.default_desc()
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        d.q = 19
        d.precision = precision
        d.rheology = _capi.BINGHAM
        d.builder = 1   # parallelPlatesBounceBack_Y: walls at y = 0 and NY-1, periodic x and z
        d.tau = self.tau
        d.eta_p = (self.tau - 0.5) / 3.0
        d.s_y = self.tau_y
        d.fx, d.fy, d.fz = self.fx, 0.0, 0.0
        d.rho0 = 1.0
        return d

    def analytic_ux(self) -> np.ndarray:
        # halfway bounce-back: walls at y = -1/2 and NY - 1/2
        h = self.ny / 2.0
        y = np.arange(self.ny)
        return bingham_profile(y - (self.ny - 1) / 2.0, h, self.fx, (self.tau - 0.5) / 3.0, self.tau_y)


def make_jobs(n, rate=20.0, seed=7, nx=(32, 256), ny=(32, 128), nz=(16, 128), bn=(0.0, 5.0),
              tau=(0.55, 1.5), fx=(1e-6, 1e-5)) -> List[JobSpec]:
    """Seeded Poisson stream of Bingham channel jobs (exponential gaps of mean 1/rate)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        j_nx = int(rng.integers(nx[0], nx[1] + 1))
        j_ny = int(rng.integers(ny[0], ny[1] + 1))
        j_nz = int(rng.integers(nz[0], nz[1] + 1))
        j_bn = float(rng.uniform(*bn))
        j_tau = float(rng.uniform(*tau))
        j_fx = float(rng.uniform(*fx))
        ty = tau_y_for_bn(j_bn, j_ny / 2.0, j_fx, (j_tau - 0.5) / 3.0)
        jobs.append(JobSpec(j_nx, j_ny, j_nz, j_bn, j_tau, j_fx, ty, t))
    return jobs


def profile_error(spec: JobSpec, ux: np.ndarray) -> float:
    """Relative L2 of the x-z averaged u_x(y) against the analytic profile."""
    prof = ux.mean(axis=(0, 2))
    ref = spec.analytic_ux()
    return float(np.linalg.norm(prof - ref) / max(np.linalg.norm(ref), 1e-300))


@dataclass
class SweepResult:
    jobs: int
    seconds: float
    node_updates: float
    latencies: List[float] = field(default_factory=list)
    errors: List[float] = field(default_factory=list)
    states: Dict[str, int] = field(default_factory=dict)
    records: List[Dict] = field(default_factory=list)      # one per job, in completion order

    def summary(self) -> Dict:
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        err = np.array(self.errors) if self.errors else np.zeros(1)
        return {"jobs": self.jobs, "seconds": round(self.seconds, 3),
                "jobs_per_s": round(self.jobs / self.seconds, 3),
                "mlups": round(self.node_updates / self.seconds / 1e6, 1),
                "p50_s": round(float(np.percentile(lat, 50)), 3), "p99_s": round(float(np.percentile(lat, 99)), 3),
                "profile_err_median": float(np.median(err)), "profile_err_max": float(err.max()),
                "states": self.states}


def sweep(jobs: List[JobSpec], resid_max=1e-6, max_steps=50000, check_every=1000, realtime=True,
          keep=None, log=None, sink=None, **batch_kw) -> SweepResult:
    """Continuous batching of a job stream: arrivals are submitted when their
    time comes (wall clock from the start) while earlier jobs run."""
    b = Batch(check_every=check_every, **batch_kw)
    ids = {}
    t_start = time.perf_counter()
    res = SweepResult(jobs=len(jobs), seconds=0.0, node_updates=0.0)
    nxt = 0
    done = 0
    results = {}
    while done < len(jobs):
        now = time.perf_counter() - t_start
        while nxt < len(jobs) and (not realtime or jobs[nxt].arrival <= now):
            ids[b.submit(jobs[nxt].desc(), resid_max, max_steps)] = nxt
            nxt += 1
        running, pending, finished = b.advance()
        if finished or (running == 0 and pending == 0):
            for jid, k in list(ids.items()):
                inf = b.info(jid)
                if inf["state"] in ("done", "max_steps", "failed"):
                    spec = jobs[k]
                    shape = (spec.nz, spec.ny, spec.nx)
                    out = b.result(jid, shape)
                    b.release(jid)
                    t_done = time.perf_counter() - t_start
                    arrival = spec.arrival if realtime else 0.0
                    res.latencies.append(t_done - arrival)
                    res.errors.append(profile_error(spec, out["ux"]))
                    res.node_updates += float(spec.nx * spec.ny * spec.nz) * inf["steps"]
                    res.states[inf["state"]] = res.states.get(inf["state"], 0) + 1
                    res.records.append({"job": k, "state": inf["state"], "steps": inf["steps"],
                                        "residual": inf["residual"], "latency_s": res.latencies[-1],
                                        "profile_err": res.errors[-1], "nodes": spec.nx * spec.ny * spec.nz,
                                        "t_done_s": t_done})
                    if sink:
                        sink(res.records[-1])
                    if keep is not None:
                        keep[k] = {"info": inf, "ux": out["ux"], "rho": out["rho"]}
                    del ids[jid]
                    done += 1
                    if log:
                        log(f"job {k} {inf['state']} after {inf['steps']} steps, residual {inf['residual']:.2e}")
        if running == 0 and pending == 0 and nxt < len(jobs):
            time.sleep(max(0.0, min(0.05, jobs[nxt].arrival - (time.perf_counter() - t_start))))
    res.seconds = time.perf_counter() - t_start
    res.config = b.config()
    b.close()
    return res


def sequential(jobs: List[JobSpec], resid_max=1e-6, max_steps=50000, check_every=1000, log=None,
               realtime=True, sink=None) -> SweepResult:
    """The same jobs one at a time through the plain single-lattice path,
    with the same stop test and check cadence (first step direct, then
    checks after 1 + k * check_every steps), in arrival order: a job starts
    when it has arrived and the previous one is done; latency is measured
    from its arrival."""
    from .lattice import Lattice
    res = SweepResult(jobs=len(jobs), seconds=0.0, node_updates=0.0)
    t_start = time.perf_counter()
    for k, spec in enumerate(jobs):
        if realtime:
            wait = spec.arrival - (time.perf_counter() - t_start)
            if wait > 0:
                time.sleep(wait)
        lat = Lattice(spec.desc())
        lat.init()
        steps = 0
        lat.step(1)
        steps = 1
        state = "max_steps"
        while True:
            lat.step(check_every)
            steps += check_every
            lat.compute_macroscopics()     # moments of the current state, as the batched check
            r, m = lat.reduce()
            if not math.isfinite(m) or not math.isfinite(r) or m < 0:
                state = "failed"
                break
            if r < resid_max:
                state = "done"
                break
            if steps >= max_steps:
                break
        lat.sync()
        m = lat.macroscopics()
        lat.close()
        t_done = time.perf_counter() - t_start
        res.latencies.append(t_done - (spec.arrival if realtime else 0.0))
        res.errors.append(profile_error(spec, m["ux"]))
        res.node_updates += float(spec.nx * spec.ny * spec.nz) * steps
        res.states[state] = res.states.get(state, 0) + 1
        res.records.append({"job": k, "state": state, "steps": steps, "residual": r,
                            "latency_s": res.latencies[-1], "profile_err": res.errors[-1],
                            "nodes": spec.nx * spec.ny * spec.nz, "t_done_s": t_done})
        if sink:
            sink(res.records[-1])
        if log:
            log(f"sequential job {k} {state} after {steps} steps")
    res.seconds = time.perf_counter() - t_start
    return res


def stream_slice(jobs: List[JobSpec], first: int, count: int) -> List[JobSpec]:
    """Jobs [first, first + count) of a stream, arrivals shifted so the slice
    starts at its first job's arrival."""
    part = jobs[first:first + count]
    t0 = part[0].arrival if part else 0.0
    return [dataclasses.replace(j, arrival=j.arrival - t0) for j in part]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=200)
    ap.add_argument("--first", type=int, default=0,
                    help="run jobs [first, first + jobs) of the seeded stream (a part of a longer sweep)")
    ap.add_argument("--stream", type=int, default=0, help="length of the seeded stream (default: first + jobs)")
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--max-steps", type=int, default=5000)
    ap.add_argument("--resid", type=float, default=1e-6)
    ap.add_argument("--check-every", type=int, default=1000)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--tune-cache", default=None)
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--sequential-only", action="store_true")
    ap.add_argument("--records", default=None, help="write one JSON line per job to this file")
    ap.add_argument("--resident-nodes", type=float, default=float(1 << 24),
                    help="admit while the resident lattices hold at most this many nodes (0: no limit)")
    a = ap.parse_args()
    jobs = stream_slice(make_jobs(max(a.stream, a.first + a.jobs), a.rate, a.seed), a.first, a.jobs)
    progress = lambda s: print(s, flush=True)   # noqa: E731
    out = {"metric": "serving sweep: jobs/s, aggregate MLUPS, p50/p99 time from arrival to result",
           "workload": {"jobs": a.jobs, "first": a.first, "rate_per_s": a.rate, "seed": a.seed,
                        "max_steps": a.max_steps, "resid_max": a.resid, "check_every": a.check_every,
                        "max_resident_nodes": int(a.resident_nodes),
                        "lattice": "Nx U[32,256], Ny U[32,128], Nz U[16,128]; Bn U[0,5], tau U[0.55,1.5], "
                                   "Fx U[1e-6,1e-5]; D3Q19 FP32 Bingham channel"}}
    rec_file = open(a.records, "w") if a.records else None

    def sink_for(mode):
        if rec_file is None:
            return None

        def sink(r):   # one line per finished job, written as it finishes
            rec_file.write(json.dumps(dict(r, mode=mode, job=r["job"] + a.first)) + "\n")
            rec_file.flush()
        return sink

    if not a.sequential_only:
        t = time.perf_counter()
        res = sweep(jobs, a.resid, a.max_steps, a.check_every, realtime=True,
                    hbm_budget=int(a.budget_gb * (1 << 30)), max_jobs=256, tune_cache=a.tune_cache,
                    max_resident_nodes=int(a.resident_nodes), log=progress, sink=sink_for("batched"))
        out["batched"] = res.summary()
        out["batch_config"] = res.config
        progress(f"batched sweep done in {time.perf_counter() - t:.1f} s")
    if not a.no_sequential:
        seq = sequential(jobs, a.resid, a.max_steps, a.check_every, log=progress, sink=sink_for("sequential"))
        out["sequential"] = seq.summary()
        if "batched" in out:
            out["speedup_jobs_per_s"] = round(out["batched"]["jobs_per_s"] / out["sequential"]["jobs_per_s"], 2)
    if rec_file:
        rec_file.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
