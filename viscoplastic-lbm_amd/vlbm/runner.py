"""Driver of one simulation, following the time loop of src/CUDA/main.cu:270-548.

Per step s (counted from ini_step):
  * save = macr_save and (s - ini_step) % macr_save == 0, report likewise with
    data_report; neither on the first step (main.cu:272-287);
  * the step kernel stores the macroscopics it computes (those of the state
    entering step s) when saving, reporting or on the last step (main.cu:296);
  * on save/report steps: MLUPS since the last sync, the residual reference
    moves to this sync (main.cu:396-436), binary files labelled s
    (main.cu:460-465), report and the stop test residual < resid_max or
    mean density < 0 when data_stop is set (main.cu:468-485,
    src/CUDA/treatData.cpp:451-459);
  * checkpoint_save: every checkpoint_save steps the checkpoint folder is
    written after the step, holding the label of the step just run
    (main.cu:452-456), and once more after the loop (main.cu:546-548);
    load_checkpoint: the run starts from that folder's checkpoint and its
    stored step instead of the equilibrium initialisation (main.cu:199-203);
  * data_save: every report also updates the x-z plane averages of u_z per y
    without resetting them first, avg <- (avg + sum_xz u_z) / (NX NZ)
    (treatData, src/CUDA/treatData.cpp:57-78), and writes them to
    ID_avgUz<step>.csv (saveTreatData, treatData.cpp:476-497); once more after
    the loop, labelled with the loop counter (main.cu:553-558);
  * after the loop the saved macroscopics are written once more, labelled with
    the loop counter, so the final file holds the moments computed during the
    last step (main.cu:523-527), then ID_info.txt.
Steps between sync points run as one batch of kernel launches.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

from .config import SimConfig
from .lattice import Lattice


@dataclass
class RunResult:
    steps: int
    seconds: float
    mlups: float
    reports: List[dict] = field(default_factory=list)
    stopped: Optional[str] = None


def _next_sync(step: int, cfg: SimConfig) -> int:
    """First step >= step that saves or reports (or the last step)."""
    last = cfg.steps - 1
    best = last
    for every in (cfg.macr_save, cfg.data_report, cfg.checkpoint_save):
        if every and every > 0:
            aux = step - cfg.ini_step
            k = max(1, math.ceil(aux / every)) * every + cfg.ini_step
            best = min(best, k)
    return max(step, best)


def var_filename(path: str, sim_id: str, var: str, step: int, ext: str) -> str:
    """getVarFilename (src/CUDA/lbmReport.cpp:55-83): the step zero-padded to
    six digits (step 0 gets six zeros before it)."""
    zeros = 6 if step == 0 else sum(1 for k in range(1, 7) if step * 10 ** k < 1000000)
    return os.path.join(path, sim_id, f"{sim_id}_{var}" + "0" * zeros + str(step) + ext)


def save_avg_uz(path: str, sim_id: str, step: int, avg_uz) -> None:
    with open(var_filename(path, sim_id, "avgUz", step, ".csv"), "w") as f:
        f.write("y\tavg uz\n")
        for y, v in enumerate(avg_uz):
            f.write("%d\t%.6e\n" % (y, v))


def run(cfg: SimConfig, write_files: bool = True, log: Callable[[str], None] = print,
        lattice: Optional[Lattice] = None) -> RunResult:
    lat = lattice or Lattice.from_config(cfg)
    folder = os.path.join(cfg.path_save, cfg.id_sim)
    if write_files:
        os.makedirs(folder, exist_ok=True)
    result = RunResult(steps=0, seconds=0.0, mlups=0.0)
    nodes = cfg.NX * cfg.NY * cfg.NZ
    if cfg.load_checkpoint:
        step = lat.load_checkpoint(cfg.load_checkpoint, cfg.id_sim)
    else:
        lat.init()
        step = cfg.ini_step
    first_step = step
    avg_uz = [0.0] * cfg.NY
    lat.sync()
    t_start = time.perf_counter()
    t_sync, last_sync_step = t_start, step
    while step < cfg.steps:
        target = _next_sync(step, cfg)
        aux = target - cfg.ini_step
        save = bool(cfg.macr_save) and aux != 0 and aux % cfg.macr_save == 0
        rep = bool(cfg.data_report) and aux != 0 and aux % cfg.data_report == 0
        ckpt = bool(cfg.checkpoint_save) and aux != 0 and aux % cfg.checkpoint_save == 0
        last = target + 1 >= cfg.steps
        lat.step(target - step + 1, save_last=save or rep or last)
        step = target
        if save or rep:
            lat.sync()
            now = time.perf_counter()
            mlups = nodes * (step - last_sync_step) / 1e6 / max(now - t_sync, 1e-12)
            t_sync, last_sync_step = now, step
            residual, mean_rho = lat.reduce()
            if save and write_files:
                lat.save_macr_bin(cfg.path_save, cfg.id_sim, step)
            if rep:
                entry = {"step": step, "residual": residual, "avg_rho": mean_rho, "mlups": mlups}
                if cfg.data_save:
                    uz = lat.macroscopics()["uz"]
                    plane = uz.sum(axis=(0, 2))
                    avg_uz = [(a + p) / (cfg.NX * cfg.NZ) for a, p in zip(avg_uz, plane)]
                    entry["avg_uz_mid"] = avg_uz[cfg.NY // 2]
                    if write_files:
                        save_avg_uz(cfg.path_save, cfg.id_sim, step, avg_uz)
                result.reports.append(entry)
                log(f"step {step}: residual {residual:.4e}, avg density {mean_rho:.4e}, MLUPS {mlups:.1f}")
                if not math.isfinite(mean_rho):
                    result.stopped = "non-finite density"
                    log("Stopping: non-finite density")
                    break
                if cfg.data_stop and (residual < cfg.resid_max or mean_rho < 0):
                    result.stopped = "residual" if residual < cfg.resid_max else "negative density"
                    log(f"Stopping because of LBM ({result.stopped})")
                    break
        if ckpt and write_files:
            lat.save_checkpoint(cfg.path_save, cfg.id_sim, step)
        step += 1
    lat.sync()
    elapsed = time.perf_counter() - t_start
    total = step - first_step
    result.steps = total
    result.seconds = elapsed
    result.mlups = nodes * total / 1e6 / max(elapsed, 1e-12)
    if write_files:
        if cfg.checkpoint_save:
            lat.save_checkpoint(cfg.path_save, cfg.id_sim, step)
        lat.save_macr_bin(cfg.path_save, cfg.id_sim, step)
        lat.save_info(cfg.path_save, cfg.id_sim, total, result.mlups, elapsed)
        if cfg.data_report and cfg.data_save:
            save_avg_uz(cfg.path_save, cfg.id_sim, step, avg_uz)
    if lattice is None:
        lat.close()
    return result
