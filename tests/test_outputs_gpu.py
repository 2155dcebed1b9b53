"""Output consumers on the GPU path: a run of the driver writes the folder the
reference's post-processing reads (fileTreat.py), the data-save CSV of the
plane-averaged u_z (treatData.cpp:57-78, 476-497), the VTK export of a
saved step, and a run resumed from a checkpoint writes the same final file
as the uninterrupted run."""
import numpy as np
import pytest

import pyoracle as po
from helpers import base_cfg, rel_l2

pytestmark = pytest.mark.gpu


def announce(name):
    print(f"\n[gpu-case] {name}", flush=True)


def channel(tmp_path, **kw):
    args = dict(precision="double", NX=4, NY=12, NZ=3, tau=0.8, FZ=1e-5, builder="parallelPlatesBounceBack_Y",
                steps=300, macr_save=100, data_report=100, data_save=True, path_save=str(tmp_path), sim_id=9)
    args.update(kw)
    return base_cfg(**args)


def test_runner_folder_reads_back_and_avg_uz(tmp_path):
    announce("runner folder, avgUz csv, VTK")
    from vlbm import post
    from vlbm.runner import run
    cfg = channel(tmp_path)
    res = run(cfg, log=lambda s: None)
    assert [r["step"] for r in res.reports] == [100, 200]
    sim = post.SimFolder(str(tmp_path / "009"))
    info = sim.info()
    assert info["Prc"] == "double" and (info["NX"], info["NY"], info["NZ"]) == (4, 12, 3)
    assert sim.steps() == [100, 200, 300]
    # the saved fields are the oracle's moments of the state entering each labelled step
    orc = po.Oracle(po.desc_from(cfg.to_desc()), 8)
    orc.init()
    avg = np.zeros(cfg.NY)
    done = 0
    for step in (100, 200):
        orc.step(step + 1 - done, save_last=True)
        done = step + 1
        om = orc.macroscopics()
        m = sim.macrs_from_step(step)
        assert rel_l2(np.swapaxes(m["uz"], 0, 2), om["uz"]) <= 1e-12
        # avg <- (avg + sum_xz uz) / (NX NZ): the accumulator is not reset
        avg = (avg + om["uz"].sum(axis=(0, 2))) / (cfg.NX * cfg.NZ)
        rows = np.loadtxt(tmp_path / "009" / f"009_avgUz{step:06d}.csv", skiprows=1)
        assert np.array_equal(rows[:, 0], np.arange(cfg.NY))
        assert np.allclose(rows[:, 1], avg, rtol=1e-6, atol=0)
    # the final file after the loop repeats the last report, labelled with the loop counter
    final = np.loadtxt(tmp_path / "009" / "009_avgUz000300.csv", skiprows=1)
    assert np.allclose(final[:, 1], avg, rtol=1e-6, atol=0)
    assert open(tmp_path / "009" / "009_avgUz000300.csv").readline() == "y\tavg uz\n"
    out = post.export_folder_vtk(str(tmp_path / "009"))
    back = post.read_vtr(out[-1])
    raw = np.fromfile(tmp_path / "009" / "009_uz000300.bin", dtype="d")
    assert np.array_equal(back["uz"], raw)


def test_resumed_run_writes_the_same_final_file(tmp_path):
    announce("runner checkpoint resume")
    from vlbm.runner import run
    full = channel(tmp_path / "a", steps=400, macr_save=0, data_report=0, data_save=False, checkpoint_save=150)
    (tmp_path / "a").mkdir()
    run(full, log=lambda s: None)
    # the in-loop checkpoint at label 300 holds the state after 301 steps,
    # the final one (label 400) the state after 400: load the final one into
    # a run with more steps and compare with a longer uninterrupted run
    longer = channel(tmp_path / "b", steps=450, macr_save=0, data_report=0, data_save=False)
    (tmp_path / "b").mkdir()
    run(longer, log=lambda s: None)
    resumed = channel(tmp_path / "a", steps=450, macr_save=0, data_report=0, data_save=False,
                      load_checkpoint=str(tmp_path / "a"))
    res = run(resumed, log=lambda s: None)
    assert res.steps == 50
    a = np.fromfile(tmp_path / "a" / "009" / "009_uz000450.bin", dtype="d")
    b = np.fromfile(tmp_path / "b" / "009" / "009_uz000450.bin", dtype="d")
    assert rel_l2(a, b) <= 1e-12
    step = np.fromfile(tmp_path / "a" / "009" / "checkpoint" / "0_curr_step.bin", dtype=np.int32)
    assert step[0] == 400


def test_rheometer_checkpoint_resumes_exactly(tmp_path):
    """The rheometer's plates read their own post-collision populations of the
    step before (the reference keeps them in popAux, which its checkpoint
    saves): a run resumed from a checkpoint continues like the uninterrupted
    one."""
    announce("rheometer checkpoint resume")
    import vlbm
    from helpers import base_cfg, seeded_velocity
    cfg = base_cfg(NX=18, NY=18, NZ=6, tau=0.7, builder="parallelPlatesRheometer", outer_rotation=0.003)
    shape = (6, 18, 18)
    init = (np.ones(shape),) + seeded_velocity(shape, (0, 0, 0), 1e-3, seed=4)
    a = vlbm.Lattice(cfg.to_desc())
    a.init(*init)
    a.step(70)
    a.save_checkpoint(str(tmp_path), "rheo")
    a.step(50)
    b = vlbm.Lattice(cfg.to_desc())
    b.init()
    assert b.load_checkpoint(str(tmp_path), "rheo") == 70
    # a checkpoint saved right after the load carries the same popAux
    (tmp_path / "again").mkdir()
    b.save_checkpoint(str(tmp_path / "again"), "rheo", 70)
    for f in ("pop", "popAux"):
        first = np.fromfile(tmp_path / "rheo" / "checkpoint" / f"0_{f}.bin", dtype="d")
        again = np.fromfile(tmp_path / "again" / "rheo" / "checkpoint" / f"0_{f}.bin", dtype="d")
        assert np.max(np.abs(first - again)) < 1e-15, f     # f - w + w round trip
    b.step(50)
    used = ((vlbm.build_map(a.desc) >> 21) & 1).astype(bool)
    pa, pb = a.get_pops()[:, :-1], b.get_pops()[:, :-1]
    err = np.abs(pa - pb)[:, used].max()
    print("max |resumed - uninterrupted|", err, flush=True)
    assert err <= 1e-15
    a.close()
    b.close()
