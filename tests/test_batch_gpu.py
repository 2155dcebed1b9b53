"""Batched serving engine (vlbm_batch_*) on the GPU.

Each job's arithmetic is independent of the other resident jobs, of the
launch shape and of graph replay, so results are compared bit for bit:
alone vs in a mix, captured sequences vs plain launches, batched vs the
single-lattice path. Admission under an HBM budget, the stop test and the
autotuner cache are checked on small jobs.
"""
import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu


def spec(nx, ny, nz, bn, tau, fx, arrival=0.0):
    from vlbm.serve import JobSpec, tau_y_for_bn
    return JobSpec(nx, ny, nz, bn, tau, fx, tau_y_for_bn(bn, ny / 2.0, fx, (tau - 0.5) / 3.0), arrival)


def run_batch(specs, submit_at=None, max_steps=60, resid=0.0, **kw):
    """Runs specs through one batch; submit_at[k] = advance call before which job k is submitted."""
    from vlbm.serve import Batch
    kw.setdefault("check_every", 20)
    kw.setdefault("steps_per_launch", 10)
    kw.setdefault("block", 256)
    kw.setdefault("grid", 64)
    kw.setdefault("hbm_budget", 1 << 30)
    b = Batch(**kw)
    submit_at = submit_at or [0] * len(specs)
    ids = {}
    out = {}
    call = 0
    while len(out) < len(specs):
        for k, s in enumerate(specs):
            if submit_at[k] == call and k not in ids.values():
                ids[b.submit(s.desc(), resid, max_steps)] = k
        b.advance()
        call += 1
        for jid, k in ids.items():
            if k in out:
                continue
            inf = b.info(jid)
            if inf["state"] in ("done", "max_steps", "failed"):
                out[k] = (inf, b.result(jid, (specs[k].nz, specs[k].ny, specs[k].nx)))
        assert call < 5000
    cfg = b.config()
    b.close()
    return out, cfg


def test_alone_and_in_a_mix_are_identical():
    a = spec(24, 20, 12, 1.5, 0.8, 5e-6)
    others = [spec(40, 32, 16, 0.0, 0.6, 1e-5), spec(17, 33, 9, 3.0, 1.2, 2e-6)]
    print("case: batch alone vs mix", flush=True)
    alone, _ = run_batch([a])
    mix, _ = run_batch([others[0], a, others[1]], submit_at=[0, 2, 1])
    ia, ra = alone[0]
    im, rm = mix[1]
    assert ia["steps"] == im["steps"] and ia["residual"] == im["residual"]
    for f in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(ra[f], rm[f]), f


def test_graph_replay_matches_plain_launches():
    jobs = [spec(24, 20, 12, 1.5, 0.8, 5e-6), spec(30, 18, 10, 0.5, 0.7, 8e-6)]
    print("case: graph vs eager", flush=True)
    g, _ = run_batch(jobs, use_graph=1, max_steps=100)
    e, _ = run_batch(jobs, use_graph=0, max_steps=100)
    for k in range(2):
        assert g[k][0]["steps"] == e[k][0]["steps"] == 101
        for f in ("rho", "ux", "uy", "uz"):
            assert np.array_equal(g[k][1][f], e[k][1][f])


def test_batched_matches_single_lattice_path():
    import vlbm
    s = spec(28, 22, 14, 2.0, 0.9, 6e-6)
    print("case: batch vs plain", flush=True)
    out, _ = run_batch([s], max_steps=60, block=128, grid=32)
    inf, res = out[0]
    lat = vlbm.Lattice(s.desc())
    lat.init()
    lat.step(inf["steps"])
    lat.compute_macroscopics()
    m = lat.macroscopics()
    lat.close()
    for f in ("rho", "ux", "uy", "uz"):
        d = np.abs(res[f] - m[f]).max() / max(np.abs(m[f]).max(), 1e-300)
        print(f, "max rel diff", d)
        assert d <= 1e-6


def test_admission_under_budget_and_stop_test():
    from vlbm.serve import Batch
    jobs = [spec(32, 24, 16, 0.0, 1.0, 1e-5) for _ in range(4)]
    probe = Batch(check_every=20, steps_per_launch=10, block=256, grid=64, hbm_budget=1 << 30)
    per_job = probe.job_bytes(jobs[0].desc())
    probe.close()
    b = Batch(check_every=20, steps_per_launch=10, block=256, grid=64, hbm_budget=2 * per_job + 4096)
    ids = [b.submit(j.desc(), 1e-3, 5000) for j in jobs]
    running, pending, _ = b.advance()
    assert running == 2 and pending == 2
    calls = 0
    while any(b.info(i)["state"] in ("pending", "running") for i in ids):
        r, p, _ = b.advance()
        assert r <= 2
        calls += 1
        assert calls < 500
    infos = [b.info(i) for i in ids]
    assert all(i["state"] == "done" for i in infos), infos
    assert all(i["residual"] < 1e-3 for i in infos)
    # identical jobs, identical histories
    assert len({i["steps"] for i in infos}) == 1
    # the last two started only after a slot freed
    assert infos[2]["t_start"] >= min(infos[0]["t_end"], infos[1]["t_end"])
    b.close()


def test_autotuner_choice_is_cached(tmp_path):
    cache = str(tmp_path / "tune.txt")
    jobs = [spec(64, 48, 32, 1.0, 0.8, 5e-6), spec(48, 40, 24, 0.0, 0.7, 5e-6)]
    print("case: autotune", flush=True)
    _, cfg1 = run_batch(jobs, steps_per_launch=0, block=0, grid=0, tune_cache=cache, max_steps=100)
    assert cfg1["block"] in (128, 256) and cfg1["grid"] > 0 and 20 % cfg1["steps_per_launch"] == 0
    text = open(cache).read()
    assert "|" in text
    _, cfg2 = run_batch(jobs, steps_per_launch=0, block=0, grid=0, tune_cache=cache, max_steps=100)
    assert cfg2 == cfg1
    assert open(cache).read() == text     # found, not tuned again


def test_autotuned_batch_keeps_each_job_s_steps_and_results():
    """The autotuner times launch shapes on real chunks of the resident jobs
    (14 candidates of block 128 and 256, two chunks of 10 steps each). The
    residual checks run after every tuning chunk and the unit table follows
    the block size, so each job stops at the same step with the same fields as
    with a fixed launch shape: with a cap of 60 the jobs finish during the
    tuning, with 400 they run through both block sizes."""
    jobs = [spec(24, 20, 12, 1.0, 0.8, 5e-6), spec(30, 18, 10, 0.0, 0.7, 6e-6), spec(16, 22, 14, 3.0, 1.1, 8e-6)]
    for cap in (60, 400):
        print(f"case: autotuned vs fixed launch shape, cap {cap}", flush=True)
        fixed, _ = run_batch(jobs, max_steps=cap)
        tuned, cfg = run_batch(jobs, max_steps=cap, block=0, grid=0)
        assert cfg["block"] in (128, 256) and cfg["grid"] > 0
        for k in range(len(jobs)):
            assert tuned[k][0]["steps"] == fixed[k][0]["steps"] == cap + 1, (k, tuned[k][0], fixed[k][0])
            for f in ("rho", "ux", "uy", "uz"):
                assert np.array_equal(tuned[k][1][f], fixed[k][1][f]), (k, cap, f)


def test_converged_channel_matches_analytic_profile():
    """Newtonian and Bingham channels run to a residual of 1e-7: the x/z
    averaged u_x(y) follows the analytic profile (halfway walls)."""
    from vlbm.serve import profile_error
    jobs = [spec(8, 16, 8, 0.0, 1.0, 1e-5), spec(8, 16, 8, 2.0, 1.0, 1e-5)]
    print("case: converged channels", flush=True)
    out, _ = run_batch(jobs, max_steps=40000, resid=1e-7, check_every=200, steps_per_launch=50)
    for k, s in enumerate(jobs):
        inf, res = out[k]
        assert inf["state"] == "done", inf
        err = profile_error(s, res["ux"])
        print("profile error", k, err, inf["steps"])
        assert err < 0.03


def test_node_budget_limits_residency_and_keeps_results():
    """max_resident_nodes admits a pending job only while the resident
    lattices stay within the node budget; results are the same as with every
    job resident at once (each job's arithmetic is independent of the mix)."""
    from vlbm.serve import Batch
    jobs = [spec(24, 20, 12, 1.0, 0.8, 5e-6), spec(24, 20, 12, 2.0, 0.9, 4e-6), spec(30, 18, 10, 0.0, 0.7, 6e-6),
            spec(16, 22, 14, 3.0, 1.1, 8e-6)]
    n = [s.nx * s.ny * s.nz for s in jobs]
    print("case: node budget", flush=True)
    free, _ = run_batch(jobs, max_steps=80)
    b = Batch(check_every=20, steps_per_launch=10, block=256, grid=64, hbm_budget=1 << 30,
              max_resident_nodes=n[0] + n[1])
    ids = [b.submit(j.desc(), 0.0, 80) for j in jobs]
    running, pending, _ = b.advance()
    assert running == 2 and pending == 2
    calls = 0
    while any(b.info(i)["state"] in ("pending", "running") for i in ids):
        r, p, _ = b.advance()
        assert r <= 2
        calls += 1
        assert calls < 500
    for k, i in enumerate(ids):
        res = b.result(i, (jobs[k].nz, jobs[k].ny, jobs[k].nx))
        assert b.info(i)["steps"] == free[k][0]["steps"]
        for f in ("rho", "ux", "uy", "uz"):
            assert np.array_equal(res[f], free[k][1][f]), (k, f)
    b.close()


def _drive(b, submits, shapes):
    """Submits (desc, resid, max_steps) in order, advances until every job has
    finished; returns {job: (info, result)}."""
    ids = [b.submit(d, r, m) for d, r, m in submits]
    out = {}
    calls = 0
    while len(out) < len(ids):
        b.advance()
        calls += 1
        assert calls < 5000
        for k, jid in enumerate(ids):
            if k not in out:
                inf = b.info(jid)
                if inf["state"] in ("done", "max_steps", "failed"):
                    out[k] = (inf, b.result(jid, shapes[k]))
    return out


def test_job_in_an_autotuned_batch_matches_the_oracle():
    """One job of a three-job FP64 batch whose launch shape the autotuner picks
    against the oracle (the reference restated on the CPU): the batch's
    result after S steps holds the moments of the state the next step
    collides, i.e. what the oracle saves during step S + 1. Bar: rel-L2 of u
    and of rho - 1 (against the larger of the rho - 1 and u norms) <= 1e-10."""
    import pyoracle as po
    from vlbm import _capi
    from vlbm.serve import Batch
    print("case: batched job vs oracle", flush=True)
    specs = [spec(24, 20, 12, 1.5, 0.8, 5e-6), spec(30, 18, 10, 0.0, 0.7, 6e-6), spec(16, 22, 14, 3.0, 1.1, 8e-6)]
    b = Batch(check_every=20, steps_per_launch=0, block=0, grid=0, hbm_budget=1 << 30, precision=_capi.FP64)
    out = _drive(b, [(s.desc(_capi.FP64), 0.0, 200) for s in specs], [(s.nz, s.ny, s.nx) for s in specs])
    cfg = b.config()
    b.close()
    assert cfg["block"] in (128, 256) and cfg["grid"] > 0
    for k, s in enumerate(specs):
        inf, res = out[k]
        assert inf["state"] == "max_steps" and inf["steps"] == 201
        orc = po.Oracle(po.desc_from(s.desc(_capi.FP64)), 8)
        orc.init()
        orc.step(inf["steps"])
        orc.step(1, save_last=True)
        om = orc.macroscopics()
        orc.close()
        u_b, u_o = np.stack([res[f] for f in ("ux", "uy", "uz")]), np.stack([om[f] for f in ("ux", "uy", "uz")])
        e_u = rel_l2(u_b, u_o)
        # a periodic channel's density stays uniform (rho - 1 is round-off): rho - 1 is measured
        # against the larger of its own norm and the velocity's
        e_r = float(np.linalg.norm(res["rho"] - om["rho"]) / max(np.linalg.norm(om["rho"] - 1), np.linalg.norm(u_o)))
        print(f"job {k}: u {e_u:.2e}  rho-1 {e_r:.2e}", flush=True)
        assert e_u <= 1e-10 and e_r <= 1e-10


def test_non_finite_or_negative_density_jobs_are_evicted():
    """The reference stops a run whose mean density is negative
    (src/CUDA/treatData.cpp:451-459); the batch retires such a job as failed at
    its next check, and also one whose sums are not finite. Injected on the
    host: a job whose initial density is NaN and one whose initial density is
    -0.5 run beside two good jobs, which finish with the results they have
    alone."""
    import math
    from vlbm.serve import Batch
    print("case: eviction of failed jobs", flush=True)
    good = [spec(24, 20, 12, 1.5, 0.8, 5e-6), spec(16, 22, 14, 3.0, 1.1, 8e-6)]
    bad_nan, bad_neg = good[0].desc(), good[1].desc()
    bad_nan.rho0 = math.nan
    bad_neg.rho0 = -0.5
    shapes = [(s.nz, s.ny, s.nx) for s in good]
    kw = dict(check_every=20, steps_per_launch=10, block=256, grid=64, hbm_budget=1 << 30)
    b = Batch(**kw)
    out = _drive(b, [(good[0].desc(), 0.0, 100), (bad_nan, 0.0, 100), (good[1].desc(), 0.0, 100),
                     (bad_neg, 0.0, 100)], [shapes[0], shapes[0], shapes[1], shapes[1]])
    b.close()
    for k in (1, 3):
        inf = out[k][0]
        assert inf["state"] == "failed" and inf["steps"] == 21, inf     # retired at its first check
    assert math.isnan(out[1][0]["mean_rho"]) and out[3][0]["mean_rho"] < 0
    alone = [_drive(Batch(**kw), [(s.desc(), 0.0, 100)], [shp])[0] for s, shp in zip(good, shapes)]
    for k, a in ((0, alone[0]), (2, alone[1])):
        assert out[k][0]["state"] == a[0]["state"] == "max_steps" and out[k][0]["steps"] == a[0]["steps"] == 101
        for f in ("rho", "ux", "uy", "uz"):
            assert np.array_equal(out[k][1][f], a[1][f]), (k, f)
