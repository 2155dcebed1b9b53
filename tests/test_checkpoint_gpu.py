"""Checkpoint folder in the reference's layout (vlbm_save_checkpoint /
vlbm_load_checkpoint; src/CUDA/simCheckpoint.cu:216-318) on the GPU.

* The files are parsed here with numpy from the reference's declarations
  (MEM_SIZE_POP, MEM_SIZE_IBM_SCALAR, MEM_SIZE_SCALAR in src/CUDA/var.h:208-213;
  the ParticleCenter struct in src/CUDA/IBM/structs/particleCenter.h:17-118)
  and compared with the oracle run on the same inputs: FP64 within 1e-10.
* The oracle continues from the GPU's checkpoint file and stays within 1e-10
  of the GPU continuing in memory (both start from the same bytes).
* A lattice loaded from a checkpoint continues the interrupted run: FP64
  within 1e-12 (the file holds f = (f - w) + w, one rounding per value),
  FP32 within 1e-4 on u (the file holds f in float, the reference's own
  precision); IBM: FP64 particles within 1e-10, FP32 lattices (whose file
  holds the particle records in float, as the reference's FP32 build) within
  1e-4; collision histories identical, since the step counter and so the
  stage labels resume from the stored step.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from checkpoint_layout import particle_center_dtype
from helpers import base_cfg, rel_l2, seeded_velocity, weights

pytestmark = pytest.mark.gpu

F = po.PARTICLE_FIELDS


def announce(name):
    print(f"\n[gpu-case] {name}", flush=True)


def read(path, dtype, count=None):
    a = np.fromfile(path, dtype=dtype)
    if count is not None:
        assert a.size == count, (path, a.size, count)
    return a


def lattice_files(folder, cfg, border=0):
    """Reference checkpoint arrays: pop (Q, NZ+1, NY, NX), bordered macroscopics."""
    t = np.float64 if cfg.precision == "double" else np.float32
    q = 19 if cfg.vel_set == "D3Q19" else 27
    nx, ny, nz = cfg.NX, cfg.NY, cfg.NZ
    pre = os.path.join(folder, "checkpoint", "0_")
    out = {"step": int(read(pre + "curr_step.bin", np.int32, 1)[0]),
           "pop": read(pre + "pop.bin", t, q * (nz + 1) * ny * nx).reshape(q, nz + 1, ny, nx),
           "popAux": read(pre + "popAux.bin", t, q * (nz + 1) * ny * nx).reshape(q, nz + 1, ny, nx)}
    nb = nx * ny * (nz + 2 * border)
    for name, fn in (("rho", "rho.bin"), ("ux", "u/x.bin"), ("uy", "u/y.bin"), ("uz", "u/z.bin")):
        a = read(pre + fn, t, nb).reshape(nz + 2 * border, ny, nx)
        assert not a[:border].any() and not a[nz + border:].any()
        out[name] = a[border:nz + border]
    if cfg.non_newtonian_model:
        out["omega"] = read(pre + "omega.bin", t, nx * ny * nz).reshape(nz, ny, nx)
    else:
        assert not os.path.exists(pre + "omega.bin")
    return out


def noisy(cfg, seed=11):
    shape = (cfg.NZ, cfg.NY, cfg.NX)
    ux, uy, uz = seeded_velocity(shape, (0.002, 0.0, 0.001), 1e-3, seed)
    return np.ones(shape), ux, uy, uz


CASES = [
    ("plates_bingham_q19", dict(NX=6, NY=24, NZ=8, tau=0.9, FX=1e-5, builder="parallelPlatesBounceBack_Y",
                                non_newtonian_model="BINGHAM", yield_stress=4e-5)),
    ("duct_zouhe_q19", dict(NX=10, NY=9, NZ=12, tau=0.8, builder="squareDuctZouHe",
                            uz_bc=[0, 0.01, 0, 0, 0, 0, 0, 0])),
    ("plates_power_law_q27", dict(NX=4, NY=16, NZ=5, tau=0.7, FX=5e-6, vel_set="D3Q27",
                                  builder="parallelPlatesBounceBack_Y", non_newtonian_model="POWERLAW",
                                  n_index=0.6, k_consistency=0.02, tau_clamp=[0.51, 2.0])),
]
# lattices with nodes that never collide: the file holds what the reference's
# push streaming leaves there, and the rheometer's extra plane feeds its
# unused planes every step
CASES_UNUSED = [
    ("rheometer_noisy", dict(NX=18, NY=18, NZ=6, tau=0.7, builder="parallelPlatesRheometer", outer_rotation=0.002)),
    ("circular_duct_ibb", dict(NX=14, NY=14, NZ=5, tau=0.8, FZ=1e-5, builder="circularDuctInterpBounceBack")),
]


@pytest.mark.parametrize("name,kw", CASES + CASES_UNUSED, ids=[c[0] for c in CASES + CASES_UNUSED])
def test_checkpoint_files_match_oracle_and_oracle_continues(name, kw, tmp_path):
    """pop.bin equals the oracle's pop buffer as a whole: every node (also
    those that never collide) and the extra plane z = NZ. The oracle then
    loads pop.bin and popAux.bin into its two buffers, as the reference's
    checkpoint load does, and continues like the GPU run in memory."""
    import vlbm
    announce(f"checkpoint files vs oracle: {name}")
    cfg = base_cfg(**kw)
    q = 19 if cfg.vel_set == "D3Q19" else 27
    init = noisy(cfg)
    lat = vlbm.Lattice(cfg.to_desc())
    orc = po.Oracle(po.desc_from(cfg.to_desc()), 8)
    lat.init(*init)
    orc.init(*init)
    k = 40
    lat.step(k, save_last=True)
    orc.step(k, save_last=True)
    lat.save_checkpoint(str(tmp_path), "007")
    files = lattice_files(str(tmp_path / "007"), cfg)
    assert files["step"] == k
    w = weights(q)[:, None, None, None]
    op = orc.get_pops()
    err = rel_l2(files["pop"] - w, op - w)
    if err > 1e-10:
        d = np.abs(files["pop"] - op)
        i, z, y, x = np.unravel_index(np.argmax(d), d.shape)
        print(f"pop.bin rel-L2 {err:.3e}; worst pop {i} at (x,y,z)=({x},{y},{z}): file {files['pop'][i, z, y, x]:.6e} "
              f"oracle {op[i, z, y, x]:.6e}", flush=True)
    assert err <= 1e-10
    # the extra plane carries pushed populations (periodic z) or the rest state
    assert rel_l2(files["pop"][:, -1] - w[:, 0], op[:, -1] - w[:, 0]) <= 1e-10
    if cfg.builder != "parallelPlatesRheometer":
        assert np.array_equal(files["popAux"], files["pop"])
    om = orc.macroscopics()
    for f in ("rho", "ux", "uy", "uz"):
        assert rel_l2(files[f], om[f]) <= 1e-10, f
    if cfg.non_newtonian_model:
        assert rel_l2(files["omega"], om["omega"]) <= 1e-10
    # the oracle starts from the file's bytes; the GPU continues in memory
    orc2 = po.Oracle(po.desc_from(cfg.to_desc()), 8)
    orc2.init()
    orc2.set_pops(files["pop"], files["popAux"])
    if cfg.non_newtonian_model == "POWERLAW":
        orc2.set_omega(files["omega"])
    n = 30
    orc2.step(n, save_last=True)
    lat.step(n, save_last=True)
    lat.sync()
    gp = lat.get_pops()
    assert rel_l2(gp - w, orc2.get_pops() - w) <= 1e-10
    # macroscopics of the colliding nodes (the others keep their initial values,
    # which differ between the two histories)
    used = ((vlbm.build_map(lat.desc) >> 21) & 1).astype(bool)
    gm, om2 = lat.macroscopics(), orc2.macroscopics()
    assert rel_l2(np.stack([gm[k][used] for k in ("ux", "uy", "uz")]),
                  np.stack([om2[k][used] for k in ("ux", "uy", "uz")])) <= 1e-10
    lat.close()


@pytest.mark.parametrize("layout", ["populations", "moments"])
@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 1e-4)])
@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_loaded_lattice_continues_the_run(name, kw, precision, tol, layout, tmp_path):
    import vlbm
    announce(f"checkpoint resume: {name} {precision} [{layout}]")
    cfg = base_cfg(precision=precision, layout=layout, **kw)
    q = 19 if cfg.vel_set == "D3Q19" else 27
    a = vlbm.Lattice(cfg.to_desc())
    a.init(*noisy(cfg))
    a.step(25)
    a.save_checkpoint(str(tmp_path), "000")
    a.step(20, save_last=True)
    a.sync()
    b = vlbm.Lattice(cfg.to_desc())
    b.init()
    assert b.load_checkpoint(str(tmp_path), "000") == 25
    assert b.steps_done == 25
    b.step(20, save_last=True)
    b.sync()
    assert b.steps_done == a.steps_done == 45
    w = weights(q)[:, None, None, None]
    pa, pb = a.get_pops(), b.get_pops()
    ma, mb = a.macroscopics(omega=bool(cfg.non_newtonian_model)), b.macroscopics(omega=bool(cfg.non_newtonian_model))
    ua = np.stack([ma["ux"], ma["uy"], ma["uz"]])
    ub = np.stack([mb["ux"], mb["uy"], mb["uz"]])
    print(f"resume: pops {rel_l2(pb[:, :-1] - w, pa[:, :-1] - w):.2e}  u {rel_l2(ub, ua):.2e}")
    assert rel_l2(ub, ua) <= tol
    assert rel_l2(mb["rho"], ma["rho"]) <= tol
    if precision == "double":
        assert rel_l2(pb[:, :-1] - w, pa[:, :-1] - w) <= tol
    if cfg.non_newtonian_model:
        assert rel_l2(mb["omega"], ma["omega"]) <= tol
    a.close()
    b.close()


def test_load_rejects_a_checkpoint_of_another_lattice(tmp_path):
    import vlbm
    announce("checkpoint size check")
    cfg = base_cfg(NX=6, NY=8, NZ=5, builder="parallelPlatesBounceBack_Y")
    a = vlbm.Lattice(cfg.to_desc())
    a.init()
    a.save_checkpoint(str(tmp_path), "000")
    b = vlbm.Lattice(base_cfg(NX=6, NY=8, NZ=6, builder="parallelPlatesBounceBack_Y").to_desc())
    b.init()
    with pytest.raises(RuntimeError, match="bytes"):
        b.load_checkpoint(str(tmp_path), "000")
    with pytest.raises(RuntimeError, match="cannot open"):
        b.load_checkpoint(str(tmp_path), "001")
    a.close()
    b.close()


# ---------------------------------------------------------------------------
# immersed boundary
# ---------------------------------------------------------------------------
def ibm_case(precision="double", layout="auto"):
    import vlbm
    from vlbm import _capi
    cfg = base_cfg(NX=32, NY=24, NZ=32, builder="parallelPlatesBounceBack_Y", tau=0.6, precision=precision,
                   FX=1e-5, layout=layout)
    idesc = _capi.default_ibm_desc(max_iterations=2, gy=-2e-5, young=5.0)
    spheres = [([10.0, 9.0, 12.0], 4.0, (0.001, 0, 0), (0, 0, 0)),
               ([17.5, 9.3, 12.4], 4.0, (0, 0, 0), (0, 0.002, 0)),
               ([24.2, 3.8, 21.7], 4.0, (0, -0.002, 0.001), (0, 0, 0))]

    def make():
        lat = vlbm.Lattice(cfg.to_desc())
        lat.init()
        lat.ibm_enable(idesc)
        for c, r, v, wv in spheres:
            xyz, S = po.sphere_polar(2 * r, c)
            lat.ibm_add_particle(c, v, wv, r, 2.0, True, xyz, S)
        return lat

    def make_oracle():
        o = po.Oracle(po.desc_from(cfg.to_desc()), 8)
        o.init()
        o.ibm_enable(po.ibm_desc_from(idesc))
        for c, r, v, wv in spheres:
            xyz, S = po.sphere_polar(2 * r, c)
            o.ibm_add_particle(c, v, wv, r, 2.0, True, xyz, S)
        return o

    return cfg, make, make_oracle


def same_or_both_nan(a, b, tol):
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(b)
    if m.any():
        assert np.abs(a[m] - b[m]).max() <= tol * max(np.abs(b[m]).max(), 1.0)


def test_ibm_checkpoint_files_match_oracle(tmp_path):
    announce("IBM checkpoint files vs oracle")
    cfg, make, make_oracle = ibm_case()
    lat, o = make(), make_oracle()
    n = 10
    lat.step(n, save_last=True)
    o.ibm_run(n, 0)
    lat.save_checkpoint(str(tmp_path), "000")
    folder = str(tmp_path / "000")
    files = lattice_files(folder, cfg, border=3)
    assert files["step"] == n
    om = o.macroscopics()
    for f in ("rho", "ux", "uy", "uz"):
        assert rel_l2(files[f], om[f]) <= 1e-10, f
    pre = os.path.join(folder, "checkpoint", "0_")
    pc = read(pre + "IBM_particles_centers.bin", particle_center_dtype(np.float64), 3)
    rec = o.ibm_particles()
    partner, tang, last = o.ibm_collision_state()
    axes = o.ibm_axes()
    for k in range(3):
        p = pc[k]
        for name in ("pos", "pos_old", "vel", "vel_old", "w", "w_avg", "w_old", "w_pos", "f", "f_old", "M",
                     "M_old", "I", "dP", "dL"):
            b = rec[k, F[name]]
            assert np.abs(p[name] - b).max() <= 1e-10 * max(np.abs(rec[:, F[name]]).max(), 1e-300), name
        assert np.array_equal(p["q_pos"], [1, 0, 0, 0]) and np.array_equal(p["q_pos_old"], [1, 0, 0, 0])
        assert p["S"] == 4.0 * np.pi * 16.0 and p["radius"] == 4.0 and p["density"] == 2.0
        assert p["volume"] == 4.0 ** 3 * 4 * np.pi / 3 and p["movable"]
        col = p["collision"]
        assert col["shape"] == 0
        assert np.array_equal(col["partner"], partner[k]) and np.array_equal(col["last"], last[k])
        same_or_both_nan(col["tang"], tang[k], 1e-10)
        same_or_both_nan(col["semiAxis"], axes[k], 1e-10)
    assert (last >= 0).sum() >= 2
    pos, nf, ndf = o.ibm_nodes()
    nn = pos.shape[1]
    idx = read(pre + "IBM_nodes_centers_idx.bin", np.uint32, nn)
    counts = np.bincount(idx, minlength=3)
    assert np.array_equal(idx, np.repeat(np.arange(3), counts))
    for grp, ref in (("IBM_nodes_pos", pos), ("IBM_nodes_f", nf), ("IBM_nodes_deltaF", ndf)):
        for a, ax in enumerate("xyz"):
            v = read(os.path.join(pre + grp, ax + ".bin"), np.float64, nn)
            assert np.abs(v - ref[a]).max() <= 1e-10 * max(np.abs(ref).max(), 1e-300), grp
    for grp in ("IBM_nodes_vel", "IBM_nodes_vel_old"):
        for ax in "xyz":
            assert not read(os.path.join(pre + grp, ax + ".bin"), np.float64, nn).any()
    _, S = po.sphere_polar(8.0, [10.0, 9.0, 12.0])
    assert np.array_equal(read(pre + "IBM_nodes_S.bin", np.float64, nn)[:counts[0]], S)
    fo = o.ibm_force()
    nz, ny, nx = cfg.NZ, cfg.NY, cfg.NX
    for a, ax in enumerate("xyz"):
        v = read(os.path.join(pre + "f", ax + ".bin"), np.float64, nx * ny * (nz + 6)).reshape(nz + 6, ny, nx)
        assert not v[:3].any() and not v[nz + 3:].any()
        assert np.abs(v[3:nz + 3] - fo[a]).max() <= 1e-10 * np.abs(fo).max()
    lat.close()


@pytest.mark.parametrize("layout", ["populations", "moments"])
@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-4)])
def test_ibm_loaded_lattice_continues_the_run(precision, tol, layout, tmp_path):
    announce(f"IBM checkpoint resume {precision} [{layout}]")
    cfg, make, make_oracle = ibm_case(precision, layout)
    a = make()
    a.step(8)
    a.save_checkpoint(str(tmp_path), "000")
    a.step(6, save_last=True)
    a.sync()
    b = make()
    assert b.load_checkpoint(str(tmp_path), "000") == 8
    b.step(6, save_last=True)
    b.sync()
    ra, rb = a.ibm_particles(), b.ibm_particles()
    for name in ("pos", "vel", "w", "f", "M"):
        e = np.abs(rb[:, F[name]] - ra[:, F[name]]).max() / max(np.abs(ra[:, F[name]]).max(), 1e-300)
        print(name, f"{e:.2e}")
        assert e <= tol, name
    pa, ta, la = a.ibm_collision_state()
    pb, tb, lb = b.ibm_collision_state()
    assert np.array_equal(pa, pb) and np.array_equal(la, lb)
    ma, mb = a.macroscopics(), b.macroscopics()
    assert rel_l2(np.stack([mb["ux"], mb["uy"], mb["uz"]]), np.stack([ma["ux"], ma["uy"], ma["uz"]])) <= tol * 10
    a.close()
    b.close()
