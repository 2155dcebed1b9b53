"""Immersed-boundary stage on the GPU (libvlbm.so through the C-ABI) against
the CPU oracle on the same inputs.

FP64 lattices: particle state, node positions, force field and macroscopics
within 1e-10 (relative) after coupled steps; collision histories identical.
FP32 lattices (particles still FP64): fluid rho/u within 1e-4, particle
positions within 1e-6 of the FP64 oracle. The uniform cell list is compared
bit-exactly with the oracle's cell rule on the same positions.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from helpers import base_cfg, rel_l2

pytestmark = pytest.mark.gpu

F = po.PARTICLE_FIELDS


def make_pair(nx, ny, nz, builder, ibm_kw, precision="double", tau=0.6, **cfg_kw):
    import vlbm
    from vlbm import _capi
    cfg = base_cfg(NX=nx, NY=ny, NZ=nz, builder=builder, tau=tau, precision=precision, **cfg_kw)
    desc = cfg.to_desc()
    lat = vlbm.Lattice(desc)
    lat.init()
    idesc = _capi.default_ibm_desc(**ibm_kw)
    lat.ibm_enable(idesc)
    o = po.Oracle(po.desc_from(desc), 8)
    o.init()
    o.ibm_enable(po.ibm_desc_from(idesc))
    return lat, o


def add_both(lat, o, center, r, vel=(0, 0, 0), w=(0, 0, 0), density=2.0, movable=True, xyz=None, S=None):
    if xyz is None:
        xyz, S = po.sphere_polar(2 * r, center)
    lat.ibm_add_particle(center, vel, w, r, density, movable, xyz, S)
    o.ibm_add_particle(center, vel, w, r, density, movable, xyz, S)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def compare_particles(lat, o, tol, fields=("pos", "vel", "w", "f", "M", "I", "w_avg", "w_pos")):
    g, c = lat.ibm_particles(), o.ibm_particles()
    assert g.shape == c.shape
    worst = {}
    for k in fields:
        scale = max(np.abs(c[:, F[k]]).max(), 1e-300)
        worst[k] = float(np.abs(g[:, F[k]] - c[:, F[k]]).max() / scale)
    print("particle rel errors:", {k: f"{v:.2e}" for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= tol, (k, v)


def compare_fluid(lat, o, tol):
    lat.sync()
    m = lat.macroscopics()
    mo = o.macroscopics()
    e_rho = rel_l2(m["rho"], mo["rho"])
    ug = np.stack([m["ux"], m["uy"], m["uz"]])
    uo = np.stack([mo["ux"], mo["uy"], mo["uz"]])
    e_u = rel_l2(ug, uo)
    fg, fo = lat.ibm_force(), o.ibm_force()
    e_f = rel_l2(fg - fo.mean(), fo - fo.mean()) if np.ptp(fo) > 0 else float(np.abs(fg - fo).max())
    print(f"fluid rel-L2: rho {e_rho:.2e}  u {e_u:.2e}  force {e_f:.2e}")
    assert e_rho <= tol and e_u <= tol and e_f <= tol * 10


@pytest.mark.parametrize("layout", ["populations", "moments"])
def test_ibm_fp64_contacts_walls_and_iterations(layout):
    """Two overlapping spheres and one pressed into the y = 0 wall, gravity,
    two IBM iterations per step, periodic x / z: 12 coupled steps."""
    ny = 24
    lat, o = make_pair(32, ny, 32, "parallelPlatesBounceBack_Y",
                       dict(max_iterations=2, gy=-2e-5, young=5.0), FX=1e-5, layout=layout)
    add_both(lat, o, [10.0, 9.0, 12.0], 4.0, vel=(0.001, 0, 0))
    add_both(lat, o, [17.5, 9.3, 12.4], 4.0, w=(0, 0.002, 0))
    add_both(lat, o, [24.2, 3.8, 21.7], 4.0, vel=(0, -0.002, 0.001))
    n = 12
    print("case: fp64 3 spheres, 12 steps", flush=True)
    lat.step(n, save_last=True)
    o.ibm_run(n, 0)
    lat.ibm_status()
    compare_particles(lat, o, 1e-10)
    compare_fluid(lat, o, 1e-10)
    pg, fg, dg = lat.ibm_nodes()
    pc, fc, dc = o.ibm_nodes()
    assert rel(pg, pc) <= 1e-12
    assert rel(fg, fc) <= 1e-10 and rel(dg, dc) <= 1e-10
    partner_g, tang_g, last_g = lat.ibm_collision_state()
    partner_c, tang_c, last_c = o.ibm_collision_state()
    assert np.array_equal(partner_g, partner_c) and np.array_equal(last_g, last_c)
    assert (last_c >= 0).sum() >= 2          # the pair and the wall contact are tracked
    assert rel(tang_g, tang_c) <= 1e-10
    # cell list: bit-exact cell rule on the positions the stage used
    dims, cells = lat.ibm_cells()
    rec = lat.ibm_particles()
    d2, c2 = po.ibm_cells((32, ny, 32), 4.0, rec[:, F["pos_old"]])
    assert np.array_equal(dims, d2) and np.array_equal(cells, c2)


@pytest.mark.parametrize("layout", ["populations", "moments"])
def test_ibm_fp64_spheres_in_circular_duct(layout):
    """Node-map lattice: a circular duct (interpolated bounce-back wall of
    radius 10.5) driven along z, two moving spheres clear of the wall, one
    across the periodic z ends, 2 iterations: 10 coupled steps. The fluid is
    compared on the colliding nodes (the map leaves the others unspecified)."""
    import vlbm
    lat, o = make_pair(24, 24, 24, "circularDuctInterpBounceBack",
                       dict(max_iterations=2, periodic_x=0, periodic_y=0, periodic_z=1), FZ=1e-5, layout=layout)
    assert lat.layout == layout
    add_both(lat, o, [12.0, 12.0, 12.0], 3.0, vel=(0, 0, 0.001), w=(0, 0.001, 0))
    add_both(lat, o, [12.5, 11.0, 3.0], 3.0, vel=(0.001, 0, 0))
    n = 10
    print(f"case: fp64 circular duct 2 spheres, 10 steps [{layout}]", flush=True)
    lat.step(n, save_last=True)
    o.ibm_run(n, 0)
    lat.ibm_status()
    compare_particles(lat, o, 1e-10)
    used = ((vlbm.build_map(lat.desc) >> 21) & 1).astype(bool)
    lat.sync()
    m, mo = lat.macroscopics(), o.macroscopics()
    e_rho = rel_l2(m["rho"][used], mo["rho"][used])
    e_u = rel_l2(np.stack([m[k][used] for k in ("ux", "uy", "uz")]),
                 np.stack([mo[k][used] for k in ("ux", "uy", "uz")]))
    fg, fo = lat.ibm_force()[:, used], o.ibm_force()[:, used]
    e_f = rel_l2(fg - fo.mean(), fo - fo.mean())
    print(f"fluid rel-L2 on colliding nodes: rho {e_rho:.2e}  u {e_u:.2e}  force {e_f:.2e}")
    assert e_rho <= 1e-10 and e_u <= 1e-10 and e_f <= 1e-9
    pg, fg, dg = lat.ibm_nodes()
    pc, fc, dc = o.ibm_nodes()
    assert rel(pg, pc) <= 1e-12
    assert rel(fg, fc) <= 1e-10 and rel(dg, dc) <= 1e-10


def test_ibm_fp64_closed_box_fibonacci_stiff_contacts():
    """Closed box, walls on every axis, four spheres with Fibonacci meshes
    made by the library, E = 1e3, gravity -z, 3 iterations: 10 steps."""
    import vlbm
    from vlbm import _capi
    lat, o = make_pair(32, 32, 32, "closedBoxBounceBack",
                       dict(max_iterations=3, periodic_x=0, periodic_y=0, periodic_z=0, gz=-1e-4,
                            young=1e3, wall_young=1e3))
    centers = [[9.0, 9.0, 5.5], [16.8, 9.5, 6.0], [9.5, 20.0, 14.0], [22.0, 22.0, 24.0]]
    for c in centers:
        lat.ibm_add_sphere(c, radius=4.0, density=2.0, mesh=_capi.MESH_FIBONACCI, n_fibonacci=201)
    pos, _, _ = lat.ibm_nodes()
    assert pos.shape == (3, 4 * 201)
    area = 4 * math.pi * 16 / 201
    for k, c in enumerate(centers):
        xyz = pos[:, k * 201:(k + 1) * 201].T
        assert np.allclose(np.linalg.norm(xyz - np.array(c), axis=1), 4.0, atol=1e-12)
        o.ibm_add_particle(c, (0, 0, 0), (0, 0, 0), 4.0, 2.0, True, xyz, np.full(201, area))
    n = 10
    print("case: fp64 closed box 4 fibonacci spheres, 10 steps", flush=True)
    lat.step(n, save_last=True)
    o.ibm_run(n, 0)
    lat.ibm_status()
    compare_particles(lat, o, 1e-10)
    compare_fluid(lat, o, 1e-10)


def test_ibm_polar_mesh_matches_reference_mesh():
    import vlbm
    from vlbm import _capi
    cfg = base_cfg(NX=24, NY=24, NZ=24, builder="taylorGreenVortex", tau=0.6)
    lat = vlbm.Lattice(cfg.to_desc())
    lat.init()
    lat.ibm_enable(_capi.default_ibm_desc(mesh_scale=0.8, periodic_y=1))
    lat.ibm_add_sphere([11.2, 12.3, 10.9], radius=5.0, mesh=_capi.MESH_POLAR)
    pos, _, _ = lat.ibm_nodes()
    xyz, S = po.sphere_polar(10.0, [11.2, 12.3, 10.9], mesh_scale=0.8)
    assert pos.shape[1] == len(S)
    assert np.allclose(pos.T, xyz, atol=1e-12)
    lat.close()


def test_ibm_contacts_only_head_on():
    """Spheres without nodes: 300 stages of pure contact dynamics (both
    histories, damping quirk included) match the oracle."""
    lat, o = make_pair(48, 24, 24, "taylorGreenVortex", dict(periodic_y=1, young=0.5, pp_rest=0.7))
    for c, v in (([14.0, 12.0, 12.0], (0.02, 0.001, 0)), ([24.0, 12.5, 12.0], (-0.02, 0, 0))):
        lat.ibm_add_particle(c, v, (0, 0, 0), 4.0, 2.0, True, np.zeros((0, 3)), np.zeros(0))
        o.ibm_add_particle(c, v, (0, 0, 0), 4.0, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    print("case: contacts only, 300 stages", flush=True)
    for s in range(300):
        lat.ibm_stage()
        o.ibm_stage(s)
    compare_particles(lat, o, 1e-10, fields=("pos", "vel", "w"))
    pg, tg, lg = lat.ibm_collision_state()
    pc, tc, lc = o.ibm_collision_state()
    assert np.array_equal(pg, pc) and np.array_equal(lg, lc)
    assert rel(tg, tc) <= 1e-10


def test_ibm_fp32_fluid_fp64_particles():
    ny = 24
    lat, o = make_pair(32, ny, 32, "parallelPlatesBounceBack_Y", dict(gy=-2e-5), precision="single", FX=1e-5)
    add_both(lat, o, [10.0, 9.0, 12.0], 4.0, vel=(0.001, 0, 0))
    add_both(lat, o, [17.5, 9.3, 12.4], 4.0, w=(0, 0.002, 0))
    n = 12
    print("case: fp32 fluid, 12 steps", flush=True)
    lat.step(n, save_last=True)
    o.ibm_run(n, 0)
    g, c = lat.ibm_particles(), o.ibm_particles()
    assert rel(g[:, F["pos"]], c[:, F["pos"]]) <= 1e-6
    assert rel(g[:, F["vel"]], c[:, F["vel"]]) <= 1e-3
    compare_fluid(lat, o, 1e-4)


@pytest.mark.parametrize("layout", ["populations", "moments"])
def test_ibm_fp32_spheres_in_circular_duct(layout):
    """FP32 fluid on the circular duct's node map (the layout large ducts take
    by default: coefficients), FP64 particles, against the FP64 oracle: 10
    coupled steps; fluid u and force field compared on the colliding nodes."""
    import vlbm
    lat, o = make_pair(24, 24, 24, "circularDuctInterpBounceBack",
                       dict(max_iterations=2, periodic_x=0, periodic_y=0, periodic_z=1), precision="single",
                       FZ=1e-5, layout=layout)
    add_both(lat, o, [12.0, 12.0, 12.0], 3.0, vel=(0, 0, 0.001), w=(0, 0.001, 0))
    add_both(lat, o, [12.5, 11.0, 3.0], 3.0, vel=(0.001, 0, 0))
    print(f"case: fp32 circular duct 2 spheres, 10 steps [{layout}]", flush=True)
    lat.step(10, save_last=True)
    o.ibm_run(10, 0)
    g, c = lat.ibm_particles(), o.ibm_particles()
    assert rel(g[:, F["pos"]], c[:, F["pos"]]) <= 1e-6
    assert rel(g[:, F["vel"]], c[:, F["vel"]]) <= 1e-3
    used = ((vlbm.build_map(lat.desc) >> 21) & 1).astype(bool)
    lat.sync()
    m, mo = lat.macroscopics(), o.macroscopics()
    e_u = rel_l2(np.stack([m[k][used] for k in ("ux", "uy", "uz")]),
                 np.stack([mo[k][used] for k in ("ux", "uy", "uz")]))
    fg, fo = lat.ibm_force()[:, used], o.ibm_force()[:, used]
    e_f = rel_l2(fg - fo.mean(), fo - fo.mean())
    print(f"fp32 vs fp64 oracle on colliding nodes: u {e_u:.2e}  force {e_f:.2e}")
    assert e_u <= 1e-4 and e_f <= 1e-3


def test_ibm_fp64_sphere_without_nodes():
    """A sphere without surface nodes has no interpolation chunk: its velocity
    update still runs every iteration (gravity, and a wall contact with a
    meshed sphere), equal to the oracle's."""
    print("case: fp64 sphere without nodes beside a meshed one, 8 steps", flush=True)
    lat, o = make_pair(32, 24, 32, "parallelPlatesBounceBack_Y", dict(gy=-2e-5))
    add_both(lat, o, [10.0, 9.0, 12.0], 4.0, vel=(0.001, 0, 0))
    add_both(lat, o, [20.0, 3.8, 16.0], 4.0, vel=(0, -0.002, 0), w=(0, 0, 0.001), xyz=np.zeros((0, 3)), S=np.zeros(0))
    lat.step(8, save_last=True)
    o.ibm_run(8, 0)
    compare_particles(lat, o, 1e-10, fields=("pos", "vel", "w", "f", "M"))
    g = lat.ibm_particles()
    assert g[1, F["vel"]][1] != -0.002            # the node-less sphere's velocity was updated
    lat.close()


def test_ibm_fp32_plain_flush_equals_atomic_flush():
    """The spreading flush writes box nodes that no other sphere's box covers
    with plain stores, the others with atomics. With boxes that do not overlap,
    every node has one contributor, so the plain run equals, bit for bit, a
    run whose stray node sends every box to the atomic flush; both follow the
    FP64 oracle within the FP32 bars."""
    print("case: fp32 three isolated spheres, plain vs atomic flush, 10 steps", flush=True)
    out = []
    for stray in (False, True):
        lat, o = make_pair(64, 32, 32, "parallelPlatesBounceBack_Y", dict(gy=-2e-5), precision="single", FX=1e-5)
        for k, x in enumerate((12.0, 32.0, 52.0)):
            c = [x, 15.0 + 0.3 * k, 16.0 + 0.2 * k]
            xyz, S = po.sphere_polar(8.0, c)
            if stray and k == 1:
                xyz = np.vstack([xyz, [c[0] + 6.0, c[1], c[2]]])
                S = np.append(S, 0.0)
            add_both(lat, o, c, 4.0, vel=(0.001 * (k - 1), 0, 0), w=(0, 0, 0.002 * k), xyz=xyz, S=S)
        lat.step(10, save_last=True)
        o.ibm_run(10, 0)
        lat.sync()
        m = lat.macroscopics()
        out.append((m, lat.ibm_force(), lat.ibm_particles()))
        g, c = lat.ibm_particles(), o.ibm_particles()
        assert rel(g[:, F["pos"]], c[:, F["pos"]]) <= 1e-6
        assert rel(g[:, F["vel"]], c[:, F["vel"]]) <= 1e-3
        compare_fluid(lat, o, 1e-4)
        lat.close()
    (m0, f0, p0), (m1, f1, p1) = out
    for k in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(m0[k], m1[k]), k
    assert np.array_equal(f0, f1)
    assert np.array_equal(p0[:, F["pos"]], p1[:, F["pos"]]) and np.array_equal(p0[:, F["vel"]], p1[:, F["vel"]])


def test_ibm_timing_split_and_status():
    from vlbm import _capi
    lat, o = make_pair(32, 32, 32, "closedBoxBounceBack",
                       dict(periodic_x=0, periodic_y=0, periodic_z=0), precision="single")
    lat.ibm_add_sphere([16.0, 16.0, 16.0], radius=4.0, mesh=_capi.MESH_FIBONACCI, n_fibonacci=201)
    ms_lbm, ms_ibm = lat.ibm_time_steps(5)
    assert ms_lbm > 0 and ms_ibm > 0
    lat.ibm_status()
    assert lat.steps_done == 5
