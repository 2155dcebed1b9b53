"""Pins the immersed-boundary part of the CPU oracle (oracle/oracle.cpp, class
Ibm) to what the mathematics fixes: Peskin's kernel moments, the exact area
of the polar sphere mesh, the rigid-body equations (Newton, Euler with the
gyroscopic term, the inertia rotation R I R^T), Hertz's normal law and
Coulomb's cap for contacts, momentum conservation of contacts and of the
force spreading, and the rotation of the Lagrangian nodes.

Every check uses a law stated independently of the oracle's code (textbook
formula, numpy linear algebra, scipy rotations), so a dropped term, a wrong
sign or a transposed operand fails one of them.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import pyoracle as po
from helpers import base_cfg

F = po.PARTICLE_FIELDS


def lattice(nx=24, ny=24, nz=24, builder="parallelPlatesBounceBack_Y", tau=0.6, **kw):
    cfg = base_cfg(NX=nx, NY=ny, NZ=nz, builder=builder, tau=tau, **kw)
    o = po.Oracle(po.desc_from(cfg.to_desc()), 8)
    o.init()
    return o


# ---------------------------------------------------------------------------
# Peskin 4-point kernel (src/CUDA/IBM/ibmGlobalFunctions.h:28-48)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("x", [0.0, 0.13, 0.5, 0.77, 0.999])
def test_peskin_kernel_moments(x):
    """Peskin (2002): sum_j phi(j-x) = 1, sum_j (j-x) phi(j-x) = 0,
    sum_j phi(j-x)^2 = 3/8 for every x; support |r| < 2."""
    js = np.arange(-4, 6)
    phi = np.array([po.stencil4(j - x) for j in js])
    r = js - x
    assert math.isclose(phi.sum(), 1.0, abs_tol=1e-14)
    assert math.isclose((r * phi).sum(), 0.0, abs_tol=1e-14)
    assert math.isclose((phi ** 2).sum(), 3 / 8, abs_tol=1e-14)
    assert np.all(phi[np.abs(r) >= 2] == 0)
    assert np.all(phi >= 0)


# ---------------------------------------------------------------------------
# sphere mesh (src/CUDA/IBM/structs/particle.cpp:698-995)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("diameter,scale", [(16, 1.0), (10, 0.5), (20, 1.0)])
def test_sphere_polar_mesh(diameter, scale):
    c = np.array([11.3, 12.0, 9.7])
    xyz, S = po.sphere_polar(diameter, c, mesh_scale=scale)
    r = diameter / 2
    n_layer = int(2 * math.sqrt(2) * r / scale + 1)
    expect = sum(int(1.5 + math.cos(math.pi * (i / n_layer - 0.5)) * n_layer * math.sqrt(3))
                 for i in range(n_layer + 1))
    assert len(S) == expect
    # rings are equal-area segments: the areas add up to the sphere's
    assert math.isclose(S.sum(), 4 * math.pi * r * r, rel_tol=1e-13)
    assert np.allclose(np.linalg.norm(xyz - c, axis=1), r, atol=1e-12)
    # poles and a symmetric distribution of heights
    assert np.allclose(xyz[0] - c, [0, 0, -r], atol=1e-12)
    assert np.allclose(xyz[-1] - c, [0, 0, r], atol=1e-12)
    assert abs((xyz[:, 2] - c[2]) @ S) < 1e-9 * S.sum() * r


def test_sphere_polar_breugem_and_coulomb():
    c = np.zeros(3)
    xyz, S = po.sphere_polar(16, c, breugem=0.3)
    assert np.allclose(np.linalg.norm(xyz, axis=1), 8 - 0.3, atol=1e-12)
    # ring areas are those of the shrunk sphere (the shift comes first)
    assert math.isclose(S.sum(), 4 * math.pi * (8 - 0.3) ** 2, rel_tol=1e-13)
    xyz2, S2 = po.sphere_polar(8, c, coulomb=2)
    assert np.allclose(np.linalg.norm(xyz2, axis=1), 4, atol=1e-12)
    # the Coulomb relaxation gives every node the nominal area / n
    assert np.allclose(S2, 4 * math.pi * 16 / len(S2))


# ---------------------------------------------------------------------------
# rigid-body update (src/CUDA/IBM/ibm.cu:755-897)
# ---------------------------------------------------------------------------
def dem_oracle(ibm_kw=None, **lat_kw):
    o = lattice(**lat_kw)
    o.ibm_enable(po.ibm_desc(**(ibm_kw or {})))
    return o


def set_fields(o, p, **vals):
    rec = o.ibm_particles()
    for k, v in vals.items():
        rec[p, F[k]] = v
    o.ibm_set_particles(rec)


def test_newton_update_closed_form():
    g = np.array([1e-4, -2e-4, 3e-5])
    o = dem_oracle(dict(gx=g[0], gy=g[1], gz=g[2], ref_float_helpers=0))
    o.ibm_add_particle([12, 12, 12], [0, 0, 0], [0, 0, 0], 4.0, 2.5, True, np.zeros((0, 3)), np.zeros(0))
    rng = np.random.default_rng(3)
    f, f_old, v_old, dP = (rng.standard_normal(3) for _ in range(4))
    set_fields(o, 0, f=f, f_old=f_old, vel_old=v_old, dP=dP)
    o.ibm_update_velocity()
    rec = o.ibm_particles()[0]
    V = 4 / 3 * math.pi * 4.0 ** 3
    rho_p, rho_f, alpha = 2.5, 1.0, 0.5
    # m dv = F_avg + dP + (m - m_fluid) g   with m = rho_p V
    expect = v_old + (alpha * f + (1 - alpha) * f_old + dP) / (rho_p * V) + (1 - rho_f / rho_p) * g
    assert np.allclose(rec[F["vel"]], expect, rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("diag", [True, False])
def test_euler_update_and_inertia_rotation(diag):
    """w = w_old + I^-1 (M_avg + dL - w_avg x (I w_avg)); then I <- R I R^T
    with R the rotation by the vector w_avg (textbook Euler equations)."""
    o = dem_oracle(dict(ref_float_helpers=0))
    o.ibm_add_particle([12, 12, 12], [0, 0, 0], [0, 0, 0], 3.0, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    rng = np.random.default_rng(5 if diag else 6)
    if diag:
        Imat = np.diag([3.0, 5.0, 7.0])
    else:
        A = rng.standard_normal((3, 3))
        Imat = A @ A.T + 3 * np.eye(3)
    I6 = [Imat[0, 0], Imat[1, 1], Imat[2, 2], Imat[0, 1], Imat[0, 2], Imat[1, 2]]
    w_old = 0.05 * rng.standard_normal(3)
    M, M_old, dL = (rng.standard_normal(3) for _ in range(3))
    set_fields(o, 0, I=I6, w_old=w_old, M=M, M_old=M_old, dL=dL)
    o.ibm_update_velocity()
    rec = o.ibm_particles()[0]
    alpha = 0.5
    w_avg = w_old  # the average is taken with the previous rotation in both slots
    A = dL + alpha * M + (1 - alpha) * M_old - np.cross(w_avg, Imat @ w_avg)
    expect_w = w_old + np.linalg.solve(Imat, A)
    assert np.allclose(rec[F["w"]], expect_w, rtol=1e-12, atol=1e-15)
    if diag:
        # Euler's equations for principal axes: I1 dw1 = M1 - (I3 - I2) w2 w3
        I1, I2, I3 = 3.0, 5.0, 7.0
        d = dL + alpha * M + (1 - alpha) * M_old
        dw = np.array([(d[0] - (I3 - I2) * w_avg[1] * w_avg[2]) / I1,
                       (d[1] - (I1 - I3) * w_avg[2] * w_avg[0]) / I2,
                       (d[2] - (I2 - I1) * w_avg[0] * w_avg[1]) / I3])
        assert np.allclose(rec[F["w"]], w_old + dw, rtol=1e-12, atol=1e-15)
    R = Rotation.from_rotvec(w_avg).as_matrix()
    In = R @ Imat @ R.T
    got = rec[F["I"]]
    assert np.allclose(got, [In[0, 0], In[1, 1], In[2, 2], In[0, 1], In[0, 2], In[1, 2]], rtol=1e-12, atol=1e-13)


def test_float_helpers_only_perturb_at_float_level():
    """The reference's float sqrt/rsqrt in FP64 builds change the inertia
    rotation by O(1e-7) relative, not more."""
    res = []
    for helpers in (0, 1):
        o = dem_oracle(dict(ref_float_helpers=helpers))
        o.ibm_add_particle([12, 12, 12], [0, 0, 0], [0.01, 0.02, -0.03], 3.0, 2.0, True, np.zeros((0, 3)),
                           np.zeros(0))
        set_fields(o, 0, I=[3, 5, 7, 0.5, 0.2, -0.4], w_old=[0.01, 0.02, -0.03])
        o.ibm_update_velocity()
        res.append(o.ibm_particles()[0][F["I"]])
    d = np.abs(res[0] - res[1]).max() / np.abs(res[0]).max()
    assert 0 < d < 1e-6


# ---------------------------------------------------------------------------
# contacts (src/CUDA/IBM/collision/ibmCollision.cu)
# ---------------------------------------------------------------------------
def hertz_pair_kn(E, nu, r1, r2):
    Es = 1 / ((1 - nu ** 2) / E + (1 - nu ** 2) / E)
    Rs = r1 * r2 / (r1 + r2)
    return 4 / 3 * Es * math.sqrt(Rs)


def test_sphere_sphere_hertz_normal_force():
    """Static overlap, no relative motion: |F| = (4/3) E* sqrt(R*) delta^1.5 along
    the line of centres, repulsive, equal and opposite; no torque."""
    o = dem_oracle(dict(young=1e3, poisson=0.3), nx=40, ny=40, nz=40)
    r1, r2, delta = 4.0, 3.0, 0.2
    c1 = np.array([15.0, 18.0, 20.0])
    n = np.array([1.0, 2.0, -0.5]) / np.linalg.norm([1.0, 2.0, -0.5])
    c2 = c1 + n * (r1 + r2 - delta)
    o.ibm_add_particle(c1, [0, 0, 0], [0, 0, 0], r1, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    o.ibm_add_particle(c2, [0, 0, 0], [0, 0, 0], r2, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    o.ibm_collisions(10)
    rec = o.ibm_particles()
    kn = hertz_pair_kn(1e3, 0.3, r1, r2)
    mag = kn * delta ** 1.5
    # particle 0 is pushed away from particle 1 (along -n), particle 1 along +n
    assert np.allclose(rec[0, F["f"]], -mag * n, rtol=1e-6)
    assert np.allclose(rec[1, F["f"]], mag * n, rtol=1e-6)
    assert np.allclose(rec[0, F["M"]], 0, atol=1e-12)
    partner, tang, last = o.ibm_collision_state()
    # the lower index keeps the history, in the first partner slot
    assert partner[0, 7] == 1 and last[0, 7] == 10
    assert np.all(partner[1] == -1)


def test_sphere_sphere_coulomb_cap_and_torque():
    """Sliding contact: tangential force capped at mu |F_n|, equal and opposite
    on the two spheres, torque r n x F_t on both.

    Direction: the reference applies -f_dirs to the lower-index sphere with
    f_tang built from that sphere's own slip (ibmCollision.cu:1856-1890), so
    the tangential force on it points ALONG its slip, not against it; the
    wall contact (force applied with +f_dirs) opposes the slip. Both kept."""
    mu = 0.1
    o = dem_oracle(dict(young=1e3, poisson=0.3, pp_friction=mu), nx=40, ny=40, nz=40)
    r = 4.0
    c1 = np.array([15.0, 20.0, 20.0])
    c2 = c1 + np.array([2 * r - 0.1, 0, 0])
    o.ibm_add_particle(c1, [0, 0.5, 0], [0, 0, 0], r, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    o.ibm_add_particle(c2, [0, 0, 0], [0, 0, 0], r, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    o.ibm_collisions(3)
    rec = o.ibm_particles()
    f0 = rec[0, F["f"]]
    n = np.array([-1.0, 0, 0])  # from sphere 1 towards sphere 0 (the reference's n = (p_i - p_j)/|.|)
    fn = (f0 @ n) * n
    ft = f0 - fn
    # sphere 0 slides in +y relative to sphere 1: the reference's tangential force on it is +y
    assert ft[1] > 0 and abs(ft[0]) < 1e-12 and abs(ft[2]) < 1e-12
    assert math.isclose(np.linalg.norm(ft), mu * np.linalg.norm(fn), rel_tol=1e-9)
    assert np.allclose(rec[0, F["f"]], -rec[1, F["f"]], rtol=1e-12)
    # torque: r n x F_t evaluated with the contact normal and the force on sphere 1
    ft1 = rec[1, F["f"]] - (rec[1, F["f"]] @ n) * n
    tq = r * np.cross(n, ft1)
    assert np.allclose(rec[0, F["M"]], tq, rtol=1e-9)
    assert np.allclose(rec[1, F["M"]], tq, rtol=1e-9)


def test_sphere_wall_hertz_and_walls_at_grid_ends():
    """Wall at y = 0 (and y = NY-1): |F| = (4/3) E* sqrt(r) delta^1.5 with the
    wall's own modulus, pointing into the domain."""
    E, nu, Ew, nuw = 1e3, 0.3, 5e3, 0.25
    o = dem_oracle(dict(young=E, poisson=nu, wall_young=Ew, wall_poisson=nuw), nx=30, ny=30, nz=30)
    r, delta = 4.0, 0.15
    o.ibm_add_particle([15, r - delta, 15], [0, 0, 0], [0, 0, 0], r, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    o.ibm_add_particle([10, 29 - (r - delta), 10], [0, 0, 0], [0, 0, 0], r, 2.0, True, np.zeros((0, 3)),
                       np.zeros(0))
    o.ibm_collisions(5)
    rec = o.ibm_particles()
    Es = 1 / ((1 - nu ** 2) / E + (1 - nuw ** 2) / Ew)
    mag = 4 / 3 * Es * math.sqrt(r) * delta ** 1.5
    assert np.allclose(rec[0, F["f"]], [0, mag, 0], rtol=1e-6, atol=1e-12)
    assert np.allclose(rec[1, F["f"]], [0, -mag, 0], rtol=1e-6, atol=1e-12)
    partner, tang, last = o.ibm_collision_state()
    # wall slots are a function of the normal (calculateWallIndex)
    assert (last[0] == 5).sum() == 1 and (last[1] == 5).sum() == 1
    assert np.argmax(last[0] == 5) != np.argmax(last[1] == 5)


def dem_head_on(rest, steps=400, v=0.02):
    o = dem_oracle(dict(young=0.5, poisson=0.3, pp_rest=rest, pp_friction=0.0), nx=64, ny=32, nz=32)
    r = 4.0
    o.ibm_add_particle([20, 16, 16], [v, 0, 0], [0, 0, 0], r, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    o.ibm_add_particle([30, 16, 16], [-v, 0, 0], [0, 0, 0], r, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    for s in range(steps):
        o.ibm_stage(s)
    rec = o.ibm_particles()
    return rec[0, F["vel"]], rec[1, F["vel"]]


def dem_wall_bounce(rest, steps=400, v=0.02):
    o = dem_oracle(dict(young=0.5, poisson=0.3, pw_rest=rest, pw_friction=0.0, wall_young=0.5), nx=32,
                   ny=32, nz=32)
    o.ibm_add_particle([16, 6, 16], [0, -v, 0], [0, 0, 0], 4.0, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    for s in range(steps):
        o.ibm_stage(s)
    return o.ibm_particles()[0, F["vel"]]


def test_head_on_collision_momentum_and_damping_sign():
    """Particles without Lagrangian nodes are pure contact bodies.

    Momentum is conserved exactly. Without damping the bounce is nearly
    elastic (the force average lags the motion by one step, which adds a few
    per cent over a ~75-step Hertz contact). The reference evaluates the
    sphere-sphere damping with the sign of the force on the partner and
    applies it with the opposite sign (ibmCollision.cu:1808-1810,1885-1895),
    so a restitution coefficient below one makes the rebound FASTER; the
    wall contact damps as it should (next test)."""
    v0, v1 = dem_head_on(1.0)
    assert np.allclose(v0 + v1, 0, atol=1e-15)
    e_elastic = -v0[0] / 0.02
    assert v0[0] < 0 and 0.95 < e_elastic < 1.15
    assert abs(v0[1]) < 1e-15 and abs(v0[2]) < 1e-15
    w0, w1 = dem_head_on(0.5)
    assert np.allclose(w0 + w1, 0, atol=1e-15)
    assert -w0[0] / 0.02 > e_elastic * 1.2


def test_wall_bounce_damping():
    e1 = dem_wall_bounce(1.0)[1] / 0.02
    e5 = dem_wall_bounce(0.5)[1] / 0.02
    assert 0.95 < e1 < 1.15
    assert 0.3 < e5 < 0.8 * e1


# ---------------------------------------------------------------------------
# interpolation / spreading (src/CUDA/IBM/ibm.cu:194-542)
# ---------------------------------------------------------------------------
def one_sphere(o, center, vel=(0, 0, 0), w=(0, 0, 0), r=4.0, movable=True):
    xyz, S = po.sphere_polar(2 * r, center)
    o.ibm_add_particle(center, vel, w, r, 2.0, movable, xyz, S)
    return xyz, S


def test_interpolation_of_uniform_field_and_force_balance():
    """Uniform fluid velocity u0 around a sphere at rest: every node sees u0
    (partition of unity), dF = 2 rho dA u0, and the spread force sums to -sum dF."""
    o = dem_oracle(dict(periodic_y=1), nx=24, ny=24, nz=24, builder="taylorGreenVortex")
    xyz, S = one_sphere(o, np.array([11.37, 12.21, 11.83]), movable=False)
    shape = (24, 24, 24)
    rho = np.full(shape, 1.02)
    u0 = np.array([0.01, -0.02, 0.005])
    o.ibm_set_macr(rho, *(np.full(shape, c) for c in u0))
    aux = o.ibm_interp_spread()
    pos, nf, ndf = o.ibm_nodes()
    assert np.allclose(ndf.T, 2 * 1.02 * S[:, None] * u0[None, :], rtol=1e-12, atol=1e-17)
    fsum = aux[3:].reshape(3, -1).sum(axis=1)
    assert np.allclose(fsum, -ndf.sum(axis=1), rtol=1e-11)
    # velocity correction is F / (2 rho) node by node
    assert np.allclose(aux[:3], 0.5 * aux[3:] / 1.02, rtol=1e-11, atol=1e-20)
    rec = o.ibm_particles()
    assert np.allclose(rec[0, F["f"]], ndf.sum(axis=1), rtol=1e-12)


def test_rotating_sphere_torque():
    """Sphere spinning in fluid at rest: dF_p = -2 rho dA (w x r_p); total
    force ~ 0 and torque -2 rho sum dA (r^2 w - r (r.w)) = -(4/3) rho S r^2 w."""
    o = dem_oracle(dict(periodic_y=1), nx=24, ny=24, nz=24, builder="taylorGreenVortex")
    w = np.array([0.001, -0.002, 0.0015])
    c = np.array([12.0, 12.0, 12.0])
    xyz, S = one_sphere(o, c, w=w)
    shape = (24, 24, 24)
    o.ibm_set_macr(np.ones(shape), np.zeros(shape), np.zeros(shape), np.zeros(shape))
    o.ibm_interp_spread()
    rec = o.ibm_particles()
    r = xyz - c
    expect_M = -2 * (S[:, None] * np.cross(r, np.cross(w, r))).sum(axis=0)
    assert np.allclose(rec[0, F["M"]], expect_M, rtol=1e-10)
    assert np.allclose(rec[0, F["M"]], -(4 / 3) * S.sum() * 16 * w, rtol=2e-2)
    assert np.linalg.norm(rec[0, F["f"]]) < 1e-3 * np.linalg.norm(rec[0, F["M"]]) / 4


def test_wall_clipping_keeps_partition_inside():
    """A node next to the y=0 wall interpolates only from in-domain nodes
    (the stencil is clipped, not wrapped)."""
    o = dem_oracle(dict(periodic_y=0), nx=16, ny=16, nz=16)
    o.ibm_add_particle([8, 1.0, 8], [0, 0, 0], [0, 0, 0], 1.0, 2.0, False, np.array([[8.2, 0.4, 8.3]]),
                       np.array([1.0]))
    shape = (16, 16, 16)
    uy = np.zeros(shape)
    uy[:, 15, :] = 1.0  # would leak in through a periodic wrap
    o.ibm_set_macr(np.ones(shape), np.zeros(shape), uy, np.zeros(shape))
    o.ibm_interp_spread()
    _, _, ndf = o.ibm_nodes()
    assert ndf[1, 0] == 0.0
    # clipped stencil: weights on y = 0, 1, 2 only (base int(0.4) - 1 = -1)
    w = sum(po.stencil4(j - 0.4) for j in (0, 1, 2))
    uy[:, 15, :] = 0.0
    uy[:, 0:3, :] = 1.0
    o2 = dem_oracle(dict(periodic_y=0), nx=16, ny=16, nz=16)
    o2.ibm_add_particle([8, 1.0, 8], [0, 0, 0], [0, 0, 0], 1.0, 2.0, False, np.array([[8.2, 0.4, 8.3]]),
                        np.array([1.0]))
    o2.ibm_set_macr(np.ones(shape), np.zeros(shape), uy, np.zeros(shape))
    o2.ibm_interp_spread()
    _, _, ndf2 = o2.ibm_nodes()
    # the interpolated density is clipped the same way: dF = 2 (w rho) (w u) dA
    assert math.isclose(ndf2[1, 0], 2 * w * w, rel_tol=1e-12)


# ---------------------------------------------------------------------------
# coupled stage: Euler sum and movement (ibm.cu:6-192, 899-1197)
# ---------------------------------------------------------------------------
def test_stage_momentum_exchange_and_force_field():
    """After one coupled step the force field holds F0 + spread force, and the
    force added to the fluid is minus the IBM force on the particle."""
    o = dem_oracle(dict(gz=-1e-4, max_iterations=2), nx=20, ny=20, nz=20)
    xyz, S = one_sphere(o, np.array([10.3, 9.8, 10.1]), vel=(0.002, 0, -0.001))
    o.ibm_run(1, 0)
    force = o.ibm_force()
    added = force.reshape(3, -1).sum(axis=1)
    _, nf, _ = o.ibm_nodes()
    assert np.allclose(added, -nf.sum(axis=1), rtol=1e-10)
    rec = o.ibm_particles()
    # no contacts: the particle force is the IBM force only
    assert np.allclose(rec[0, F["f"]], nf.sum(axis=1), rtol=1e-12)


def test_translation_and_rotation_about_z():
    """Positions advance by (v + v_old)/2; nodes rotate rigidly about the
    centre by |w_avg| (a rotation about z avoids the reference's
    (qi qj - q0 qj) entry, which only matters for x/y components)."""
    c = np.array([16.0, 16.0, 16.0])
    w = np.array([0, 0, 0.01])
    v = np.array([0.003, -0.002, 0.001])
    # no iterations and alpha = 1: no fluid force, the average rotation is w
    o2 = dem_oracle(dict(periodic_y=1, ref_float_helpers=0, max_iterations=0, alpha=1.0), nx=32, ny=32,
                    nz=32, builder="taylorGreenVortex")
    xyz2, _ = one_sphere(o2, c, vel=v, w=w, r=3.0)
    o2.ibm_stage(0)
    rec2 = o2.ibm_particles()[0]
    # with no forces and alpha = 1 the velocity stays v_old + 0 and w stays w_old
    assert np.allclose(rec2[F["vel"]], v, atol=1e-16)
    assert np.allclose(rec2[F["pos"]], c + v, atol=1e-14)
    pos, _, _ = o2.ibm_nodes()
    R = Rotation.from_rotvec(w).as_matrix()
    expect = (c + v) + (R @ (xyz2 - c).T).T
    assert np.allclose(pos.T, expect, atol=1e-12)


# ---------------------------------------------------------------------------
# semi-axis points of the particle record (ibm.cu:958-991)
# ---------------------------------------------------------------------------
def quat(w):
    """Unit quaternion of the rotation by |w| about w: (q0, qi, qj, qk)."""
    n = np.linalg.norm(w)
    return np.concatenate([[math.cos(n / 2)], np.asarray(w) / n * math.sin(n / 2)])


@pytest.mark.parametrize("w", [(0.0, 0.0, 0.02), (0.011, 0.013, 0.0), (0.004, -0.009, 0.017)])
def test_semi_axes_rotate_with_the_reference_matrix(w):
    """The three semi-axis points move to pos + M (p - pos_old) where M is the
    rotation matrix of the quaternion of w_avg (scipy) except for entry
    (z, x), which the reference computes as 2 (qi qj - q0 qj) instead of
    2 (qi qk - q0 qj). A rotation about z leaves that entry exact."""
    w = np.array(w)
    c = np.array([16.0, 15.5, 16.2])
    v = np.array([0.002, 0.001, -0.003])
    r = 3.0
    o = dem_oracle(dict(periodic_y=1, ref_float_helpers=0, max_iterations=0, alpha=1.0), nx=32, ny=32,
                   nz=32, builder="taylorGreenVortex")
    o.ibm_add_particle(c, v, w, r, 2.0, True, np.zeros((0, 3)), np.zeros(0))
    ax0 = o.ibm_axes()[0]
    assert np.array_equal(ax0, [[r, r, r], [0, 0, 0], [0, 0, 0]])   # particle.cpp:777-778
    o.ibm_move()
    q0, qi, qj, qk = quat(w)
    M = Rotation.from_rotvec(w).as_matrix()
    M[2, 0] += 2 * qi * qj - 2 * qi * qk
    expect = (c + v) + (M @ (ax0 - c).T).T
    assert np.allclose(o.ibm_axes()[0], expect, rtol=0, atol=1e-12)
    if w[0] == 0 and w[1] == 0:
        R = Rotation.from_rotvec(w).as_matrix()
        assert np.allclose(expect, (c + v) + (R @ (ax0 - c).T).T, atol=1e-14)


def test_semi_axes_become_nan_without_rotation():
    """No guard on |w_avg| = 0: q = (1, 0/0, 0/0, 0/0) makes every semi-axis
    coordinate NaN, while positions and nodes (which take the translation
    branch) stay finite."""
    c = np.array([12.0, 12.0, 12.0])
    o = dem_oracle(dict(periodic_y=1, ref_float_helpers=0, max_iterations=0, alpha=1.0), nx=24, ny=24,
                   nz=24, builder="taylorGreenVortex")
    one_sphere(o, c, vel=(0.001, 0, 0), r=3.0)
    o.ibm_move()
    assert np.isnan(o.ibm_axes()).all()
    pos, _, _ = o.ibm_nodes()
    assert np.isfinite(pos).all()
    assert np.allclose(o.ibm_particles()[0, F["pos"]], c + [0.001, 0, 0], atol=1e-15)
