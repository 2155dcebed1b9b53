"""Pins the oracle's interpolated bounce-back (oracle/oracle.cpp bc_interp_bb,
following src/CUDA/boundaryConditionsSchemes/interpolatedBounceBack.cu:6-354)
by things the oracle's own code does not contain:

* the reference's archived FP64 runs of its circular and annular duct
  validation (u_z along the north-south line and the diagonal, 7 significant
  digits; tests/golden/duct_profiles.json, written by make_duct_profiles.py);
* closed forms: Hagen-Poiseuille in the circular duct, pressure-driven flow in
  the annulus, Taylor-Couette flow between a rotating inner wall and a fixed
  outer wall, and axial Couette flow with either wall sliding along z, with
  the convergence order between two radii.

The moving-wall checks fail, on a 16-node radius, for each of: the 3/q factor
of the far-wall branch doubled, the 6 of the near-wall branch halved, the wall
term's sign flipped, the far-wall interpolation weights swapped, the near-wall
populations swapped, the rotation turned radial, and the z component of the
wall velocity mirrored.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import pyoracle as po
from helpers import base_cfg, rel_l2

GOLDEN = Path(__file__).resolve().parent / "golden" / "duct_profiles.json"
TAU = 0.8
NU = (TAU - 0.5) / 3


def duct_run(builder, n, R, r, steps, tau=TAU, F=0.0, w_in=0.0, u_in=0.0, u_out=0.0):
    """FP64 oracle run of a duct of n x n x 1 nodes; returns macroscopics, the
    used-node mask and the node coordinates relative to the duct axis."""
    cfg = base_cfg(NX=n, NY=n, NZ=1, tau=tau, FZ=F, builder=builder)
    d = po.desc_from(cfg.to_desc())
    d.duct_outer_radius = R
    d.duct_inner_radius = r
    d.inner_rotation = w_in
    d.inner_velocity = u_in
    d.outer_velocity = u_out
    o = po.Oracle(d, 8)
    o.init()
    o.step(steps, save_last=True, threads=4)
    m = o.macroscopics()
    o.close()
    used = ((po.build_map(d)[0] >> 21) & 1).astype(bool)
    y, x = np.mgrid[0:n, 0:n]
    X = x + 0.5 - n / 2.0                       # DUCT_CENTER_X = (NX-1)/2 + 1/2 (var.h:144)
    Y = y + 0.5 - n / 2.0
    return {k: v[0] for k, v in m.items()}, used, np.hypot(X, Y), X, Y


# ---------------------------------------------------------------------------
# the reference's archived runs
# ---------------------------------------------------------------------------
def _cases():
    return json.loads(GOLDEN.read_text())["cases"]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c["source"].split("/")[-3] + "-" + c["source"][-4:-1])
def test_duct_profiles_match_the_reference_archive(case):
    """The info files do not record the radii; the archive's centre value
    identifies R = (NX-1)/2 (circular 17^2: Hagen-Poiseuille F R^2 / 4 nu =
    5.859e-2 against 5.852e-2 archived) and the inner radius is var.h's
    INNER_RADIUS = OUTER_RADIUS / 4 (var.h:155). The archive has 7
    significant digits; a wrong q branch, weight or wall distance moves the
    profile by far more than 1e-6."""
    inf = case["info"]
    n = inf["NX"]
    assert inf["NY"] == n and inf["NZ"] == 1 and inf["Velocity set"] == "D3Q19"
    R = (n - 1) / 2.0
    r = R / 4.0 if case["builder"].startswith("annular") else 0.0
    m, used, *_ = duct_run(case["builder"], n, R, r, inf["Nsteps"], tau=inf["Tau"], F=inf["FZ"])
    uz = m["uz"]
    c = (n - 1) // 2
    ref = np.array(case["uz_north_south"])
    assert np.array_equal(ref[:, 0], np.arange(n))
    scale = np.abs(ref[:, 1]).max()
    np.testing.assert_allclose(uz[:, c], ref[:, 1], rtol=0, atol=1e-6 * scale)
    if "uz_southwest_northeast" in case:
        diag = np.array(case["uz_southwest_northeast"])[:, 1]
        ours = np.array([uz[k, k] for k in range(n) if used[k, k]])
        assert ours.size == diag.size
        np.testing.assert_allclose(ours, diag, rtol=0, atol=1e-6 * scale)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
def _steps(R):
    return int(6 * R * R / NU)


def test_hagen_poiseuille_second_order():
    """u_z(r) = F (R^2 - r^2) / (4 nu) in the circular duct; the error falls
    by more than 4x from R = 8 to R = 16 (Bouzidi's linear interpolation is
    second order)."""
    errs = []
    for R in (8.0, 16.0):
        n = int(2 * R) + 2
        F = 1e-4 * (8.0 / R) ** 2
        m, used, rr, _, _ = duct_run("circularDuctInterpBounceBack", n, R, 0.0, _steps(R), F=F)
        exact = F * (R * R - rr * rr) / (4 * NU)
        errs.append(rel_l2(m["uz"][used], exact[used]))
    assert errs[1] < 1e-3, errs
    assert errs[0] / errs[1] > 4.0, errs


def test_annular_poiseuille_second_order():
    """u_z = F/(4 nu) [R^2 - r^2 + (R^2 - r_i^2) ln(r/R) / ln(R/r_i)] between
    two fixed walls (inner wall branch with the outer one)."""
    errs = []
    for R in (8.0, 16.0):
        n, ri = int(2 * R) + 2, 0.375 * R
        F = 1e-4 * (8.0 / R) ** 2
        m, used, rr, _, _ = duct_run("annularDuctInterpBounceBack", n, R, ri, _steps(R), F=F)
        exact = F / (4 * NU) * (R * R - rr * rr + (R * R - ri * ri) * np.log(rr / R) / np.log(R / ri))
        errs.append(rel_l2(m["uz"][used], exact[used]))
    assert errs[1] < 5e-3, errs
    assert errs[0] / errs[1] > 3.0, errs


def test_taylor_couette_rotating_inner_wall():
    """u_theta = A r + B / r with A = -w r_i^2 / (R^2 - r_i^2), B = w r_i^2 R^2 / (R^2 - r_i^2)
    for an inner wall turning at w and a fixed outer wall. The reference
    evaluates the wall velocity at the boundary node's own polar angle
    (interpolatedBounceBack.cu:36-49), not at the link's wall crossing, so the
    error is first order: 16 % at R = 8, 6 % at R = 16."""
    errs = []
    for R in (8.0, 16.0):
        n, ri = int(2 * R) + 2, 0.375 * R
        w = 0.02 / ri
        m, used, rr, X, Y = duct_run("annularDuctInterpBounceBack", n, R, ri, _steps(R), w_in=w)
        A = -w * ri * ri / (R * R - ri * ri)
        B = w * ri * ri * R * R / (R * R - ri * ri)
        ut = A * rr + B / rr
        ex, ey = -ut * Y / rr, ut * X / rr
        num = np.sum((m["ux"][used] - ex[used]) ** 2 + (m["uy"][used] - ey[used]) ** 2)
        errs.append(float(np.sqrt(num / np.sum(ex[used] ** 2 + ey[used] ** 2))))
        assert np.abs(m["uz"][used]).max() < 1e-12
    assert errs[1] < 0.08, errs
    assert errs[0] / errs[1] > 2.0, errs


@pytest.mark.parametrize("wall", ["inner", "outer"])
def test_axial_couette_sliding_wall(wall):
    """One wall sliding along z at U, the other fixed, no force:
    u_z = U ln(r/R) / ln(r_i/R) (inner wall moving) or U ln(r/r_i) / ln(R/r_i)
    (outer wall moving). The wall terms of the links with c_z != 0 leave an
    in-plane flow of a few 1e-4 U."""
    R, U = 16.0, 0.02
    n, ri = int(2 * R) + 2, 0.375 * R
    kw = {"u_in": U} if wall == "inner" else {"u_out": U}
    m, used, rr, _, _ = duct_run("annularDuctInterpBounceBack", n, R, ri, _steps(R), **kw)
    exact = U * (np.log(rr / R) / np.log(ri / R) if wall == "inner" else np.log(rr / ri) / np.log(R / ri))
    assert rel_l2(m["uz"][used], exact[used]) < 2e-3
    assert np.abs(m["ux"][used]).max() < 1e-3 * U and np.abs(m["uy"][used]).max() < 1e-3 * U
