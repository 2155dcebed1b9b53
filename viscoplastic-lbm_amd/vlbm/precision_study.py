"""Precision study: D3Q27 regularized power-law channel on one GPU in FP64,
FP32, and FP16- and bfloat16-deviation storage, each 16-bit type with
stochastic rounding and with rounding to nearest.

Case: n = 0.6, consistency K chosen so that the nominal tau is 0.7 at the
mean shear rate (K gamma^(n-1) = (0.7 - 1/2)/3 with gamma = u_max/h, solved
by fixed-point iteration on the analytic profile), apparent tau clamped to
[0.51, 2.0], periodic x and z, bounce-back walls in y (halfway: walls at
y = -1/2 and NY - 1/2), F_x = 5e-6, rho = 1, u = U(-1e-4, 1e-4) per
component from a fixed seed. The same initial fields go to every run.

Reports per storage type: MLUPS, relative L2 of u and of the apparent
viscosity against the FP64 run. With these parameters the analytic channel
velocity peaks near 1.5 lattice units and the viscous time h^2/nu is ~5.5e5
steps, so after 10,000 steps the flow is far from that profile; the analytic
comparison is made on a second, converged case: a 32-wide channel whose force
is chosen for a peak velocity of 0.02 with the same n, the same nominal-tau
rule for K and the same clamp, run for ten viscous times. There the
x/z-averaged u_x(y) is compared with the power-law and the clamped-law
profiles, and the shear stress rho nu_app du_x/dy with F s (force balance,
s the distance from the centre line).

    python -m vlbm.precision_study [--n 384] [--steps 10000]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _capi
from .lattice import Lattice

N_INDEX = 0.6
TAU_NOMINAL = 0.7
FX = 5e-6
TAU_CLAMP = (0.51, 2.0)


def powerlaw_profile(s, h, F, K, n):
    """u(s) = n/(n+1) (F/K)^(1/n) (h^((n+1)/n) - |s|^((n+1)/n))."""
    e = (n + 1.0) / n
    return n / (n + 1.0) * (F / K) ** (1.0 / n) * (h ** e - np.abs(s) ** e)


def clamped_profile(s, h, F, K, n, tau_clamp=TAU_CLAMP, m=20001):
    """Channel velocity of the power law with the apparent viscosity clamped
    to [(tau_min - 1/2)/3, (tau_max - 1/2)/3]: shear stress F|s| by force
    balance, shear rate from the clamped law, u integrated from the wall."""
    nu_lo, nu_hi = (tau_clamp[0] - 0.5) / 3.0, (tau_clamp[1] - 0.5) / 3.0
    t = np.linspace(0.0, h, m)
    stress = F * t
    gd = (stress / K) ** (1.0 / n)                       # unclamped power law
    with np.errstate(divide="ignore"):
        nu = np.where(gd > 0, K * gd ** (n - 1.0), np.inf)
    gd = np.where(nu > nu_hi, stress / nu_hi, gd)
    gd = np.where(nu < nu_lo, stress / nu_lo, gd)
    # u(t) = integral_t^h gd dt'
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (gd[1:] + gd[:-1]) * np.diff(t))])
    u = cum[-1] - cum
    return np.interp(np.abs(s), t, u)


def consistency_for(h, F=FX, n=N_INDEX, tau=TAU_NOMINAL):
    """K with K (u_max/h)^(n-1) = (tau - 1/2)/3, u_max from the analytic profile."""
    nu = (tau - 0.5) / 3.0
    K = nu
    for _ in range(200):
        umax = powerlaw_profile(0.0, h, F, K, n)
        K_new = nu / (umax / h) ** (n - 1.0)
        if abs(K_new - K) < 1e-15 * K:
            break
        K = K_new
    return K


def force_for_peak(h, umax, n=N_INDEX, tau=TAU_NOMINAL):
    """(F, K): K from the nominal-tau rule at u_max/h, F giving that peak."""
    K = (tau - 0.5) / 3.0 / (umax / h) ** (n - 1.0)
    e = (n + 1.0) / n
    F = K * (umax * (n + 1.0) / n / h ** e) ** n
    return F, K


def make_desc(n, precision, K, ny=None, nxz=None, F=FX, rounding=_capi.ROUND_STOCHASTIC):
    d = _capi.default_desc()
    d.rounding = rounding
    d.nx = d.ny = d.nz = n
    if ny is not None:
        d.ny = ny
        d.nx = d.nz = nxz
    d.q = 27
    d.precision = precision
    d.rheology = _capi.POWERLAW
    d.builder = 1            # parallelPlatesBounceBack_Y
    d.tau = TAU_NOMINAL
    d.n_index = N_INDEX
    d.k_consistency = K
    d.gamma_0 = 0.0
    d.omega_clamp = 1
    d.tau_min, d.tau_max = TAU_CLAMP
    d.fx, d.fy, d.fz = F, 0.0, 0.0
    d.rho0 = 1.0
    return d


def initial_fields(n, seed=4242):
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    return (np.ones(shape), rng.uniform(-1e-4, 1e-4, shape), rng.uniform(-1e-4, 1e-4, shape),
            rng.uniform(-1e-4, 1e-4, shape))


def run(n, steps, precision, K, init, log, rounding=_capi.ROUND_STOCHASTIC):
    lat = Lattice(make_desc(n, precision, K, rounding=rounding))
    lat.init(*init)
    lat.step(10)
    lat.sync()
    log(f"[precision_study] storage {precision} B: timing {steps} steps")
    t0 = time.perf_counter()
    lat.step(steps - 10, save_last=True)
    lat.sync()
    wall = time.perf_counter() - t0
    m = lat.macroscopics(omega=True)
    lat.close()
    return m, n ** 3 * (steps - 10) / wall / 1e6


def profiles(m, n, F=FX):
    """x/z-averaged u_x(y) and shear stress rho nu_app du_x/dy."""
    ux = m["ux"].mean(axis=(0, 2))
    om = m["omega"].mean(axis=(0, 2))
    rho = m["rho"].mean(axis=(0, 2))
    nu = (1.0 / om - 0.5) / 3.0
    dudy = np.gradient(ux)
    stress = rho * nu * dudy
    return ux, stress


def validation(precision, rounding=_capi.ROUND_STOCHASTIC, steps=64000, ny=32, nxz=4, umax=0.02):
    h = ny / 2.0
    F, K = force_for_peak(h, umax)
    lat = Lattice(make_desc(0, precision, K, ny=ny, nxz=nxz, F=F, rounding=rounding))
    lat.init()
    lat.step(steps, save_last=True)
    lat.sync()
    m = lat.macroscopics(omega=True)
    lat.close()
    y = np.arange(ny) - (ny - 1) / 2.0
    ux, st = profiles(m, ny, F)
    u_an = powerlaw_profile(y, h, F, K, N_INDEX)
    u_cl = clamped_profile(y, h, F, K, N_INDEX)
    stress_an = -F * y
    inner = np.abs(y) < h - 2
    return {"profile_rel_l2_vs_power_law": float(np.linalg.norm(ux - u_an) / np.linalg.norm(u_an)),
            "profile_rel_l2_vs_clamped_law": float(np.linalg.norm(ux - u_cl) / np.linalg.norm(u_cl)),
            "stress_rel_l2_vs_force_balance": float(np.linalg.norm(st[inner] - stress_an[inner]) /
                                                     np.linalg.norm(stress_an[inner])),
            "F_x": F, "K": K, "steps": steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=384)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=4242)
    ap.add_argument("--no-validation", action="store_true")
    a = ap.parse_args()
    log = lambda s: print(s, file=sys.stderr, flush=True)   # noqa: E731
    n = a.n
    h = n / 2.0
    K = consistency_for(h)
    init = initial_fields(n, a.seed)
    results = {}
    ref = None
    SR, RN = _capi.ROUND_STOCHASTIC, _capi.ROUND_NEAREST
    precisions = (("fp64", _capi.FP64, RN), ("fp32", _capi.FP32, RN), ("fp16_dev_stochastic", _capi.FP16, SR),
                  ("fp16_dev_nearest", _capi.FP16, RN), ("bf16_dev_stochastic", _capi.BF16, SR),
                  ("bf16_dev_nearest", _capi.BF16, RN))
    for name, prec, rnd in precisions:
        m, mlups = run(n, a.steps, prec, K, init, log, rnd)
        entry = {"mlups": round(mlups, 1), "max_ux": float(m["ux"].max())}
        if ref is None:
            ref = m
        else:
            uf = np.stack([m["ux"], m["uy"], m["uz"]])
            ur = np.stack([ref["ux"], ref["uy"], ref["uz"]])
            nu_f = (1.0 / m["omega"] - 0.5) / 3.0
            nu_r = (1.0 / ref["omega"] - 0.5) / 3.0
            entry["u_rel_l2_vs_fp64"] = float(np.linalg.norm(uf - ur) / np.linalg.norm(ur))
            entry["nu_rel_l2_vs_fp64"] = float(np.linalg.norm(nu_f - nu_r) / np.linalg.norm(nu_r))
        results[name] = entry
        log(f"[precision_study] {name}: {entry}")
        del m
    out = {"study": "D3Q27 regularized power-law channel, storage precision",
           "lattice": [n, n, n], "steps": a.steps, "n_index": N_INDEX, "K": K,
           "tau_clamp": list(TAU_CLAMP), "F_x": FX, "results": results}
    if not a.no_validation:
        out["converged_channel_32"] = {name: validation(prec, rnd) for name, prec, rnd in precisions}
        log(f"[precision_study] validation: {out['converged_channel_32']}")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
