"""Shared test helpers: configurations, velocity set of the reference, norms."""
from __future__ import annotations

import numpy as np

# Velocity set of the reference (src/CUDA/velocitySets/D3Q19.h:7-29 and
# D3Q27.h:7-36 comment tables): one row per population, (cx, cy, cz).
C27 = np.array([
    (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, -1), (0, 1, 1), (0, -1, -1),
    (1, -1, 0), (-1, 1, 0), (1, 0, -1), (-1, 0, 1), (0, 1, -1), (0, -1, 1),
    (1, 1, 1), (-1, -1, -1), (1, 1, -1), (-1, -1, 1), (1, -1, 1), (-1, 1, -1), (-1, 1, 1), (1, -1, -1),
], dtype=np.int64)


def cvec(q: int) -> np.ndarray:
    return C27[:q]


def weights(q: int) -> np.ndarray:
    if q == 19:
        return np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    return np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)


def opp(i: int) -> int:
    return 0 if i == 0 else (i + 1 if i % 2 else i - 1)


def feq(q, rho, u):
    """Second-order equilibrium w rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u), closed form."""
    c, w = cvec(q), weights(q)
    cu = c @ np.asarray(u, dtype=float)
    return w * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * np.dot(u, u))


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def base_cfg(**kw):
    """SimConfig with pyvisc's required keys filled in; keyword overrides."""
    from vlbm.config import SimConfig
    cfg = dict(precision="double", NX=8, NY=8, NZ=8, tau=0.9, vel_set="D3Q19", use_ibm=False,
               non_newtonian_model=None, sim_id=0, path_save="TEST", steps=100, macr_save=0,
               data_report=0)
    cfg.update(kw)
    return SimConfig(**cfg)


def seeded_velocity(shape, u0, amp, seed):
    """u = u0 + U(-amp, amp) per component from numpy's PCG64 (seeded)."""
    rng = np.random.default_rng(seed)
    ux = u0[0] + rng.uniform(-amp, amp, size=shape)
    uy = u0[1] + rng.uniform(-amp, amp, size=shape)
    uz = u0[2] + rng.uniform(-amp, amp, size=shape)
    return ux, uy, uz
