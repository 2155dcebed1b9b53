"""vlbm_get_pops_noalloc (no device allocation; its values pass through one
fixed device array per GPU in chunks of nodes) against vlbm_get_pops (which
stages them in buffers it allocates) on the same lattice: equal bit for bit,
on two calls in a row, the second on a smaller lattice. The larger lattices
span several chunks (55,040 nodes for D3Q19, 38,656 for D3Q27); each layout,
storage precision, the rheometer (whose vlbm_get_pops materializes into a
buffer of its own) and the state right after init / load (read in place) are
covered."""
import ctypes as C

import numpy as np
import pytest

import vlbm
from helpers import base_cfg, seeded_velocity
from vlbm._capi import check, lib

pytestmark = pytest.mark.gpu


def announce(name):
    print(f"\n[gpu-case] {name}", flush=True)


def pops_noalloc(lat):
    nz, ny, nx = lat.shape
    out = np.full((lat.q, nz + 1, ny, nx), np.nan)
    check(lib.vlbm_get_pops_noalloc(lat._h, out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def make(kw, steps, noise=1e-3, seed=5):
    cfg = base_cfg(**kw)
    shape = (cfg.NZ, cfg.NY, cfg.NX)
    lat = vlbm.Lattice(cfg.to_desc())
    lat.init(np.ones(shape), *seeded_velocity(shape, (0.0, 0.0, 0.005), noise, seed))
    if steps:
        lat.step(steps, save_last=True)
    lat.sync()
    return lat


DUCT = dict(tau=0.8, builder="squareDuctZouHe", non_newtonian_model="BINGHAM", yield_stress=1e-5,
            uz_bc=[0, 0.01, 0, 0, 0, 0, 0, 0])
PAIRS = [
    # name, larger lattice (several chunks), smaller lattice, steps
    ("d3q19_fp32_bingham_duct", dict(precision="single", NX=64, NY=48, NZ=40, **DUCT),
     dict(precision="single", NX=16, NY=16, NZ=12, **DUCT), 30),
    ("d3q27_fp64_power_law_moments", dict(precision="double", NX=40, NY=36, NZ=30, vel_set="D3Q27", tau=0.7,
                                          FX=5e-6, builder="parallelPlatesBounceBack_Y", layout="moments",
                                          non_newtonian_model="POWERLAW", n_index=0.6, k_consistency=0.02,
                                          tau_clamp=[0.51, 2.0]),
     dict(precision="double", NX=8, NY=8, NZ=8, vel_set="D3Q27", tau=0.7, FX=5e-6, layout="moments",
          builder="parallelPlatesBounceBack_Y", non_newtonian_model="POWERLAW", n_index=0.6, k_consistency=0.02,
          tau_clamp=[0.51, 2.0]), 20),
    ("rheometer_fp64", dict(precision="double", NX=40, NY=40, NZ=40, tau=0.7, builder="parallelPlatesRheometer",
                            outer_rotation=0.002),
     dict(precision="double", NX=18, NY=18, NZ=6, tau=0.7, builder="parallelPlatesRheometer", outer_rotation=0.002),
     25),
    ("fp16_plates", dict(precision="half", NX=64, NY=64, NZ=16, tau=0.8, FX=1e-5,
                         builder="parallelPlatesBounceBack_Y"),
     dict(precision="half", NX=8, NY=16, NZ=4, tau=0.8, FX=1e-5, builder="parallelPlatesBounceBack_Y"), 15),
    ("circular_duct_fp32", dict(precision="single", NX=48, NY=48, NZ=30, tau=0.8, FZ=1e-5,
                                builder="circularDuctInterpBounceBack"),
     dict(precision="single", NX=14, NY=14, NZ=5, tau=0.8, FZ=1e-5, builder="circularDuctInterpBounceBack"), 20),
]


@pytest.mark.parametrize("name,big,small,steps", PAIRS, ids=[p[0] for p in PAIRS])
def test_noalloc_pops_equal_get_pops_twice(name, big, small, steps):
    announce(f"noalloc pops: {name}")
    a = make(big, steps)
    b = make(small, steps, seed=6)
    na = pops_noalloc(a)          # first call on the larger lattice
    nb = pops_noalloc(b)          # second call, smaller lattice: the array keeps the first call's values
    assert np.array_equal(na, a.get_pops())
    assert np.array_equal(nb, b.get_pops())
    assert np.isfinite(na).all() and np.isfinite(nb).all()
    a.close()
    b.close()


@pytest.mark.parametrize("name,big,small,steps", PAIRS[2:3] + PAIRS[4:5], ids=[PAIRS[2][0], PAIRS[4][0]])
def test_noalloc_pops_of_a_state_read_in_place(name, big, small, steps):
    """Right after init, and right after a load, the state is read in place
    (no gather, no boundary pass): both entry points return it as given."""
    announce(f"noalloc pops in place: {name}")
    a = make(big, 0)
    na = pops_noalloc(a)
    assert np.array_equal(na, a.get_pops())
    b = make(small, steps, seed=6)
    loaded = b.get_pops()
    b.load_pops(loaded)
    nb = pops_noalloc(b)
    assert np.array_equal(nb, b.get_pops())
    assert np.array_equal(nb[:, :-1], loaded[:, :-1])
    a.close()
    b.close()
