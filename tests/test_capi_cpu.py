"""CPU-side checks of the C-ABI library: symbols, host-built maps, config."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import pyoracle as po
import vlbm
from vlbm import _capi
from helpers import base_cfg

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "vlbm.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(vlbm_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    lib = C.CDLL(str(_capi.LIB_PATH))
    for n in names:
        assert hasattr(lib, n), n


def test_defaults_follow_var_h():
    d = vlbm.default_desc()
    assert (d.nx, d.ny, d.nz, d.q, d.tau, d.rho0) == (128, 128, 384, 19, 0.6, 1.0)
    assert (d.eta_p, d.s_y, d.n_index) == (0.3, 1e-4, 0.5)
    assert list(d.ux_bc) == [0, 0.1, 0, 0, 0, 0, 0, 0]
    assert list(d.uy_bc) == [0, 0.05, -0.05, 0, 0, 0, 0, 0]


SIZES = [(4, 5, 6), (1, 3, 4), (3, 1, 5), (2, 2, 2), (7, 4, 1), (16, 9, 5)]


@pytest.mark.parametrize("builder", [b for b in vlbm.BUILDERS if "Duct" not in b or "square" in b])
@pytest.mark.parametrize("size", SIZES)
def test_box_maps_bit_exact_vs_oracle(builder, size):
    nx, ny, nz = size
    cfg = base_cfg(NX=nx, NY=ny, NZ=nz, builder=builder)
    d = cfg.to_desc()
    mine = vlbm.build_map(d)
    ref = po.build_map(po.desc_from(d))
    assert np.array_equal(mine, ref)


@pytest.mark.parametrize("builder", ["circularDuctInterpBounceBack", "annularDuctInterpBounceBack"])
@pytest.mark.parametrize("n", [12, 17, 24, 33])
def test_duct_maps_bit_exact_vs_oracle(builder, n):
    cfg = base_cfg(NX=n, NY=n, NZ=3, builder=builder)
    d = cfg.to_desc()
    mine = vlbm.build_map(d)
    ref = po.build_map(po.desc_from(d))
    assert np.array_equal(mine, ref)


def test_invalid_descriptors_are_rejected():
    d = vlbm.default_desc()
    d.q = 21
    with pytest.raises(vlbm.VlbmError) as e:
        vlbm.build_map(d)
    assert e.value.code == 1 and "D3Q19 or D3Q27" in str(e.value)
    d = base_cfg(vel_set="D3Q27", builder="parallelPlatesZouHe").to_desc()
    with pytest.raises(vlbm.VlbmError) as e:
        vlbm.Lattice(d)
    assert e.value.code == 5


def test_yaml_surface(tmp_path):
    text = """
- sim_id: 1
  precision: single
  steps: 10000
  NX: 32
  NY: 32
  NZ: 64
  FX: 1e-5
  FY: 0
  FZ: 0
  tau: 0.9
  vel_set: D3Q19
  use_ibm: false
  non_newtonian_model: BINGHAM
  path_save: results
  macr_save: 1000
  data_report: 1000
  data_stop: false
  data_save: false
  pop_save: false
  ini_step: 0
  n_gpus: 1
  resid_max: 1e-5
  builder: parallelPlatesBounceBack_Y
  yield_stress: 4.0e-5
- sim_id: 2
  precision: double
"""
    f = tmp_path / "c.yaml"
    f.write_text(text)
    cfg = vlbm.read_configs(str(f))
    assert cfg.id_sim == "001" and cfg.FX == 1e-5 and cfg.resid_max == 1e-5
    d = cfg.to_desc()
    assert (d.nx, d.ny, d.nz, d.q, d.precision, d.rheology, d.builder) == (32, 32, 64, 19, 4, 1, 1)
    assert np.isclose(d.eta_p, (0.9 - 0.5) / 3) and d.s_y == 4e-5
    bad = tmp_path / "bad.yaml"
    bad.write_text(text.replace("precision: single", "precision: quad"))
    with pytest.raises(ValueError, match="Unable to load YAML"):
        vlbm.read_configs(str(bad))
