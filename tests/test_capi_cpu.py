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


@pytest.mark.parametrize("builder", [b for b in vlbm.BUILDERS
                                     if ("Duct" not in b and "Rheometer" not in b) or "square" in b])
@pytest.mark.parametrize("size", SIZES)
def test_box_maps_bit_exact_vs_oracle(builder, size):
    nx, ny, nz = size
    cfg = base_cfg(NX=nx, NY=ny, NZ=nz, builder=builder)
    d = cfg.to_desc()
    mine = vlbm.build_map(d)
    ref = po.build_map(po.desc_from(d))
    assert np.array_equal(mine, ref)


@pytest.mark.parametrize("builder", ["circularDuctInterpBounceBack", "annularDuctInterpBounceBack",
                                     "parallelPlatesRheometer"])
@pytest.mark.parametrize("n", [12, 17, 24, 33])
@pytest.mark.parametrize("nz", [3, 6])
def test_duct_maps_bit_exact_vs_oracle(builder, n, nz):
    """Circle masks, interpolated walls and, for the rheometer, the unused
    end planes (z = 0, NZ-1) and the plate planes next to them."""
    cfg = base_cfg(NX=n, NY=n, NZ=nz, builder=builder)
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


@pytest.mark.parametrize("kw,code", [(dict(vel_set="D3Q27"), 5), (dict(layout="moments"), 5), (dict(NZ=2), 1),
                                     (dict(NX=17, NY=17), 1)])
def test_rheometer_refusals(kw, code):
    """The plate condition is D3Q19 (parallelPlatesRheometer.cu defines no
    D3Q27 lines), reads populations of the idle lattice (not kept by the
    coefficient layout) and needs the two unused planes plus a fluid plane.
    An odd x odd cross-section puts a plate node on the axis, where the
    reference's wall velocity is 0/0."""
    cfg = dict(NX=12, NY=12, NZ=5, builder="parallelPlatesRheometer")
    cfg.update(kw)
    with pytest.raises(vlbm.VlbmError) as e:
        vlbm.Lattice(base_cfg(**cfg).to_desc())
    assert e.value.code == code and "rheometer" in str(e.value)


def test_yaml_surface_of_the_reference_example():
    """The reference's own example (examples/all_options.yaml, committed as
    tests/golden/all_options.yaml) reads into the descriptor pyvisc would
    render into var.h: ID 001, FP32, 32^3 D3Q19 at tau 0.9, no force,
    Newtonian, the default builder, the report cadence and residual."""
    from pathlib import Path
    cfg = vlbm.read_configs(str(Path(__file__).resolve().parent / "golden" / "all_options.yaml"))
    assert cfg.id_sim == "001" and cfg.steps == 10000 and cfg.macr_save == 1000 and cfg.data_report == 1000
    assert cfg.resid_max == 1e-5 and cfg.path_save == "results" and not cfg.data_stop and cfg.ini_step == 0
    d = cfg.to_desc()
    assert (d.nx, d.ny, d.nz, d.q, d.precision, d.rheology) == (32, 32, 32, 19, 4, 0)
    assert (d.fx, d.fy, d.fz) == (0.0, 0.0, 0.0) and np.isclose(d.tau, 0.9)


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


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors have the sizes and field offsets of include/vlbm.h."""
    import subprocess
    from vlbm import serve
    structs = {"vlbm_desc": _capi.VlbmDesc, "vlbm_ibm_desc": _capi.VlbmIbmDesc,
               "vlbm_batch_desc": serve.VlbmBatchDesc, "vlbm_job_info": serve.VlbmJobInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vlbm.h"', 'int main(void) {']
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.split("\n") if l)
    for cname, cls in structs.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_ibm_and_batch_descriptor_defaults():
    d = _capi.default_ibm_desc()
    # ibmVar.h / ibmBoundaryCondition.h constants
    assert (d.max_iterations, d.thickness, d.alpha, d.fluid_density) == (1, 1.0, 0.5, 1.0)
    assert (d.periodic_x, d.periodic_y, d.periodic_z) == (1, 0, 1)
    assert (d.pp_friction, d.pw_friction, d.pp_rest, d.pw_rest) == (0.1, 0.1, 0.9, 0.9)
    assert (d.young, d.poisson, d.wall_young, d.wall_poisson) == (0.5, 0.3, 100.0, 0.3)
    from vlbm import serve
    b = serve.batch_desc()
    assert b.q == 19 and b.precision == _capi.FP32 and b.use_graph == 1 and b.check_every % 2 == 0


@pytest.mark.parametrize("kw,msg", [
    (dict(max_jobs=0), "max_jobs"),
    (dict(check_every=7), "check_every"),
    (dict(check_every=100, steps_per_launch=30), "steps_per_launch"),
    (dict(block=64), "block"),
    (dict(q=15), "velocity set"),
])
def test_batch_rejects_bad_descriptors(kw, msg):
    """Checked before any device call, so this runs without a GPU."""
    from vlbm import serve
    with pytest.raises(vlbm.VlbmError, match=msg):
        serve.Batch(**kw)


def test_sweep_job_generator_and_bingham_number():
    from vlbm import serve
    jobs = serve.make_jobs(50, rate=20.0, seed=3)
    arr = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(arr) > 0)
    assert 1.0 < arr[-1] < 5.0                       # ~50 / 20 s
    for j in jobs:
        assert 32 <= j.nx <= 256 and 32 <= j.ny <= 128 and 16 <= j.nz <= 128
        assert 0.55 <= j.tau <= 1.5 and 1e-6 <= j.fx <= 1e-5 and 0 <= j.bn <= 5
        h = j.ny / 2.0
        nu = (j.tau - 0.5) / 3.0
        assert j.tau_y < j.fx * h                    # the plug stays inside the channel
        if j.bn > 0:
            U = serve.mean_velocity(h, j.fx, nu, j.tau_y)
            assert abs(j.tau_y * 2 * h / (nu * U) - j.bn) < 1e-6 * max(1.0, j.bn)
    # Newtonian limit: the profile is the Poiseuille parabola
    s = np.linspace(-4, 4, 9)
    assert np.allclose(serve.bingham_profile(s, 4.0, 1e-5, 0.1, 0.0), 1e-5 / 0.2 * (16 - s * s))


def test_reference_particle_center_layout():
    """The checkpoint's ParticleCenter record: 8 dfloat3, 2 dfloat4, 4 dfloat3,
    dfloat6, 2 dfloat3, 4 dfloat, bool, then CollisionData (int, 3 dfloat3,
    int[13], dfloat3[13], int[13]); offsets by hand for both dfloat types."""
    from checkpoint_layout import particle_center_dtype
    for t, s, size in ((np.float64, 8, 992), (np.float32, 4, 544)):
        d = particle_center_dtype(t)
        assert d.itemsize == size
        assert d.fields["q_pos"][1] == 24 * s
        assert d.fields["I"][1] == 44 * s
        assert d.fields["S"][1] == 56 * s
        assert d.fields["movable"][1] == 60 * s
        assert d.fields["collision"][1] == 61 * s    # bool padded to the next dfloat boundary
        c = d.fields["collision"][0]
        assert c.fields["semiAxis"][1] == s          # int padded to a dfloat
        assert c.fields["partner"][1] == s + 9 * s
        assert c.fields["tang"][1] == -(-(10 * s + 52) // s) * s
        assert c.fields["last"][1] == c.fields["tang"][1] + 39 * s


def test_checkpoint_keys_in_config():
    from helpers import base_cfg
    cfg = base_cfg(checkpoint_save=500, load_checkpoint="RUNS")
    assert cfg.checkpoint_save == 500 and cfg.load_checkpoint == "RUNS"
    with pytest.raises(ValueError):
        base_cfg(checkpoint_save=-1)
