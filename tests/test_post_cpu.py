"""Output consumers on the CPU: reading a simulation folder the way
src/Post Processing/fileTreat.py does, the VTK export of dataSave.py:5-48 and
the CSV writer, on synthetic folders written here in the reference's format
(lbmReport.cpp:55-172 file names and raw layout, getSimInfoString key lines)."""
import os

import numpy as np
import pytest

from vlbm import post
from vlbm.runner import var_filename


def make_folder(tmp_path, prc="float", nx=5, ny=4, nz=3, steps=(0, 100, 2000)):
    sim = "007"
    folder = tmp_path / sim
    folder.mkdir()
    t = np.float32 if prc == "float" else np.float64
    with open(folder / f"{sim}_info.txt", "w") as f:
        f.write("---------------------------- SIMULATION INFORMATION ----------------------------\n")
        f.write(f"      Simulation ID: {sim}\n       Velocity set: D3Q19\n          Precision: {prc}\n")
        f.write(f"                 NX: {nx}\n                 NY: {ny}\n                 NZ: {nz}\n")
        f.write(f"           NZ_TOTAL: {nz}\n                Tau: 9.000000e-01\n")
        f.write("               Umax: 1.000000e-01\n                 FX: 1.000000e-05\n")
        f.write("                 FY: 0.000000e+00\n                 FZ: 0.000000e+00\n")
        f.write("             Nsteps: 2000\n")
    rng = np.random.default_rng(3)
    fields = {}
    for s in steps:
        for name in ("rho", "ux", "uy", "uz"):
            a = rng.standard_normal((nz, ny, nx)).astype(t)          # [z][y][x] on disk
            a.tofile(var_filename(str(tmp_path), sim, name, s, ".bin"))
            fields[(s, name)] = a
    return str(folder), fields


@pytest.mark.parametrize("prc", ["float", "double"])
def test_sim_folder_reads_like_file_treat(tmp_path, prc):
    folder, fields = make_folder(tmp_path, prc)
    sim = post.SimFolder(folder)
    info = sim.info()
    assert info["ID"] == "007" and info["Prc"] == prc
    assert (info["NX"], info["NY"], info["NZ"], info["NZ_TOTAL"]) == (5, 4, 3, 3)
    assert info["Tau"] == 0.9 and info["FX"] == 1e-5 and info["Nsteps"] == 2000
    assert sim.steps() == [0, 100, 2000]
    m = sim.macrs_from_step(100)
    assert set(m) == {"ux", "uy", "uz", "rho"}
    for name, arr in m.items():
        assert arr.shape == (5, 4, 3)                     # [x][y][z]
        src = fields[(100, name)]
        assert arr[2, 1, 0] == src[0, 1, 2]
        assert np.array_equal(arr, np.swapaxes(src, 0, 2))


@pytest.mark.parametrize("points", [True, False])
def test_vtk_export_grid_and_data(tmp_path, points):
    folder, fields = make_folder(tmp_path, "double")
    sim = post.SimFolder(folder)
    info = sim.info()
    m = sim.macrs_from_step(2000)
    if not points:
        m = {k: v for k, v in m.items()}
    out = post.save_vtk3d(m, str(tmp_path / "case"), info, points=points, norm_val=0)
    back = post.read_vtr(out)
    nx, ny, nz = 5, 4, 3
    if points:
        # point data: NX points spaced 1/(NX-1) so that x runs over [0, 1]
        assert np.allclose(back["x_coordinates"], np.linspace(0, 1, nx))
        assert np.allclose(back["y_coordinates"], np.arange(ny) / (nx - 1))
        assert list(back["_extent"]) == [0, nx - 1, 0, ny - 1, 0, nz - 1]
        assert not back["_cell"]
    else:
        assert np.allclose(back["x_coordinates"], np.arange(nx + 1) / nx)
        assert list(back["_extent"]) == [0, nx, 0, ny, 0, nz]
        assert back["_cell"]
    # VTK orders x fastest: the file's own [z][y][x] bytes
    for name in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(back[name], fields[(2000, name)].ravel())


def test_vtk_rejects_mismatched_shapes(tmp_path):
    folder, _ = make_folder(tmp_path, "float")
    info = post.SimFolder(folder).info()
    with pytest.raises(ValueError):
        post.save_vtk3d({"rho": np.zeros((4, 4, 3), np.float32)}, str(tmp_path / "x"), info)


def test_export_folder_names_steps(tmp_path):
    folder, _ = make_folder(tmp_path, "float", steps=(10, 20))
    out = post.export_folder_vtk(folder)
    assert [os.path.basename(p) for p in out] == ["007macr000010.vtr", "007macr000020.vtr"]


def test_macr_csv(tmp_path):
    fn = str(tmp_path / "a.csv")
    post.save_macr_csv(fn, np.array([1.0, 2.5, -3e-7]))
    assert open(fn).read().splitlines() == ["0,1.000000e+00", "1,2.500000e+00", "2,-3.000000e-07"]
    post.save_macr_csv(fn, np.array([1.0, 2.0]), normalize_dist=True)
    assert open(fn).read().splitlines() == ["0.000000e+00,1.000000e+00", "5.000000e-01,2.000000e+00"]
    post.save_macr_csv(fn, np.eye(2))
    assert np.array_equal(np.loadtxt(fn, delimiter=","), np.eye(2))
    with pytest.raises(ValueError):
        post.save_macr_csv(fn, np.zeros((2, 2, 2)))


def test_var_filename_padding():
    """getVarFilename: zero-padded to 6 digits; step 0 gets six zeros in front."""
    assert var_filename("P", "000", "ux", 0, ".bin") == os.path.join("P", "000", "000_ux0000000.bin")
    assert var_filename("P", "000", "ux", 7, ".bin").endswith("000_ux000007.bin")
    assert var_filename("P", "000", "avgUz", 123456, ".csv").endswith("000_avgUz123456.csv")
    assert var_filename("P", "000", "ux", 1234567, ".bin").endswith("000_ux1234567.bin")
