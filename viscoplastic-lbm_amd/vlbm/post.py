"""Reading and exporting a simulation folder, as the reference's
post-processing does (src/Post Processing/fileTreat.py, dataSave.py).

* ``SimFolder(path)``: the ``ID_info.txt`` keys (getSimInfo,
  fileTreat.py:56-143), the saved steps (getMacrSteps, fileTreat.py:37-53) and
  the binary fields of one step as ``[x][y][z]`` arrays (readFileMacr3D /
  getMacrsFromStep, fileTreat.py:146-200: raw native-endian values of the
  info file's precision, ``[NZ_TOTAL][NY][NX]`` on disk, axes swapped on read).
* ``save_vtk3d``: the fields of one step on a rectilinear grid in VTK's XML
  format (``.vtr``), point- or cell-centred, with the grid spacing rule of
  saveVTK3D (dataSave.py:5-48). The reference calls pyevtk's gridToVTK, which
  is not installed here; this writer produces the same grid from the VTK XML
  file-format specification (RectilinearGrid, raw appended data with UInt64
  block headers, little endian).
* ``save_macr_csv``: saveMacrCsv (dataSave.py:51-74).
"""
from __future__ import annotations

import glob
import os
import re
import sys
from typing import Dict, List, Optional

import numpy as np

MACR_NAMES = ["ux", "uy", "uz", "rho"]        # fileTreat.py:8


class SimFolder:
    """One simulation's output folder (``PATH/ID/``)."""

    def __init__(self, path: str, macr_names: Optional[List[str]] = None):
        self.path = path
        self.macr_names = list(macr_names or MACR_NAMES)
        self._info: Dict = {}

    def info(self) -> Dict:
        """Keys of the first ``*info*.txt``: the last token of the first line
        containing the key, as getSimInfo reads it (NZ is matched as 'NZ:')."""
        if self._info:
            return self._info
        files = sorted(glob.glob(os.path.join(self.path, "*info*.txt")))
        if not files:
            raise FileNotFoundError(f"no info file in {self.path}")
        with open(files[0]) as f:
            lines = [l.strip() for l in f.readlines()]
        keys = [("ID", "Simulation ID", str), ("Prc", "Precision", str), ("NX", "NX", int), ("NY", "NY", int),
                ("NZ", "NZ:", int), ("NZ_TOTAL", "NZ_TOTAL", int), ("Tau", "Tau", float), ("Umax", "Umax", float),
                ("Nsteps", "Nsteps", int), ("FX", "FX", float), ("FY", "FY", float), ("FZ", "FZ", float)]
        for key, match, conv in keys:
            hits = [l.split(" ")[-1] for l in lines if match in l]
            if hits:
                try:
                    self._info[key] = conv(hits[0])
                except ValueError:
                    print(f"Not able to get {key} from info file", file=sys.stderr)
            else:
                print(f"Not able to get {key} from info file", file=sys.stderr)
        return self._info

    def filenames(self, macr: str) -> List[str]:
        return sorted(glob.glob(os.path.join(self.path, "*" + macr + "*.bin")))

    def steps(self) -> List[int]:
        """Steps of the files of the first field name (the digits after it)."""
        name = self.macr_names[0]
        out = set()
        for fn in self.filenames(name):
            tail = os.path.basename(fn).split(name)[-1][:-4]
            if re.fullmatch(r"\d+", tail):
                out.add(int(tail))
        return sorted(out)

    def dtype(self):
        prc = self.info()["Prc"]
        if prc == "double":
            return np.float64
        if prc == "float":
            return np.float32
        raise ValueError(f"unknown precision {prc!r}")

    def read_macr_3d(self, filename: str) -> np.ndarray:
        """File -> array indexed [x][y][z]."""
        info = self.info()
        vec = np.fromfile(filename, dtype=self.dtype())
        vec3d = vec.reshape((info["NZ_TOTAL"], info["NY"], info["NX"]), order="C")
        return np.swapaxes(vec3d, 0, 2)

    def macrs_from_step(self, step: int) -> Dict[str, np.ndarray]:
        """Every field name's file of the step (name followed by the step
        zero-padded to 6 digits)."""
        out = {}
        for name in self.macr_names:
            want = "%s%06d.bin" % (name, step)
            for fn in self.filenames(name):
                if os.path.basename(fn).endswith(want):
                    out[name] = self.read_macr_3d(fn)
                    break
        return out


def _grid(n: int, norm: float, points: bool, dtype) -> np.ndarray:
    d = 1.0 / norm
    last = (n - 1) / norm if points else n / norm
    return np.arange(0, last + 0.1 * d, d, dtype=dtype)


def save_vtk3d(macrs: Dict[str, np.ndarray], filename: str, info: Dict, points: bool = True,
               norm_val: float = 1) -> str:
    """Writes ``filename.vtr``; arrays are [x][y][z] as SimFolder returns them.
    norm_val = 0 normalises distances by NX (NX - 1 for point data)."""
    if norm_val == 0:
        norm_val = info["NX"] - (1 if points else 0)
    dtype = np.float64 if info["Prc"] == "double" else np.float32
    x = _grid(info["NX"], norm_val, points, dtype)
    y = _grid(info["NY"], norm_val, points, dtype)
    z = _grid(info["NZ_TOTAL"], norm_val, points, dtype)
    nx, ny, nz = len(x) - (0 if points else 1), len(y) - (0 if points else 1), len(z) - (0 if points else 1)
    for k, v in macrs.items():
        if v.shape != (nx, ny, nz):
            raise ValueError(f"{k}: shape {v.shape}, grid expects {(nx, ny, nz)}")
    vtk_type = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32"}
    blocks: List[bytes] = []
    offset = 0
    arrays_xml = []

    def add(name: str, data: np.ndarray) -> str:
        nonlocal offset
        raw = np.ascontiguousarray(data).tobytes(order="C")
        blocks.append(np.uint64(len(raw)).tobytes() + raw)
        xml = (f'<DataArray type="{vtk_type[data.dtype]}" Name="{name}" NumberOfComponents="1" '
               f'format="appended" offset="{offset}"/>')
        offset += 8 + len(raw)
        return xml

    coords = [add("x_coordinates", x), add("y_coordinates", y), add("z_coordinates", z)]
    for k, v in macrs.items():
        # VTK orders points x fastest: the [x][y][z] array in Fortran order
        arrays_xml.append(add(k, np.asarray(v).ravel(order="F")))
    ext = f"0 {len(x) - 1} 0 {len(y) - 1} 0 {len(z) - 1}"
    data_tag = "PointData" if points else "CellData"
    head = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
        f'<RectilinearGrid WholeExtent="{ext}">\n'
        f'<Piece Extent="{ext}">\n'
        '<Coordinates>\n' + "\n".join(coords) + '\n</Coordinates>\n'
        f'<{data_tag}>\n' + "\n".join(arrays_xml) + f'\n</{data_tag}>\n'
        '</Piece>\n</RectilinearGrid>\n<AppendedData encoding="raw">\n_')
    out = filename + ".vtr"
    with open(out, "wb") as f:
        f.write(head.encode())
        for b in blocks:
            f.write(b)
        f.write(b'\n</AppendedData>\n</VTKFile>\n')
    return out


def read_vtr(filename: str) -> Dict[str, np.ndarray]:
    """Arrays of a ``.vtr`` written by save_vtk3d (raw appended, UInt64 headers)."""
    with open(filename, "rb") as f:
        data = f.read()
    start = data.index(b'<AppendedData encoding="raw">')
    mark = data.index(b"_", start) + 1
    head = data[:start].decode()
    out = {}
    for m in re.finditer(r'<DataArray type="(Float32|Float64)" Name="([^"]+)"[^>]*offset="(\d+)"', head):
        t = np.float32 if m.group(1) == "Float32" else np.float64
        off = mark + int(m.group(3))
        n = int(np.frombuffer(data[off:off + 8], dtype=np.uint64)[0])
        out[m.group(2)] = np.frombuffer(data[off + 8:off + 8 + n], dtype=t).copy()
    ext = re.search(r'WholeExtent="([^"]+)"', head).group(1).split()
    out["_extent"] = np.array([int(e) for e in ext])
    out["_cell"] = "<CellData>" in head
    return out


def save_macr_csv(filename: str, macr: np.ndarray, normalize_dist: bool = False) -> None:
    """1D: rows 'position,value' (position as %d, or i/len as %.6e when
    normalised); 2D: the matrix as CSV."""
    macr = np.asarray(macr)
    with open(filename, "w") as f:
        if macr.ndim == 1:
            if not normalize_dist:
                arr = [[i, macr[i]] for i in range(len(macr))]
                np.savetxt(f, arr, fmt=["%d", "%.6e"], delimiter=",")
            else:
                arr = [(i / len(macr), macr[i]) for i in range(len(macr))]
                np.savetxt(f, arr, fmt=["%.6e", "%.6e"], delimiter=",")
        elif macr.ndim == 2:
            np.savetxt(f, macr, delimiter=",")
        else:
            raise ValueError("saveMacrCsv takes a 1D or 2D array")


def export_folder_vtk(path: str, out_prefix: Optional[str] = None, points: bool = True,
                      macr_names: Optional[List[str]] = None) -> List[str]:
    """Every saved step of a folder to ``<ID>macr<step:06d>.vtr`` (exampleVTK.py)."""
    sim = SimFolder(path, macr_names)
    info = sim.info()
    written = []
    for step in sim.steps():
        macr = sim.macrs_from_step(step)
        name = (out_prefix or os.path.join(path, info["ID"])) + "macr" + str(step).zfill(6)
        written.append(save_vtk3d(macr, name, info, points=points))
    return written
