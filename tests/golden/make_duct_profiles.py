"""Writes tests/golden/duct_profiles.json from the reference's archived duct runs.

The reference keeps, for its interpolated bounce-back validation, the setup
(`<id>_info.txt`) and the u_z profiles along the north-south line
(`<id>uzN-S<steps>.csv`) and, for the circular duct, the south-west to
north-east diagonal (`<id>uzSW-NE<steps>.csv`) of D3Q19 regularized FP64 runs:
    doc/Simulations/Analysis/Regularized/D3Q19/circularDuctInterp/data/20{0,1}*
    doc/Simulations/Analysis/Regularized/D3Q19/annularDuctInterp/data/00{0,1}*
This script copies those numbers (inputs and outputs only) into one JSON
fixture. It runs in a container holding the reference tree:

    python tests/golden/make_duct_profiles.py [/root/reference]
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
BASE = REF / "doc/Simulations/Analysis/Regularized/D3Q19"
CASES = [("circularDuctInterp", "200"), ("circularDuctInterp", "201"),
         ("annularDuctInterp", "000"), ("annularDuctInterp", "001")]
KEYS = {"NX": int, "NY": int, "NZ": int, "Tau": float, "FX": float, "FY": float, "FZ": float,
        "Nsteps": int, "Velocity set": str, "Precision": str}


def info(path: Path) -> dict:
    out = {}
    for line in path.read_text().splitlines():
        if ":" not in line:
            continue
        k, v = (s.strip() for s in line.split(":", 1))
        if k in KEYS and k not in out:
            out[k] = KEYS[k](v.split()[0])
    return out


def profile(path: Path) -> list:
    rows = []
    for line in path.read_text().split():
        i, v = line.split(",")
        rows.append([int(i), float(v)])
    return rows


def main():
    cases = []
    for folder, sid in CASES:
        d = BASE / folder / "data"
        inf = info(d / f"{sid}_info.txt")
        steps = inf["Nsteps"]
        case = {"source": f"doc/Simulations/Analysis/Regularized/D3Q19/{folder}/data/{sid}*",
                "builder": "circularDuctInterpBounceBack" if folder.startswith("circular")
                else "annularDuctInterpBounceBack",
                "info": inf, "uz_north_south": profile(d / f"{sid}uzN-S{steps}.csv")}
        diag = d / f"{sid}uzSW-NE{steps}.csv"
        if diag.exists():
            case["uz_southwest_northeast"] = profile(diag)
        cases.append(case)
    out = Path(__file__).resolve().parent / "duct_profiles.json"
    out.write_text(json.dumps({"cases": cases}, indent=1) + "\n")
    print(f"wrote {out} ({len(cases)} cases)")


if __name__ == "__main__":
    main()
