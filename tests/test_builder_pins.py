"""Node-type maps of every builder against the rules written in the reference's
builder files (src/CUDA/boundaryConditionsBuilders/*.cu), restated here as
per-node expectations and packed with the bit layout of
src/CUDA/structs/nodeTypeMap.h:15-108. Both the oracle's map and the
library's host map (vlbm_build_map) must equal them bit for bit.

Bits a builder never sets are taken as zero: the reference allocates the map
with cudaMalloc (src/CUDA/structs/populations.h:51) and never clears it, and
the driver hands out zeroed memory.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import base_cfg

DIRS = ["N", "S", "W", "E", "F", "B", "NW", "NE", "NF", "NB", "SW", "SE", "SF", "SB", "WF", "WB",
        "EF", "EB", "NWF", "NWB", "NEF", "NEB", "SWF", "SWB", "SEF", "SEB"]   # nodeTypeMap.h:60-87
SCHEMES = {"NULL": 0, "VEL_ZOUHE": 1, "VEL_BB": 2, "PRES_ZOUHE": 3, "FREE_SLIP": 4, "BB": 5,
           "INTERP_BB": 6, "SPECIAL": 7}                                         # nodeTypeMap.h:37-44


def word(used=True, save=False, scheme="NULL", d=None, ux=0, uy=0, uz=0, rho=0, interp=0):
    """nodeTypeMap.h bit fields: rho 0-2, uz 3-5, uy 6-8, ux 9-11, geometry 12
    (CONCAVE = 0), direction 13-17, scheme 18-20, used 21, save 22, unknown
    populations of the interpolated bounce-back 23-30."""
    dcode = 0 if d is None else DIRS.index(d)
    return (rho | uz << 3 | uy << 6 | ux << 9 | dcode << 13 | SCHEMES[scheme] << 18 | int(used) << 21 |
            int(save) << 22 | interp << 23)


def expected_box(builder, nx, ny, nz):
    out = np.zeros((nz, ny, nx), dtype=np.uint32)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                s, n, w, e, b, f = y == 0, y == ny - 1, x == 0, x == nx - 1, z == 0, z == nz - 1
                wd = word()
                if builder == "taylorGreenVortex":             # every node fluid, no BC
                    pass
                elif builder == "parallelPlatesBounceBack_X":  # "W, E: Wall", everything else periodic
                    if w:
                        wd = word(scheme="BB", d="W")
                    elif e:
                        wd = word(scheme="BB", d="E")
                elif builder == "parallelPlatesBounceBack_Z":  # "B, F: Wall"
                    if b:
                        wd = word(scheme="BB", d="B")
                    elif f:
                        wd = word(scheme="BB", d="F")
                elif builder == "parallelPlatesBounceBackFreeSlip":
                    # the header says "N: wall; S: symmetry" but the chain sets
                    # bounce-back at y = 0 and free slip at y = NY-1
                    if s:
                        wd = word(scheme="BB", d="S")
                    elif n:
                        wd = word(scheme="FREE_SLIP", d="N")
                elif builder == "couetteZouHe":                # "N, S: moving wall" (uz index 1 / 2)
                    if s:
                        wd = word(scheme="VEL_ZOUHE", d="S", uz=2)
                    elif n:
                        wd = word(scheme="VEL_ZOUHE", d="N", uz=1)
                elif builder == "parallelPlatesZouHe":
                    # N, S velocity Zou-He on every node of the plates; B, F
                    # pressure Zou-He (rho index 1 in, 2 out) between them; W, E periodic
                    if s:
                        wd = word(scheme="VEL_ZOUHE", d="S")
                    elif n:
                        wd = word(scheme="VEL_ZOUHE", d="N")
                    elif b:
                        wd = word(scheme="PRES_ZOUHE", d="B", rho=1)
                    elif f:
                        wd = word(scheme="PRES_ZOUHE", d="F", rho=2)
                elif builder == "inflowFreeSlip":               # "N, S: free slip; B: inflow; F: outflow"
                    if s:
                        wd = word(scheme="FREE_SLIP", d="S")
                    elif n:
                        wd = word(scheme="FREE_SLIP", d="N")
                    elif b:
                        wd = word(scheme="VEL_BB", d="B", uz=1)
                    elif f:
                        wd = word(scheme="VEL_BB", d="F", uz=1)
                elif builder == "lidDrivenBounceBack":
                    # "N: ux=UMAX; W,E,S: wall; B,F: periodic"; the x-y edges
                    # (all z) are bounce-back with the edge's direction, the
                    # lid is velocity bounce-back with ux index 1
                    if s and w:
                        wd = word(scheme="BB", d="SW")
                    elif s and e:
                        wd = word(scheme="BB", d="SE")
                    elif n and w:
                        wd = word(scheme="BB", d="NW")
                    elif n and e:
                        wd = word(scheme="BB", d="NE")
                    elif s:
                        wd = word(scheme="BB", d="S")
                    elif n:
                        wd = word(scheme="VEL_BB", d="N", ux=1)
                    elif w:
                        wd = word(scheme="BB", d="W")
                    elif e:
                        wd = word(scheme="BB", d="E")
                else:
                    raise KeyError(builder)
                out[z, y, x] = wd
    return out


BOX = ["taylorGreenVortex", "parallelPlatesBounceBack_X", "parallelPlatesBounceBack_Z",
       "parallelPlatesBounceBackFreeSlip", "couetteZouHe", "parallelPlatesZouHe", "inflowFreeSlip",
       "lidDrivenBounceBack"]


def _maps(builder, nx, ny, nz, **kw):
    import vlbm
    cfg = base_cfg(NX=nx, NY=ny, NZ=nz, builder=builder, **kw)
    return po.build_map(po.desc_from(cfg.to_desc())), vlbm.build_map(cfg.to_desc())


@pytest.mark.parametrize("builder", BOX)
@pytest.mark.parametrize("shape", [(4, 5, 6), (7, 3, 5)])
def test_box_builders_follow_their_files(builder, shape):
    nx, ny, nz = shape
    exp = expected_box(builder, nx, ny, nz)
    orc, lib = _maps(builder, nx, ny, nz)
    np.testing.assert_array_equal(orc, exp)
    np.testing.assert_array_equal(lib, exp)


# ---------------------------------------------------------------------------
# duct builders: distances with distPoints2D (globalFunctions.h:158-162, the
# first difference cast to float) and thresholds R - sqrt((float)2),
# R - 2*sqrt((float)2) (float arithmetic, then widened)
# ---------------------------------------------------------------------------
SQ2 = float(np.sqrt(np.float32(2)))
SQ2X2 = float(np.float32(2) * np.float32(np.sqrt(np.float32(2))))
D2Q9 = [(1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]


def dist2(x1, y1, x2, y2):
    return float(np.sqrt(float(np.float32(x1 - x2)) * (x1 - x2) + (y1 - y2) * (y1 - y2)))


def unknown_bits(xn, yn, nx, ny, xc, yc, outside):
    bits = 0
    for i, (dx, dy) in enumerate(D2Q9):
        xa, ya = xn - dx, yn - dy
        if not (xa < nx and xa > 0 and ya < ny and ya > 0) or outside(dist2(xa, ya, xc, yc)):
            bits |= 1 << i
    return bits


def expected_duct(builder, nx, ny, nz, R, r):
    """circularDuctInterpBounceBack.cu:35-121, annularDuctInterpBounceBack.cu,
    parallelPlatesRheometer.cu: nodes at (x + 1/2, y + 1/2) against the centre
    ((NX-1)/2 + 1/2, (NY-1)/2 + 1/2) (var.h:144-145); the wall ring of
    sqrt(2) is interpolated bounce-back (direction N outside, S inside,
    setIsInsideNodeInterpoBB), the next ring saves post-collision values."""
    xc, yc = (nx - 1) / 2.0 + 0.5, (ny - 1) / 2.0 + 0.5
    out = np.zeros((nz, ny, nx), dtype=np.uint32)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                xn, yn = x + 0.5, y + 0.5
                dn = dist2(xn, yn, xc, yc)
                if builder == "circularDuctInterpBounceBack":
                    # no axial direction; the unknown bits are set on both rings
                    ring_bits = unknown_bits(xn, yn, nx, ny, xc, yc, lambda da: da > R)
                    if dn > R:
                        w = word(used=False)
                    elif dn > R - SQ2:
                        w = word(scheme="INTERP_BB", d="N", interp=ring_bits)
                    elif dn > R - SQ2X2:
                        w = word(save=True, interp=ring_bits)
                    else:
                        w = word()
                elif builder == "annularDuctInterpBounceBack":
                    bits = unknown_bits(xn, yn, nx, ny, xc, yc, lambda da: da > R or da < r)
                    if dn > R or dn < r:
                        w = word(used=False, d="F", interp=bits)
                    elif dn > R - SQ2:
                        w = word(scheme="INTERP_BB", d="N", interp=bits)
                    elif dn > R - SQ2X2:
                        w = word(save=True, d="F", interp=bits)
                    elif dn < r + SQ2:
                        w = word(scheme="INTERP_BB", d="S", interp=bits)
                    elif dn < r + SQ2X2:
                        w = word(save=True, d="F", interp=bits)
                    else:
                        w = word(d="F", interp=bits)
                else:   # parallelPlatesRheometer: planes 0 and NZ-1 unused, plates at 1 and NZ-2
                    plate = "B" if z == 1 else ("F" if z == nz - 2 else None)
                    sch = "SPECIAL" if plate else "NULL"
                    if dn > R or z == 0 or z == nz - 1:
                        w = word(used=False)
                    elif dn > R - SQ2:
                        w = word(scheme="INTERP_BB", d="N",
                                 interp=unknown_bits(xn, yn, nx, ny, xc, yc, lambda da: da > R))
                    elif dn > R - SQ2X2:
                        w = word(save=True, scheme=sch, d=plate,
                                 interp=unknown_bits(xn, yn, nx, ny, xc, yc, lambda da: da > R))
                    else:
                        w = word(scheme=sch, d=plate)
                out[z, y, x] = w
    return out


@pytest.mark.parametrize("builder,shape,radii", [
    ("circularDuctInterpBounceBack", (17, 17, 2), (8.0, 0.0)),
    ("circularDuctInterpBounceBack", (20, 14, 2), (5.5, 0.0)),
    ("annularDuctInterpBounceBack", (24, 24, 2), (10.5, 10.5 / 4)),
    ("annularDuctInterpBounceBack", (17, 17, 1), (8.0, 2.0)),
    ("parallelPlatesRheometer", (18, 18, 6), (7.5, 0.0)),
    ("parallelPlatesRheometer", (20, 20, 5), (8.5, 0.0)),
])
def test_duct_builders_follow_their_files(builder, shape, radii):
    nx, ny, nz = shape
    R, r = radii
    exp = expected_duct(builder, nx, ny, nz, R, r)
    orc, lib = _maps(builder, nx, ny, nz, duct_outer_radius=R, duct_inner_radius=r)
    np.testing.assert_array_equal(orc, exp)
    np.testing.assert_array_equal(lib, exp)
