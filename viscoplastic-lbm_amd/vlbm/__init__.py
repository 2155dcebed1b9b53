"""MI355X-native VISCOPLASTIC-LBM time-step library (host-side Python mirror).

The compute path is libvlbm.so (HIP, gfx950) behind the C-ABI of
include/vlbm.h; this package binds it, reads pyvisc-style YAML and provides
the driver loop of the reference's main.cu.
"""
from . import _capi
from ._capi import VlbmDesc, VlbmError, default_desc, FP64, FP32, FP16, BF16, NEWTONIAN, BINGHAM, POWERLAW
from .config import SimConfig, read_configs, from_dict, BUILDERS
from .lattice import Lattice, build_map

__all__ = [
    "VlbmDesc", "VlbmError", "default_desc", "FP64", "FP32", "FP16", "BF16", "NEWTONIAN", "BINGHAM", "POWERLAW",
    "SimConfig", "read_configs", "from_dict", "BUILDERS", "Lattice", "build_map",
]
