"""fluidx3d_amd: MI355X-native lattice Boltzmann core with the FluidX3D LBM API.

The compute path is libfx_lbm.so (hand-written HIP for gfx950, C ABI in
include/fx_lbm.h).  `LBM` mirrors the reference's host class on top of that ABI.
"""
from ._abi import (FxError, relaxation_rate, device_count, device_info, make_config, bandwidth_bytes_per_cell,
                   memory_bytes_per_cell, Context)
from .lbm import LBM, TYPE_S, TYPE_E, TYPE_T, TYPE_F, TYPE_I, TYPE_G, TYPE_X, TYPE_Y

__all__ = ["FxError", "relaxation_rate", "device_count", "device_info", "make_config", "bandwidth_bytes_per_cell",
           "memory_bytes_per_cell", "Context", "LBM", "TYPE_S", "TYPE_E", "TYPE_T", "TYPE_F", "TYPE_I", "TYPE_G",
           "TYPE_X", "TYPE_Y"]
