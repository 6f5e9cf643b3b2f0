"""MI355X-native (gfx950) WENO reconstruction path: HIP kernels behind a C-ABI
(include/weno_c.h), with a Fortran drop-in for the reference's module and a
Python mirror of the same call surface."""
from ._lib import (ACCUMULATE, BF16, DIR_X, DIR_Y, DIR_Z, F32, F64, FUSE_COPY, FUSE_MP, LIB_PATH, ONLY_L, ONLY_R,
                   SRC_FIELD, SRC_RESHAPED, STREAMING, Block, Coeffs, Flux, Grid, WenoError, check, lib, make_grid)
from .device import DeviceArray
from . import fields

__all__ = [
    "ACCUMULATE", "Flux", "BF16", "DIR_X", "DIR_Y", "DIR_Z", "F32", "F64", "FUSE_COPY", "FUSE_MP", "LIB_PATH", "ONLY_L",
    "ONLY_R", "SRC_FIELD", "SRC_RESHAPED", "STREAMING", "Block", "Coeffs", "Grid", "WenoError", "check", "lib",
    "make_grid", "DeviceArray", "fields",
]
