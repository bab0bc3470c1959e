"""qz_amd — MI355X-native block-sorting engine for qz's BSC stream path.

The compute path is the HIP library qz_amd/libqzbsc.so (C-ABI in
include/qzbsc.h); this package loads it through ctypes and offers the
Python forms of its entry points: `bsc` mirrors qz-lib's bsc.rs wrapper,
`bwt` exposes the BWT stages and their batch/device-resident forms.
"""
from . import bsc, bwt  # noqa: F401
from ._native import lib, LIB_PATH, NativeMissing  # noqa: F401

__all__ = ["bsc", "bwt", "lib", "LIB_PATH", "NativeMissing"]
