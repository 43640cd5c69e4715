"""MI355X-native MPPI engine: HIP kernels for gfx950 behind the C-ABI of include/mppi_amd.h.

Python is a thin host layer over that ABI (ctypes); the C++ plugin API lives in include/mppi_amd/.
"""
from . import capi  # noqa: F401
from .capi import MPPIError, load  # noqa: F401

__version__ = "0.1.0"
