"""icarat2_amd -- MI355X (gfx950) implementation of ICARAT's linear-elastic SIMP
compliance loop on a voxel hexa8 mesh.

The product is the shared library libicarat_hip.so (C ABI in
include/icarat_hip.h, C++ facade in include/icarat_hip.hpp); this package
holds its sources (csrc/), its build recipe (build.py) and a ctypes binding
(_capi.py)."""
from . import _capi  # noqa: F401
from ._capi import Context, cantilever, device_available, gaussian_field  # noqa: F401

__all__ = ["Context", "cantilever", "device_available", "gaussian_field"]
