"""MI355X-native build-and-intersect library with the RadeonRays C API.

The product is libradeonrays.so (HIP kernels for gfx950 behind the rr* C
ABI, include/radeonrays*.h).  This package holds its sources (csrc/), the
ctypes binding (rr.py), a torch-backed convenience wrapper (api.py) and the
seeded workload generators (scenes.py).
"""
from . import rr  # noqa: F401

__all__ = ["rr", "api", "scenes"]
