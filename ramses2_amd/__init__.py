"""MI355X (gfx950) implementation of RAMSES II's fine-level Godunov hydro step.

The compute path is the HIP library ramses2_amd/lib/libramses2_hip.so behind
the C-ABI of include/ramses2_hip.h; this package is its host side.
"""
from .hydro import HipHydro, HydroParams, UniformLevel, newdt  # noqa: F401
from .mesh import UniformMesh, prim_to_cons  # noqa: F401

__all__ = ["HipHydro", "HydroParams", "UniformLevel", "UniformMesh", "newdt", "prim_to_cons"]
