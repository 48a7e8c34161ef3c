"""mwdiem_amd: MI355X (gfx950) implementation of MWDiEM's per-timestep DEM loop
for convex polyhedra -- cell-hash candidates, FP64 GJK/EPA fine contact, the
spring-dashpot/Coulomb contact law with a deterministic per-particle
reduction, and the quaternion Verlet integrator -- as hand-written HIP kernels
behind a C ABI (include/mwdiem_hip.h), with drop-in replacements for the
reference's gpu_* modules in mwdiem_amd.dropin.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
__version__ = "0.1.0"
