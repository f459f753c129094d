"""MI355X (gfx950) backend for Mechanic's per-tick rigid-body step.

`GpuPhysics` mirrors `mechanic_gpu::GpuPhysics`; the kernels and the host
runtime live in libmechanic_hip.so (built by `mechanic_amd.build`).
"""
from . import _lib
from ._lib import (CONSTRAINT_NON_CONVERGENCE_FLAG, INVALID_NUMERIC_FLAG, MANIFOLD_OVERFLOW_FLAG,
                   MAX_BEARINGS, MAX_BODIES, MAX_COLLIDERS, MAX_CONTACT_PAIRS, PAIR_OVERFLOW_FLAG,
                   SNAPSHOT_RING_SIZE, LibraryMissing)
from .physics import (DeviceError, GpuImpulseError, GpuPhysics, GpuPhysicsConfig, GpuPhysicsError,
                      GpuReadbackError, GpuTickReadback, device_count)
from .scenes import Scene, boxes, cuboid_pile, named_pile

__all__ = [
    "GpuPhysics", "GpuPhysicsConfig", "GpuPhysicsError", "GpuImpulseError", "GpuReadbackError", "GpuTickReadback",
    "DeviceError", "LibraryMissing", "Scene", "cuboid_pile", "named_pile", "boxes", "device_count",
    "PAIR_OVERFLOW_FLAG", "INVALID_NUMERIC_FLAG", "CONSTRAINT_NON_CONVERGENCE_FLAG", "MANIFOLD_OVERFLOW_FLAG",
    "MAX_BODIES", "MAX_BEARINGS", "MAX_COLLIDERS", "MAX_CONTACT_PAIRS", "SNAPSHOT_RING_SIZE",
]
