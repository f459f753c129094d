"""Seeded synthetic scenes in the row format GpuPhysics uploads.

The reference builds its benchmark scenes through the construction graph
(crates/mechanic-bench/src/main.rs:442-524); these helpers produce the same row
shapes directly: one compound per cuboid with one collider
(crates/mechanic-core/src/compile.rs:131-148, 333-334, 889-895).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Scene:
    """Rows of a CompiledCreation lowered for upload (device.rs:424-523)."""

    body_count: int
    poses: ctypes.Array
    inverse_masses: ctypes.Array
    masses: ctypes.Array
    inertias: ctypes.Array
    colliders: ctypes.Array
    suppressed_pairs: list = field(default_factory=list)
    body_components: object = None
    root_flags: object = None
    # Articulated rows (mechanic_amd.articulated.compile_mechanisms).
    bearings: object = None
    bearing_count: int = 0
    mechanism_bodies: object = None
    preorder: object = None
    drives: object = None
    coordinate_count: int = 0

    @property
    def collider_count(self) -> int:
        return len(self.colliders)

    def pose_array(self) -> np.ndarray:
        """(N, 8) float32 view: position xyzw, rotation xyzw."""
        return np.ctypeslib.as_array(self.poses).view(np.float32).reshape(self.body_count, 8)

    def collider_array(self) -> np.ndarray:
        """(C, 16) float32 view of the collider rows (metadata as raw bits)."""
        return np.ctypeslib.as_array(self.colliders).view(np.float32).reshape(len(self.colliders), 16)


# Constants of the pile configurations: 0.25 m cubes, density 1000 kg/m^3,
# 0.27 m lattice spacing, lowest faces 0.5 m above the ground, +-0.02 m jitter,
# uniform yaw.
PILE_CONFIGS = {
    "pile_1k": dict(nx=10, ny=10, nz=10, seed=0),
    "pile_100k": dict(nx=50, ny=40, nz=50, seed=1),
}


def cuboid_pile(nx: int, ny: int, nz: int, seed: int = 0, edge: float = 0.25, spacing: float = 0.27,
                drop_height: float = 0.5, jitter: float = 0.02, density: float = 1000.0, yaw: bool = True) -> Scene:
    """A jittered, yawed lattice of cuboids dropping onto the ground plane y = 0."""
    lib = _lib.load()
    n = nx * ny * nz
    params = _lib.PileParams(nx, ny, nz, edge, spacing, drop_height, jitter, density, seed, 1 if yaw else 0, 0)
    poses = (_lib.Transform * n)()
    inv = (ctypes.c_float * n)()
    masses = (_lib.Mass * n)()
    inertias = (_lib.SpatialInertia * n)()
    colliders = (_lib.Collider * n)()
    rc = lib.mh_scene_cuboid_pile(ctypes.byref(params), poses, inv, masses, inertias, colliders, n)
    if rc != _lib.MH_OK:
        raise ValueError(f"mh_scene_cuboid_pile failed: {rc}")
    return Scene(n, poses, inv, masses, inertias, colliders)


def named_pile(name: str, **overrides) -> Scene:
    cfg = dict(PILE_CONFIGS[name])
    cfg.update(overrides)
    return cuboid_pile(**cfg)


def boxes(centers, half_extents, rotations=None, inverse_mass=1.0, inverse_inertia=6.0) -> Scene:
    """Hand-placed single-collider bodies, for small deterministic cases.

    `inverse_mass` 0 makes a body static (compile.rs:812-819)."""
    centers = np.asarray(centers, dtype=np.float32).reshape(-1, 3)
    n = len(centers)
    half = np.broadcast_to(np.asarray(half_extents, dtype=np.float32), (n, 3))
    rot = np.zeros((n, 4), np.float32)
    rot[:, 3] = 1.0
    if rotations is not None:
        rot = np.asarray(rotations, dtype=np.float32).reshape(n, 4)
    inv_m = np.broadcast_to(np.asarray(inverse_mass, dtype=np.float32), (n,))
    inv_i = np.broadcast_to(np.asarray(inverse_inertia, dtype=np.float32), (n,))
    poses = (_lib.Transform * n)()
    inv = (ctypes.c_float * n)()
    masses = (_lib.Mass * n)()
    inertias = (_lib.SpatialInertia * n)()
    colliders = (_lib.Collider * n)()
    for i in range(n):
        poses[i].position[:] = [float(centers[i, 0]), float(centers[i, 1]), float(centers[i, 2]), 0.0]
        poses[i].rotation[:] = [0.0, 0.0, 0.0, 1.0]
        inv[i] = float(inv_m[i])
        masses[i].inverse_mass[0] = float(inv_m[i])
        masses[i].inverse_inertia_x[0] = float(inv_i[i])
        masses[i].inverse_inertia_y[1] = float(inv_i[i])
        masses[i].inverse_inertia_z[2] = float(inv_i[i])
        if inv_m[i] > 0:
            inertias[i].mass[0] = 1.0 / float(inv_m[i])
        colliders[i].local_rotation[:] = [float(x) for x in rot[i]]
        colliders[i].half_extents[:] = [float(half[i, 0]), float(half[i, 1]), float(half[i, 2]), 0.0]
        colliders[i].metadata[:] = [i, i, 0, 0]
    return Scene(n, poses, inv, masses, inertias, colliders)


def batch_scenes(count: int = 256, seed_base: int = 1000, smallest: int = 64, largest: int = 4096,
                 size_seed: int = 3) -> list:
    """The batched configuration: `count` independent jittered drops like the
    1,000-cuboid pile, scene s seeded with seed_base + s and holding
    round(exp(U[ln smallest, ln largest])) cuboids (the bottom layers of a
    near-cubic lattice)."""
    rng = np.random.default_rng(size_seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(smallest), np.log(largest), count))).astype(int)
    out = []
    for s, n in enumerate(sizes):
        side = int(np.ceil(np.cbrt(n)))
        ny = int(np.ceil(n / (side * side)))
        full = cuboid_pile(side, ny, side, seed=seed_base + s)
        out.append(truncate(full, int(n)))
    return out


def truncate(scene: Scene, n: int) -> Scene:
    """The first n bodies of a single-collider-per-body scene."""
    poses = (_lib.Transform * n)(*list(scene.poses)[:n])
    inv = (ctypes.c_float * n)(*list(scene.inverse_masses)[:n])
    masses = (_lib.Mass * n)(*list(scene.masses)[:n])
    inertias = (_lib.SpatialInertia * n)(*list(scene.inertias)[:n])
    cols = (_lib.Collider * n)(*list(scene.colliders)[:n])
    return Scene(n, poses, inv, masses, inertias, cols)
