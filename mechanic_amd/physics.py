"""Host-side mirror of `mechanic_gpu::GpuPhysics` over the C ABI.

Method names, argument meaning and error behaviour follow
crates/mechanic-gpu/src/device.rs:199-1663; the error types follow
device.rs:113-189. All device work runs in libmechanic_hip.so; this module only
marshals rows. There is no CPU fallback: without the built library every
constructor raises `LibraryMissing`.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .scenes import Scene


class GpuPhysicsError(Exception):
    """GpuPhysicsError (device.rs:130-173)."""

    def __init__(self, kind: str, required: int = 0, capacity: int = 0, provided: int = 0, message: str = ""):
        super().__init__(message or kind)
        self.kind, self.required, self.capacity, self.provided = kind, required, capacity, provided


class GpuImpulseError(Exception):
    """GpuImpulseError (device.rs:175-189)."""

    def __init__(self, kind: str, body_index: int = 0, body_count: int = 0, message: str = ""):
        super().__init__(message or kind)
        self.kind, self.body_index, self.body_count = kind, body_index, body_count


class GpuReadbackError(Exception):
    """GpuReadbackError (device.rs:113-128)."""

    def __init__(self, kind: str, slot: Optional[int] = None, message: str = ""):
        super().__init__(message or kind)
        self.kind, self.slot = kind, slot


class DeviceError(RuntimeError):
    """A HIP runtime failure inside the library."""


_PHYSICS_KINDS = {
    _lib.MH_ERR_BODY_CAPACITY: "BodyCapacity",
    _lib.MH_ERR_BEARING_CAPACITY: "BearingCapacity",
    _lib.MH_ERR_COLLIDER_CAPACITY: "ColliderCapacity",
    _lib.MH_ERR_DRIVE_STATE_COUNT: "DriveStateCount",
    _lib.MH_ERR_COORDINATE_STATE_COUNT: "CoordinateStateCount",
}


@dataclass
class GpuPhysicsConfig:
    """GpuPhysicsConfig (device.rs:22-41) plus the constants this backend exposes."""

    collisions_enabled: bool = True
    mechanism_self_collisions: bool = True
    solver_iterations: int = 8
    friction: float = 0.05  # CONTACT_FRICTION, collision.wgsl:66
    pair_capacity: int = _lib.MAX_CONTACT_PAIRS
    use_graph: bool = True
    warm_start: int = _lib.WARM_START_CACHE  # WARM_START_CACHE / _LOG / _LEARNED
    residual: bool = False  # measure the contact residual after each tick's solve
    # Learned warm start fallback: a tick whose RMS residual after the learned
    # passes exceeds the limit (N s) runs more passes (0 disables the check;
    # 0 passes: as many as the learned run's).
    learned_residual_limit: float = 0.0
    learned_fallback_passes: int = 0
    # Loop-closure Newton solve: Newton steps (0: 12), PCG iterations per step
    # (0: 8) and PCG tolerance on sqrt(r.z / r0.z) (0: every iteration).
    closure_newton_steps: int = 0
    closure_pcg_iterations: int = 0
    closure_pcg_tolerance: float = 0.0
    closure_components: bool = False  # one PCG per closure component (one workgroup each)
    # Project the coordinate speeds onto the loops' velocity constraint before
    # the tree reconstruction (needs closure_components).
    closure_velocity_projection: bool = False

    def to_c(self) -> _lib.Config:
        return _lib.Config(int(self.collisions_enabled), int(self.mechanism_self_collisions),
                           int(self.solver_iterations), float(self.friction), int(self.pair_capacity),
                           int(self.warm_start), int(self.use_graph), int(self.residual),
                           float(self.learned_residual_limit), int(self.learned_fallback_passes),
                           int(self.closure_newton_steps), int(self.closure_pcg_iterations),
                           float(self.closure_pcg_tolerance), int(self.closure_components),
                           int(self.closure_velocity_projection))


@dataclass
class GpuTickReadback:
    """GpuTickReadback (device.rs:74-92)."""

    gpu_tick_ms: Optional[float]
    kernel_timings: Optional[dict]
    error_flags: int
    pair_count: int
    contact_count: int
    active_contact_count: int
    anchor_residual_meters: float
    axis_residual_degrees: float


def _check(rc: int, what: str) -> None:
    if rc == _lib.MH_OK:
        return
    msg = _lib.load().mh_status_string(rc).decode()
    if rc == _lib.MH_ERR_INVALID_SNAPSHOT_SLOT:
        raise GpuReadbackError("InvalidSnapshotSlot", message=f"{what}: {msg}")
    if rc == _lib.MH_ERR_DEVICE:
        raise GpuReadbackError("DevicePoll", message=f"{what}: {msg}")
    raise DeviceError(f"{what}: {msg} ({rc})")


class GpuPhysics:
    """One uploaded creation stepped on the GPU (device.rs:199-1663)."""

    def __init__(self, handle: int, body_count: int, config: GpuPhysicsConfig):
        self._h = ctypes.c_void_p(handle)
        self._body_count = body_count
        self.config = config

    # --- construction (device.rs:381-811) ---
    @classmethod
    def new(cls, scene: Scene) -> "GpuPhysics":
        return cls.new_with_config(scene, GpuPhysicsConfig())

    @staticmethod
    def _creation(scene: Scene, keep: list) -> "_lib.Creation":
        supp = None
        if scene.suppressed_pairs:
            pairs = sorted((min(a, b), max(a, b)) for a, b in scene.suppressed_pairs)
            supp = (_lib.Pair * len(pairs))(*[_lib.Pair(a, b) for a, b in pairs])
            keep.append(supp)

        def ptr(x):
            if x is None:
                return None
            if isinstance(x, np.ndarray):
                arr = np.ascontiguousarray(x, dtype=np.uint32)
                keep.append(arr)
                return arr.ctypes.data
            return ctypes.cast(x, ctypes.c_void_p).value

        return _lib.Creation(
            scene.body_count, len(scene.colliders), scene.bearing_count, len(scene.suppressed_pairs),
            ptr(scene.poses), ptr(scene.inverse_masses), ptr(scene.masses), ptr(scene.inertias),
            ptr(scene.colliders), ptr(scene.bearings), ptr(supp), ptr(scene.body_components), ptr(scene.root_flags),
            ptr(scene.mechanism_bodies), ptr(scene.preorder), ptr(scene.drives), scene.coordinate_count, 0)

    @classmethod
    def new_batch(cls, scenes, config: Optional[GpuPhysicsConfig] = None) -> "GpuPhysics":
        """One handle stepping many independent scenes (mh_batch_create); rows
        are concatenated scene after scene, see `scene_range`."""
        lib = _lib.load()
        config = config or GpuPhysicsConfig()
        keep: list = []
        rows = (_lib.Creation * len(scenes))(*[cls._creation(sc, keep) for sc in scenes])
        cfg = config.to_c()
        handle = ctypes.c_void_p()
        err = _lib.ErrorDetail()
        rc = lib.mh_batch_create(rows, len(scenes), ctypes.byref(cfg), ctypes.byref(handle), ctypes.byref(err))
        if rc in _PHYSICS_KINDS:
            raise GpuPhysicsError(_PHYSICS_KINDS[rc], err.required, err.capacity, err.provided, err.message.decode())
        if rc != _lib.MH_OK:
            raise DeviceError(f"mh_batch_create: {err.message.decode()} ({rc})")
        gpu = cls(handle.value, sum(sc.body_count for sc in scenes), config)
        gpu._coordinate_count = 0
        return gpu

    def scene_count(self) -> int:
        return int(_lib.load().mh_batch_scene_count(self._h))

    def scene_range(self, scene: int):
        """(first body, body count, first collider, collider count) of one scene."""
        vals = [ctypes.c_uint32() for _ in range(4)]
        _check(_lib.load().mh_batch_scene_range(self._h, scene, *[ctypes.byref(v) for v in vals]), "scene_range")
        return tuple(v.value for v in vals)

    def set_tuning(self, **fields) -> None:
        t = _lib.Tuning()
        for k, v in fields.items():
            setattr(t, k, int(v))
        _check(_lib.load().mh_physics_set_tuning(self._h, ctypes.byref(t)), "set_tuning")

    def get_tuning(self) -> dict:
        t = _lib.Tuning()
        _check(_lib.load().mh_physics_get_tuning(self._h, ctypes.byref(t)), "get_tuning")
        return {n: getattr(t, n) for n, _ in _lib.Tuning._fields_}

    @classmethod
    def new_with_config(cls, scene: Scene, config: GpuPhysicsConfig) -> "GpuPhysics":
        lib = _lib.load()
        keep = []
        creation = cls._creation(scene, keep)
        cfg = config.to_c()
        handle = ctypes.c_void_p()
        err = _lib.ErrorDetail()
        rc = lib.mh_physics_create(ctypes.byref(creation), ctypes.byref(cfg), ctypes.byref(handle), ctypes.byref(err))
        if rc in _PHYSICS_KINDS:
            raise GpuPhysicsError(_PHYSICS_KINDS[rc], err.required, err.capacity, err.provided, err.message.decode())
        if rc != _lib.MH_OK:
            raise DeviceError(f"mh_physics_create: {err.message.decode()} ({rc})")
        gpu = cls(handle.value, scene.body_count, config)
        gpu._coordinate_count = scene.coordinate_count
        return gpu

    @staticmethod
    def check_capacity(body_count: int, collider_count: int, bearing_count: int = 0) -> None:
        """The capacity checks new_with_config runs before any device work (device.rs:405-422)."""
        lib = _lib.load()
        creation = _lib.Creation(body_count, collider_count, bearing_count, 0)
        handle = ctypes.c_void_p()
        err = _lib.ErrorDetail()
        if body_count <= _lib.MAX_BODIES and collider_count <= _lib.MAX_COLLIDERS and bearing_count <= _lib.MAX_BEARINGS:
            return
        rc = lib.mh_physics_create(ctypes.byref(creation), None, ctypes.byref(handle), ctypes.byref(err))
        raise GpuPhysicsError(_PHYSICS_KINDS.get(rc, str(rc)), err.required, err.capacity, err.provided,
                              err.message.decode())

    def close(self) -> None:
        if self._h and self._h.value:
            _lib.load().mh_physics_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- GpuPhysics methods ---
    def body_count(self) -> int:
        return self._body_count

    def has_gpu_timestamps(self) -> bool:
        return bool(_lib.load().mh_physics_has_gpu_timestamps(self._h))

    def apply_impulse(self, body_index: int, world_point, impulse) -> None:
        """apply_impulse (device.rs:821-856)."""
        lib = _lib.load()
        p = (ctypes.c_float * 3)(*[float(x) for x in world_point])
        j = (ctypes.c_float * 3)(*[float(x) for x in impulse])
        err = _lib.ErrorDetail()
        rc = lib.mh_physics_apply_impulse(self._h, body_index, p, j, ctypes.byref(err))
        if rc == _lib.MH_ERR_BODY_INDEX_OUT_OF_RANGE:
            raise GpuImpulseError("BodyIndexOutOfRange", err.body_index, err.body_count, err.message.decode())
        if rc == _lib.MH_ERR_NON_FINITE:
            raise GpuImpulseError("NonFinite", message=err.message.decode())
        _check(rc, "apply_impulse")

    def dispatch_tick(self, tick_index: int) -> int:
        """dispatch_tick (device.rs:859-961); returns the snapshot slot."""
        slot = ctypes.c_uint8()
        _check(_lib.load().mh_physics_dispatch_tick(self._h, tick_index, ctypes.byref(slot)), "dispatch_tick")
        return slot.value

    def dispatch_ticks(self, first_tick: int, count: int) -> None:
        _check(_lib.load().mh_physics_dispatch_ticks(self._h, first_tick, count), "dispatch_ticks")

    def profile_tick(self, tick_index: int) -> dict:
        t = _lib.PhaseTimings()
        _check(_lib.load().mh_physics_profile_tick(self._h, tick_index, ctypes.byref(t)), "profile_tick")
        return {name: getattr(t, name) for name, _ in _lib.PhaseTimings._fields_}

    def debug_tick(self, tick_index: int, stop: int) -> None:
        _check(_lib.load().mh_physics_debug_tick(self._h, tick_index, stop), "debug_tick")

    def synchronize(self) -> None:
        _check(_lib.load().mh_physics_synchronize(self._h), "synchronize")

    def read_last_tick(self) -> GpuTickReadback:
        """read_last_tick (device.rs:1439-1524)."""
        r = _lib.TickReadback()
        _check(_lib.load().mh_physics_read_last_tick(self._h, ctypes.byref(r)), "read_last_tick")
        timings = None
        if r.has_timings:
            timings = dict(integration_ms=r.integration_ms, mechanism_ms=r.mechanism_ms,
                           broadphase_ms=r.broadphase_ms, narrowphase_ms=r.narrowphase_ms,
                           contact_solver_ms=r.contact_solver_ms, bearings_ms=r.bearings_ms,
                           snapshot_ms=r.snapshot_ms)
        return GpuTickReadback(r.gpu_tick_ms if r.has_timings else None, timings, r.error_flags, r.pair_count,
                               r.contact_count, r.active_contact_count, r.anchor_residual_meters,
                               r.axis_residual_degrees)

    def read_snapshot_transforms(self, slot: int) -> np.ndarray:
        """read_snapshot_transforms (device.rs:1534-1570) as an (N, 8) float32 array."""
        if not 0 <= slot < _lib.SNAPSHOT_RING_SIZE:
            raise GpuReadbackError("InvalidSnapshotSlot", slot)
        out = (_lib.Transform * max(self._body_count, 1))()
        _check(_lib.load().mh_physics_read_snapshot(self._h, slot, out, self._body_count), "read_snapshot")
        return np.ctypeslib.as_array(out).view(np.float32).reshape(-1, 8)[: self._body_count].copy()

    def snapshot(self, slot: int):
        """snapshot (device.rs:1578-1580): device pointers of one ring slot, or None."""
        pos, rot = ctypes.c_void_p(), ctypes.c_void_p()
        rc = _lib.load().mh_physics_snapshot_device_ptrs(self._h, slot, ctypes.byref(pos), ctypes.byref(rot))
        if rc == _lib.MH_ERR_INVALID_SNAPSHOT_SLOT:
            return None
        _check(rc, "snapshot")
        return pos.value, rot.value

    def read_bodies(self):
        """(poses, velocities) as (N, 8) float32 arrays."""
        n = max(self._body_count, 1)
        poses = (_lib.Transform * n)()
        vels = (_lib.Velocity * n)()
        _check(_lib.load().mh_physics_read_bodies(self._h, poses, vels, n), "read_bodies")
        p = np.ctypeslib.as_array(poses).view(np.float32).reshape(-1, 8)[: self._body_count].copy()
        v = np.ctypeslib.as_array(vels).view(np.float32).reshape(-1, 8)[: self._body_count].copy()
        return p, v

    def write_bodies(self, poses: Optional[np.ndarray], velocities: Optional[np.ndarray]) -> None:
        n = self._body_count
        pp = vp = None
        if poses is not None:
            pa = np.ascontiguousarray(poses, dtype=np.float32).reshape(n, 8)
            pp = pa.ctypes.data_as(ctypes.POINTER(_lib.Transform))
        if velocities is not None:
            va = np.ascontiguousarray(velocities, dtype=np.float32).reshape(n, 8)
            vp = va.ctypes.data_as(ctypes.POINTER(_lib.Velocity))
        _check(_lib.load().mh_physics_write_bodies(self._h, pp, vp, n), "write_bodies")

    def read_buffer(self, buffer: int, nbytes: int) -> bytes:
        buf = (ctypes.c_uint8 * max(nbytes, 1))()
        written = ctypes.c_size_t()
        _check(_lib.load().mh_physics_read_buffer(self._h, buffer, buf, nbytes, ctypes.byref(written)),
               "read_buffer")
        return bytes(buf)[: written.value]

    def write_mechanism_drives(self, drives) -> None:
        """write_mechanism_drives (device.rs:1622-1644)."""
        n = len(drives)
        arr = (_lib.MechanismDrive * max(n, 1))(*drives)
        err = _lib.ErrorDetail()
        rc = _lib.load().mh_physics_write_drives(self._h, arr, n, ctypes.byref(err))
        if rc in _PHYSICS_KINDS:
            raise GpuPhysicsError(_PHYSICS_KINDS[rc], err.required, err.capacity, err.provided, err.message.decode())
        _check(rc, "write_mechanism_drives")

    def initialize_mechanism_coordinates(self, coordinates) -> None:
        """initialize_mechanism_coordinates (device.rs:1646-1662)."""
        n = len(coordinates)
        arr = (_lib.MechanismCoordinate * max(n, 1))(*coordinates)
        err = _lib.ErrorDetail()
        rc = _lib.load().mh_physics_init_coordinates(self._h, arr, n, ctypes.byref(err))
        if rc in _PHYSICS_KINDS:
            raise GpuPhysicsError(_PHYSICS_KINDS[rc], err.required, err.capacity, err.provided, err.message.decode())
        _check(rc, "initialize_mechanism_coordinates")


    # --- learned warm start and contact residual (no counterpart in the reference) ---
    def set_mlp(self, mlp) -> None:
        """Attaches a `WarmStartNet` (None detaches) for WARM_START_LEARNED handles."""
        _check(_lib.load().mh_physics_set_mlp(self._h, mlp._h if mlp is not None else None), "set_mlp")
        self._mlp = mlp  # keep the network alive while attached

    def read_training_rows(self):
        """(features (R, 16), targets (R, 3)) of the last tick of a WARM_START_LOG handle."""
        lib = _lib.load()
        rows = ctypes.c_size_t()
        cap = max(int(self.config.pair_capacity), 1)
        f = np.empty((cap, _lib.MLP_FEATURES), np.float32)
        t = np.empty((cap, _lib.MLP_OUTPUTS), np.float32)
        fp = f.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        tp = t.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        _check(lib.mh_physics_read_training_rows(self._h, fp, tp, cap, ctypes.byref(rows)), "read_training_rows")
        return f[: rows.value].copy(), t[: rows.value].copy()

    def read_residual(self) -> dict:
        """Contact residual of the last tick of a handle created with residual=True."""
        r = _lib.ContactResidual()
        _check(_lib.load().mh_physics_read_residual(self._h, ctypes.byref(r)), "read_residual")
        return dict(rms=r.rms, mean_penetration=r.mean_penetration, max_penetration=r.max_penetration,
                    active_contacts=r.active_contacts, fallback_passes=r.fallback_passes)

    def read_learned_stats(self) -> dict:
        """Fallback counts of a learned-warm-start handle."""
        r = _lib.LearnedStats()
        _check(_lib.load().mh_physics_read_learned_stats(self._h, ctypes.byref(r)), "read_learned_stats")
        return {name: getattr(r, name) for name, _ in _lib.LearnedStats._fields_}


    def read_coordinates(self) -> np.ndarray:
        """(Q, 2) float32 array of (angle, angular velocity) per coordinate."""
        q = int(getattr(self, "_coordinate_count", 0))
        arr = (_lib.MechanismCoordinate * max(q, 1))()
        _check(_lib.load().mh_physics_read_coordinates(self._h, arr, q), "read_coordinates")
        return np.ctypeslib.as_array(arr).view(np.float32).reshape(-1, 2)[:q].copy()


def device_count() -> int:
    n = ctypes.c_int()
    _lib.load().mh_device_count(ctypes.byref(n))
    return n.value
