"""ctypes view of the C ABI in include/mechanic_hip.h.

The row structures mirror the reference's Pod rows
(crates/mechanic-gpu/src/abi.rs:16-296); their sizes are checked against
abi.rs:298-315 at import time.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char, c_double, c_float, c_int, c_size_t, c_uint8, c_uint32, c_uint64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
# The built library; MECHANIC_HIP_LIB names another build of it (an experiment,
# a different install). There is no fallback: a missing library is an error.
LIB_PATH = os.environ.get("MECHANIC_HIP_LIB") or os.path.join(HERE, "libmechanic_hip.so")

F4 = c_float * 4
U4 = c_uint32 * 4

MH_OK = 0
MH_ERR_BODY_CAPACITY = 1
MH_ERR_BEARING_CAPACITY = 2
MH_ERR_COLLIDER_CAPACITY = 3
MH_ERR_DRIVE_STATE_COUNT = 4
MH_ERR_COORDINATE_STATE_COUNT = 5
MH_ERR_BODY_INDEX_OUT_OF_RANGE = 6
MH_ERR_NON_FINITE = 7
MH_ERR_INVALID_SNAPSHOT_SLOT = 8
MH_ERR_DEVICE = 9
MH_ERR_INVALID_ARGUMENT = 10
MH_ERR_NO_DEVICE = 11
MH_ERR_UNSUPPORTED = 12

PAIR_OVERFLOW_FLAG = 1 << 0
INVALID_NUMERIC_FLAG = 1 << 1
CONSTRAINT_NON_CONVERGENCE_FLAG = 1 << 2
MANIFOLD_OVERFLOW_FLAG = 1 << 3

MAX_BODIES = 131072
MAX_BEARINGS = 262144
MAX_COLLIDERS = 131072
MAX_CONTACT_PAIRS = 2097152
MAX_BATCH_BODIES = 1048576
MAX_BATCH_SCENES = 4096
SNAPSHOT_RING_SIZE = 3
INVALID_INDEX = 0xFFFFFFFF

BUFFER_MORTON_KEYS = 1
BUFFER_COLLIDER_AABBS = 2
BUFFER_NODE_CHILDREN = 3
BUFFER_NODE_AABBS = 4
BUFFER_PAIRS = 5
BUFFER_CONTACTS = 6
BUFFER_DIAGNOSTICS = 7
BUFFER_WORLD_MASSES = 8
BUFFER_MANIFOLDS = 9
BUFFER_NODE_PARENTS = 10
BUFFER_BODY_SLOTS = 11

WARM_START_CACHE = 0
WARM_START_LOG = 1
WARM_START_LEARNED = 2
MLP_FEATURES = 16
MLP_OUTPUTS = 3


class TickConfig(ctypes.Structure):
    _fields_ = [(n, c_uint32) for n in ("body_count", "tick_index", "snapshot_slot", "collider_count")] + [
        (n, c_float) for n in ("delta_seconds", "gravity_y", "linear_damping", "angular_damping")
    ] + [(n, c_uint32) for n in ("bearing_count", "suppression_count", "pair_capacity", "flags", "hash_capacity",
                                 "solver_iterations", "reserved_a", "reserved_b")]


class Diagnostics(ctypes.Structure):
    _fields_ = [(n, c_uint32) for n in ("error_flags", "pair_count", "contact_count", "max_anchor_micrometers",
                                        "max_axis_microdegrees", "active_contact_count", "reserved_b", "reserved_c")]


class Transform(ctypes.Structure):
    _fields_ = [("position", F4), ("rotation", F4)]


class Velocity(ctypes.Structure):
    _fields_ = [("linear", F4), ("angular", F4)]


class Mass(ctypes.Structure):
    _fields_ = [("inverse_mass", F4), ("inverse_inertia_x", F4), ("inverse_inertia_y", F4),
                ("inverse_inertia_z", F4)]


class SpatialInertia(ctypes.Structure):
    _fields_ = [("mass", F4), ("inertia_x", F4), ("inertia_y", F4), ("inertia_z", F4)]


class Collider(ctypes.Structure):
    _fields_ = [("local_center", F4), ("local_rotation", F4), ("half_extents", F4), ("metadata", U4)]


class Bearing(ctypes.Structure):
    _fields_ = [("local_anchor_a", F4), ("local_anchor_b", F4), ("local_axis_a", F4), ("local_axis_b", F4),
                ("metadata", U4)]


class Pair(ctypes.Structure):
    _fields_ = [("collider_a", c_uint32), ("collider_b", c_uint32)]


class Contact(ctypes.Structure):
    _fields_ = [("metadata", U4), ("normal_penetration", F4), ("arm_a_impulse", F4), ("arm_b", F4)]


class PersistentManifold(ctypes.Structure):
    _fields_ = [("pair_tick", U4), ("normal_penetration", F4), ("point_impulse", F4)]


class MechanismBody(ctypes.Structure):
    _fields_ = [("metadata", U4), ("traversal", U4), ("bind_relative_position", F4),
                ("bind_relative_rotation", F4)]


class MechanismCoordinate(ctypes.Structure):
    _fields_ = [("angle", c_float), ("angular_velocity", c_float)]


class MechanismDrive(ctypes.Structure):
    _fields_ = [("mode", c_uint32)] + [(n, c_float) for n in (
        "max_acceleration", "max_speed", "target_speed", "target_angle", "min_angle", "max_angle", "padding")]


class LinkState(ctypes.Structure):
    _fields_ = [("position", F4), ("rotation", F4), ("metadata", U4)]


class ContractionNode(ctypes.Structure):
    _fields_ = [("metadata", U4)]


class Creation(ctypes.Structure):
    _fields_ = [
        ("body_count", c_uint32), ("collider_count", c_uint32), ("bearing_count", c_uint32),
        ("suppression_count", c_uint32),
        ("poses", c_void_p), ("inverse_masses", c_void_p), ("masses", c_void_p), ("inertias", c_void_p),
        ("colliders", c_void_p), ("bearings", c_void_p), ("suppressed_pairs", c_void_p),
        ("body_components", c_void_p), ("root_flags", c_void_p),
        ("mechanism_bodies", c_void_p), ("preorder", c_void_p), ("drives", c_void_p),
        ("coordinate_count", c_uint32), ("reserved", c_uint32),
    ]


class Config(ctypes.Structure):
    _fields_ = [("collisions_enabled", c_uint32), ("mechanism_self_collisions", c_uint32),
                ("solver_iterations", c_uint32), ("friction", c_float), ("pair_capacity", c_uint32),
                ("warm_start_mode", c_uint32), ("use_graph", c_uint32), ("residual", c_uint32),
                ("learned_residual_limit", c_float), ("learned_fallback_passes", c_uint32),
                ("closure_newton_steps", c_uint32), ("closure_pcg_iterations", c_uint32),
                ("closure_pcg_tolerance", c_float), ("closure_components", c_uint32),
                ("closure_velocity_projection", c_uint32)]


class Tuning(ctypes.Structure):
    _fields_ = [(n, c_uint32) for n in ("sort_tile", "sort_levels", "narrow_blocks", "prepare_blocks",
                                        "solve_blocks", "scan_blocks", "sort_kind", "traverse_kind",
                                        "solve_kind")]


class ContactResidual(ctypes.Structure):
    _fields_ = [("rms", c_double), ("mean_penetration", c_double), ("max_penetration", c_double),
                ("active_contacts", c_uint32), ("fallback_passes", c_uint32)]


class LearnedStats(ctypes.Structure):
    _fields_ = [("fallback_this_tick", c_uint32), ("fallback_ticks", c_uint32), ("nonfinite_this_tick", c_uint32),
                ("nonfinite_total", c_uint32)]


class TickReadback(ctypes.Structure):
    _fields_ = [(n, c_double) for n in ("gpu_tick_ms", "integration_ms", "mechanism_ms", "broadphase_ms",
                                        "narrowphase_ms", "contact_solver_ms", "bearings_ms", "snapshot_ms")] + [
        (n, c_uint32) for n in ("error_flags", "pair_count", "contact_count", "active_contact_count")] + [
        ("anchor_residual_meters", c_float), ("axis_residual_degrees", c_float), ("tick_index", c_uint64),
        ("snapshot_slot", c_uint32), ("has_timings", c_uint32)]


class PhaseTimings(ctypes.Structure):
    _fields_ = [(n, c_double) for n in ("integrate_ms", "morton_ms", "sort_ms", "lbvh_ms", "traverse_ms",
                                        "narrowphase_ms", "prepare_ms", "solve_ms", "persist_ms", "snapshot_ms",
                                        "total_ms", "mechanism_ms", "validation_ms", "solve_kernel_ms",
                                        "apply_kernel_ms", "solve_kernel_event_ms")]


class ErrorDetail(ctypes.Structure):
    _fields_ = [("code", c_int), ("required", c_uint64), ("capacity", c_uint64), ("provided", c_uint64),
                ("body_index", c_uint32), ("body_count", c_uint32), ("message", c_char * 160)]


class PileParams(ctypes.Structure):
    _fields_ = [("nx", c_uint32), ("ny", c_uint32), ("nz", c_uint32), ("edge", c_float), ("spacing", c_float),
                ("drop_height", c_float), ("jitter", c_float), ("density", c_float), ("seed", c_uint64),
                ("yaw", c_uint32), ("reserved", c_uint32)]


ROW_SIZES = {  # abi.rs:298-315
    TickConfig: 64, Diagnostics: 32, Transform: 32, Velocity: 32, Mass: 64, SpatialInertia: 64, Collider: 64,
    Bearing: 80, Pair: 8, Contact: 64, PersistentManifold: 48, MechanismBody: 64, MechanismCoordinate: 8,
    MechanismDrive: 32, LinkState: 48, ContractionNode: 16,
}
for _t, _s in ROW_SIZES.items():
    assert ctypes.sizeof(_t) == _s, (_t.__name__, ctypes.sizeof(_t), _s)

_lib = None


class LibraryMissing(RuntimeError):
    """The HIP library has not been built (run __graft_entry__.build())."""


def load() -> ctypes.CDLL:
    """Loads libmechanic_hip.so; raises LibraryMissing if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is missing: build the HIP library first")
    lib = ctypes.CDLL(LIB_PATH)
    H = c_void_p
    sig = {
        "mh_config_default": (Config, []),
        "mh_status_string": (ctypes.c_char_p, [c_int]),
        "mh_device_count": (c_int, [POINTER(c_int)]),
        "mh_device_arch": (c_int, [ctypes.c_char_p, c_size_t]),
        "mh_device_round_i32": (c_int, [POINTER(c_float), POINTER(ctypes.c_int32), c_size_t]),
        "mh_physics_create": (c_int, [POINTER(Creation), POINTER(Config), POINTER(H), POINTER(ErrorDetail)]),
        "mh_physics_destroy": (None, [H]),
        "mh_physics_body_count": (c_uint32, [H]),
        "mh_physics_has_gpu_timestamps": (c_int, [H]),
        "mh_physics_apply_impulse": (c_int, [H, c_uint32, POINTER(c_float), POINTER(c_float), POINTER(ErrorDetail)]),
        "mh_physics_dispatch_tick": (c_int, [H, c_uint64, POINTER(c_uint8)]),
        "mh_physics_dispatch_ticks": (c_int, [H, c_uint64, c_uint32]),
        "mh_physics_profile_tick": (c_int, [H, c_uint64, POINTER(PhaseTimings)]),
        "mh_physics_read_last_tick": (c_int, [H, POINTER(TickReadback)]),
        "mh_physics_read_snapshot": (c_int, [H, c_uint8, POINTER(Transform), c_size_t]),
        "mh_physics_snapshot_device_ptrs": (c_int, [H, c_uint8, POINTER(c_void_p), POINTER(c_void_p)]),
        "mh_physics_read_bodies": (c_int, [H, POINTER(Transform), POINTER(Velocity), c_size_t]),
        "mh_physics_write_bodies": (c_int, [H, POINTER(Transform), POINTER(Velocity), c_size_t]),
        "mh_physics_read_buffer": (c_int, [H, c_int, c_void_p, c_size_t, POINTER(c_size_t)]),
        "mh_physics_synchronize": (c_int, [H]),
        "mh_physics_debug_tick": (c_int, [H, c_uint64, c_int]),
        "mh_physics_write_drives": (c_int, [H, POINTER(MechanismDrive), c_size_t, POINTER(ErrorDetail)]),
        "mh_physics_init_coordinates": (c_int, [H, POINTER(MechanismCoordinate), c_size_t, POINTER(ErrorDetail)]),
        "mh_physics_read_coordinates": (c_int, [H, POINTER(MechanismCoordinate), c_size_t]),
        "mh_batch_create": (c_int, [POINTER(Creation), c_uint32, POINTER(Config), POINTER(H), POINTER(ErrorDetail)]),
        "mh_batch_scene_count": (c_uint32, [H]),
        "mh_batch_scene_range": (c_int, [H, c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32),
                                         POINTER(c_uint32)]),
        "mh_physics_set_tuning": (c_int, [H, POINTER(Tuning)]),
        "mh_physics_get_tuning": (c_int, [H, POINTER(Tuning)]),
        "mh_mlp_create": (c_int, [c_uint64, POINTER(H)]),
        "mh_mlp_destroy": (None, [H]),
        "mh_mlp_param_count": (c_uint32, []),
        "mh_mlp_net_floats": (c_uint32, []),
        "mh_mlp_get_net": (c_int, [H, POINTER(c_float), c_size_t]),
        "mh_mlp_set_net": (c_int, [H, POINTER(c_float), c_size_t]),
        "mh_mlp_set_dataset": (c_int, [H, POINTER(c_float), POINTER(c_float), c_size_t]),
        "mh_mlp_train": (c_int, [H, c_uint32, c_uint32, c_float, POINTER(c_float)]),
        "mh_mlp_gradient": (c_int, [H, POINTER(c_float), POINTER(c_float)]),
        "mh_mlp_predict": (c_int, [H, POINTER(c_float), c_size_t, POINTER(c_float)]),
        "mh_mlp_gradient_noalloc": (c_int, [H, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p]),
        "mh_physics_set_mlp": (c_int, [H, H]),
        "mh_physics_read_training_rows": (c_int, [H, POINTER(c_float), POINTER(c_float), c_size_t,
                                                  POINTER(c_size_t)]),
        "mh_physics_read_residual": (c_int, [H, POINTER(ContactResidual)]),
        "mh_physics_read_learned_stats": (c_int, [H, POINTER(LearnedStats)]),
        "mh_scene_cuboid_pile": (c_int, [POINTER(PileParams), POINTER(Transform), POINTER(c_float), POINTER(Mass),
                                         POINTER(SpatialInertia), POINTER(Collider), c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
