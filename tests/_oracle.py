"""ctypes wrapper of the CPU oracle (oracle/liboracle.so), the checker of the tests."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    lib = ctypes.CDLL(ORACLE_LIB)
    P, u32, u64, f32, i32, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_int, ctypes.c_size_t
    lib.oracle_create.restype = P
    lib.oracle_create.argtypes = [u32, u32, u32, P, P, P, P, P, P, P, u32, f32, u32, u32, u32, u32, i32]
    lib.oracle_destroy.argtypes = [P]
    lib.oracle_step.argtypes = [P, u64]
    lib.oracle_step_partial.argtypes = [P, u64, i32]
    lib.oracle_set_threads.argtypes = [P, i32]
    lib.oracle_apply_impulse.argtypes = [P, u32, P, P]
    lib.oracle_get_bodies.argtypes = [P, P, P]
    lib.oracle_get_bodies_f64.argtypes = [P, P, P]
    lib.oracle_set_bodies.argtypes = [P, P, P]
    lib.oracle_get_diag.argtypes = [P, P]
    lib.oracle_set_closure.restype = u64
    lib.oracle_set_closure.argtypes = [P, u32, u32, f32, i32, i32]
    lib.oracle_set_serial_limit.argtypes = [P, u32]
    lib.oracle_phase_ms.restype = i32
    lib.oracle_phase_ms.argtypes = [P, P, i32]
    for name in ("oracle_get_keys", "oracle_get_children", "oracle_get_parents", "oracle_get_pairs",
                 "oracle_get_contacts", "oracle_get_world_masses"):
        getattr(lib, name).restype = sz
        getattr(lib, name).argtypes = [P, P, sz]
    lib.oracle_get_aabbs.restype = sz
    lib.oracle_get_aabbs.argtypes = [P, P, sz, i32]
    lib.oracle_morton_code.restype = u32
    lib.oracle_morton_code.argtypes = [f32, f32, f32]
    lib.oracle_set_mechanisms.argtypes = [P, u32, P, P, P, u32, P]
    lib.oracle_get_coordinates.argtypes = [P, P]
    lib.oracle_set_coordinates.argtypes = [P, P]
    lib.oracle_set_drives.restype = i32
    lib.oracle_set_drives.argtypes = [P, u32, P]
    lib.oracle_fk_newton.argtypes = [P, i32]
    lib.oracle_closure_residuals.restype = sz
    lib.oracle_closure_residuals.argtypes = [P, P, sz]
    lib.oracle_obb_sat_f64.restype = i32
    lib.oracle_obb_sat_f64.argtypes = [P] * 7
    _lib = lib
    return lib


def _addr(x):
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    return ctypes.cast(x, ctypes.c_void_p).value


class Oracle:
    def __init__(self, scene, iterations=8, friction=0.05, pair_capacity=1 << 21, collisions=True,
                 self_collisions=True, fp64=False, threads=1):
        lib = load()
        self.n = scene.body_count
        self.c = len(scene.colliders)
        self.fp64 = fp64
        supp = np.array(sorted((min(a, b), max(a, b)) for a, b in scene.suppressed_pairs), dtype=np.uint32).reshape(-1, 2)
        self._keep = [supp]
        comps = None if scene.body_components is None else np.ascontiguousarray(scene.body_components, np.uint32)
        roots = None if scene.root_flags is None else np.ascontiguousarray(scene.root_flags, np.uint32)
        self._keep += [comps, roots]
        self.h = lib.oracle_create(self.n, self.c, len(supp), _addr(scene.poses), _addr(scene.inverse_masses),
                                   _addr(scene.masses), _addr(scene.colliders), _addr(supp) if len(supp) else None,
                                   _addr(comps), _addr(roots), iterations, friction, pair_capacity,
                                   int(collisions), int(self_collisions), int(fp64), threads)
        self.q = 0
        if getattr(scene, "bearing_count", 0):
            pre = np.ascontiguousarray(scene.preorder, np.uint32)
            self._keep.append(pre)
            self.q = scene.coordinate_count
            lib.oracle_set_mechanisms(self.h, scene.bearing_count, _addr(scene.bearings), _addr(scene.mechanism_bodies),
                                      _addr(pre), self.q, _addr(scene.drives) if self.q else None)

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_destroy(self.h)
            self.h = None

    def step(self, tick):
        load().oracle_step(self.h, tick)

    def step_partial(self, tick, stop):
        load().oracle_step_partial(self.h, tick, stop)

    def set_threads(self, t):
        load().oracle_set_threads(self.h, t)

    def apply_impulse(self, body, point, impulse):
        p = np.asarray(point, np.float32)
        j = np.asarray(impulse, np.float32)
        load().oracle_apply_impulse(self.h, body, p.ctypes.data, j.ctypes.data)

    def bodies(self):
        p = np.zeros((self.n, 8), np.float32)
        v = np.zeros((self.n, 8), np.float32)
        load().oracle_get_bodies(self.h, p.ctypes.data, v.ctypes.data)
        return p, v

    def bodies_f64(self):
        assert self.fp64
        p = np.zeros((self.n, 8), np.float64)
        v = np.zeros((self.n, 8), np.float64)
        load().oracle_get_bodies_f64(self.h, p.ctypes.data, v.ctypes.data)
        return p, v

    def set_bodies(self, poses, velocities):
        p = np.ascontiguousarray(poses, np.float32)
        v = np.ascontiguousarray(velocities, np.float32)
        load().oracle_set_bodies(self.h, p.ctypes.data, v.ctypes.data)

    def coordinates(self):
        out = np.zeros((max(self.q, 1), 2), np.float32)
        load().oracle_get_coordinates(self.h, out.ctypes.data)
        return out[: self.q]

    def set_coordinates(self, rows):
        a = np.ascontiguousarray(rows, np.float32).reshape(-1, 2)
        load().oracle_set_coordinates(self.h, a.ctypes.data)

    def set_drives(self, rows):
        """write_mechanism_drives (device.rs:1622-1644): one GpuMechanismDrive
        row per coordinate."""
        arr = np.zeros((max(len(rows), 1), 8), np.float32)
        for q, r in enumerate(rows):
            arr[q] = np.frombuffer(bytes(r), np.float32)
        if load().oracle_set_drives(self.h, len(rows), arr.ctypes.data) != 0:
            raise ValueError(f"DriveStateCount: {len(rows)} rows for {self.q} coordinates")

    def set_serial_limit(self, bearings):
        """Largest bearing count that takes the serial small-mechanism sweeps."""
        load().oracle_set_serial_limit(self.h, bearings)

    def set_closure(self, steps=0, iterations=0, tolerance=0.0, components=False, velocity_projection=False):
        """Closure Newton steps, PCG iterations per step and PCG tolerance
        (0 keeps 12 / 8 / none), one PCG per closure component instead of one
        global system, and the velocity-level closure projection before the
        tree reconstruction; returns the PCG iterations run so far."""
        return load().oracle_set_closure(self.h, steps, iterations, tolerance, int(components),
                                         int(velocity_projection))

    def fk_newton(self, steps=0):
        load().oracle_fk_newton(self.h, steps)

    def closure_residuals(self):
        lib = load()
        n = lib.oracle_closure_residuals(self.h, None, 0)
        out = np.zeros((max(n, 1), 6), np.float64)
        lib.oracle_closure_residuals(self.h, out.ctypes.data, n)
        return out[:n]

    PHASES = ("integrate", "mechanism", "world_masses", "morton", "sort", "lbvh", "traverse", "narrowphase",
              "prepare", "solve", "validate")

    def phase_ms(self, reset=False):
        """Wall time (ms) of each tick phase accumulated since the last reset."""
        out = np.zeros(len(self.PHASES), np.float64)
        load().oracle_phase_ms(self.h, out.ctypes.data, int(reset))
        return dict(zip(self.PHASES, out.tolist()))

    def diag(self):
        d = np.zeros(8, np.uint32)
        load().oracle_get_diag(self.h, d.ctypes.data)
        return d

    def _get(self, fn, width, dtype):
        lib = load()
        n = getattr(lib, fn)(self.h, None, 0)
        out = np.zeros((max(n, 1), width), dtype)
        getattr(lib, fn)(self.h, out.ctypes.data, n)
        return out[:n]

    def keys(self):
        return self._get("oracle_get_keys", 1, np.uint64).reshape(-1)

    def children(self):
        return self._get("oracle_get_children", 2, np.uint32)

    def parents(self):
        return self._get("oracle_get_parents", 1, np.uint32).reshape(-1)

    def pairs(self):
        return self._get("oracle_get_pairs", 2, np.uint32)

    def contacts(self):
        return self._get("oracle_get_contacts", 16, np.float32)

    def world_masses(self):
        return self._get("oracle_get_world_masses", 4, np.float32).reshape(-1, 12)

    def aabbs(self, nodes=False):
        lib = load()
        n = lib.oracle_get_aabbs(self.h, None, 0, int(nodes))
        out = np.zeros((max(n, 1), 8), np.float32)
        lib.oracle_get_aabbs(self.h, out.ctypes.data, n, int(nodes))
        return out[:n]


def morton_code(x, y, z):
    return load().oracle_morton_code(x, y, z)


def obb_sat_f64(ca, ra, ha, cb, rb, hb):
    arrs = [np.asarray(v, np.float64) for v in (ca, ra, ha, cb, rb, hb)]
    out = np.zeros(7, np.float64)
    near = load().oracle_obb_sat_f64(*[a.ctypes.data for a in arrs], out.ctypes.data)
    return out[:3], out[3], out[4:7], near
