"""Host-side checks that need no GPU: the C ABI library loads and exports every
declared entry point, row layouts match the reference ABI, capacity errors come
back before any device work, and the seeded scene generator is deterministic."""
import ctypes
import os
import re

import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"MH_API\s+[\w\s\*]+?\b(mh_\w+)\s*\(", text)
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_row_sizes_match_reference_abi():
    # abi.rs:298-315
    for row, size in _lib.ROW_SIZES.items():
        assert ctypes.sizeof(row) == size, row.__name__
    assert ctypes.sizeof(_lib.Collider) % 16 == 0


def test_default_config_mirrors_reference():
    cfg = _lib.load().mh_config_default()
    assert cfg.collisions_enabled == 1 and cfg.mechanism_self_collisions == 1 and cfg.solver_iterations == 8
    assert abs(cfg.friction - 0.05) < 1e-9 and cfg.pair_capacity == 2097152


@pytest.mark.parametrize("bodies,colliders,bearings,kind", [
    (131073, 10, 0, "BodyCapacity"),
    (10, 131073, 0, "ColliderCapacity"),
    (10, 10, 262145, "BearingCapacity"),
])
def test_capacity_errors_before_device_work(bodies, colliders, bearings, kind):
    with pytest.raises(mh.GpuPhysicsError) as err:
        mh.GpuPhysics.check_capacity(bodies, colliders, bearings)
    assert err.value.kind == kind
    assert err.value.required == max(bodies, colliders, bearings) - 0 or err.value.required > 0


def test_status_strings():
    lib = _lib.load()
    for code in range(0, 13):
        assert lib.mh_status_string(code)


def test_pile_generator_is_seeded_and_shaped():
    a = mh.cuboid_pile(5, 4, 3, seed=42)
    b = mh.cuboid_pile(5, 4, 3, seed=42)
    c = mh.cuboid_pile(5, 4, 3, seed=43)
    pa, pb, pc = a.pose_array(), b.pose_array(), c.pose_array()
    assert a.body_count == 60
    assert np.array_equal(pa, pb) and not np.array_equal(pa, pc)
    # Layer y: lowest faces 0.5 m above ground at the first layer, +-0.02 m jitter.
    ys = pa[:, 1].reshape(4, 3, 5)
    for j in range(4):
        assert np.all(np.abs(ys[j] - (0.5 + 0.125 + 0.27 * j)) <= 0.02 + 1e-6)
    col = a.collider_array()
    q = col[:, 4:8]
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    assert np.allclose(q[:, [0, 2]], 0.0)  # yaw only
    assert np.allclose(col[:, 8:11], 0.125)
    meta = col[:, 12:16].view(np.uint32)
    assert np.array_equal(meta[:, 0], np.arange(60))
    inv = np.ctypeslib.as_array(a.inverse_masses)
    assert np.allclose(inv, 1.0 / 15.625)


def test_named_configs():
    s = mh.named_pile("pile_1k")
    assert s.body_count == 1000 and len(s.colliders) == 1000
