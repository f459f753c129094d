"""Host-side checks of scene batches that need no GPU: the batched scene
generator and mh_batch_create's argument validation (which runs before any
device call)."""
import ctypes

import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd import _lib
from mechanic_amd.articulated import hinge_chain
from mechanic_amd.physics import GpuPhysics, GpuPhysicsError
from mechanic_amd.scenes import batch_scenes, cuboid_pile


def test_batch_scene_sizes_are_log_uniform():
    scenes = batch_scenes(count=256)
    sizes = np.array([s.body_count for s in scenes])
    assert sizes.min() >= 64 and sizes.max() <= 4096
    logs = np.log(sizes)
    # Log-uniform on [ln 64, ln 4096]: mean and spread of the logs.
    assert abs(logs.mean() - (np.log(64) + np.log(4096)) / 2) < 0.25
    assert abs(logs.std() - (np.log(4096) - np.log(64)) / np.sqrt(12)) < 0.2
    # Each scene is a drop of its own seed: distinct first poses.
    firsts = {tuple(s.pose_array()[0]) for s in scenes}
    assert len(firsts) == len(scenes)


def _create(scenes):
    keep = []
    rows = (_lib.Creation * max(len(scenes), 1))(*[GpuPhysics._creation(s, keep) for s in scenes])
    h = ctypes.c_void_p()
    err = _lib.ErrorDetail()
    rc = _lib.load().mh_batch_create(rows, len(scenes), None, ctypes.byref(h), ctypes.byref(err))
    return rc, err


def test_batch_create_rejects_bad_input_before_device_work():
    rc, _ = _create([])
    assert rc == _lib.MH_ERR_INVALID_ARGUMENT
    rc, _ = _create([cuboid_pile(2, 2, 2), hinge_chain(2)])
    assert rc == _lib.MH_ERR_UNSUPPORTED
