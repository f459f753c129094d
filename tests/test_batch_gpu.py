"""GPU tests of scene batches: every scene of a batch steps bit-identically to
the same scene in a handle of its own (and to the oracle for small scenes),
the launch tuning never changes results, and the 256-scene configuration runs."""
import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd.physics import GpuPhysics, GpuPhysicsConfig
from mechanic_amd.scenes import batch_scenes, cuboid_pile, truncate
from tests._oracle import Oracle

pytestmark = pytest.mark.gpu

CFG = GpuPhysicsConfig(friction=0.6, solver_iterations=8, pair_capacity=1 << 20)


def run_batch(scenes, ticks, tuning=None):
    with GpuPhysics.new_batch(scenes, CFG) as gpu:
        if tuning:
            gpu.set_tuning(**tuning)
        for t in range(1, ticks + 1):
            gpu.dispatch_tick(t)
        rb = gpu.read_last_tick()
        p, v = gpu.read_bodies()
        ranges = [gpu.scene_range(s) for s in range(len(scenes))]
    return p, v, ranges, rb


def run_solo(scene, ticks):
    cfg = GpuPhysicsConfig(friction=0.6, solver_iterations=8, pair_capacity=1 << 18)
    with GpuPhysics.new_with_config(scene, cfg) as gpu:
        for t in range(1, ticks + 1):
            gpu.dispatch_tick(t)
        rb = gpu.read_last_tick()
        p, v = gpu.read_bodies()
    return p, v, rb


def test_batch_scenes_match_solo_runs():
    scenes = batch_scenes(count=12, smallest=64, largest=1500, size_seed=11)
    ticks = 60
    p, v, ranges, rb = run_batch(scenes, ticks)
    assert rb.error_flags == 0
    pairs = contacts = 0
    for s, scene in enumerate(scenes):
        first, count, cfirst, ccount = ranges[s]
        assert count == scene.body_count and ccount == len(scene.colliders)
        sp, sv, srb = run_solo(scene, ticks)
        assert np.array_equal(p[first:first + count].view(np.uint32), sp.view(np.uint32)), s
        assert np.array_equal(v[first:first + count].view(np.uint32), sv.view(np.uint32)), s
        pairs += srb.pair_count
        contacts += srb.contact_count
    # The batch's broadphase found exactly the scenes' own pairs, no cross-scene ones.
    assert rb.pair_count == pairs and rb.contact_count == contacts


def test_small_batch_matches_oracle():
    scenes = [truncate(cuboid_pile(4, 4, 4, seed=1000 + s), n) for s, n in enumerate((64, 40, 17))]
    ticks = 40
    p, v, ranges, rb = run_batch(scenes, ticks)
    for s, scene in enumerate(scenes):
        ora = Oracle(scene, iterations=8, friction=0.6, threads=4)
        for t in range(1, ticks + 1):
            ora.step(t)
        op, ov = ora.bodies()
        first, count, _, _ = ranges[s]
        assert np.array_equal(p[first:first + count].view(np.uint32), op.view(np.uint32)), s
        assert np.array_equal(v[first:first + count].view(np.uint32), ov.view(np.uint32)), s


def test_tuning_never_changes_results():
    scenes = batch_scenes(count=8, smallest=64, largest=800, size_seed=5)
    base = run_batch(scenes, 30)
    tuned = run_batch(scenes, 30, dict(sort_tile=512, sort_levels=2, narrow_blocks=512, prepare_blocks=256,
                                       solve_blocks=256, scan_blocks=128))
    assert np.array_equal(base[0].view(np.uint32), tuned[0].view(np.uint32))
    assert np.array_equal(base[1].view(np.uint32), tuned[1].view(np.uint32))


@pytest.mark.parametrize("tuning", [dict(sort_kind=1), dict(traverse_kind=1), dict(traverse_kind=2),
                                    dict(sort_kind=1, traverse_kind=1), dict(solve_kind=1)])
def test_broadphase_variants_match_default(tuning):
    """The key sorts, pair walks and solver pass organisations are drop-ins for
    each other: same pairs, bit-identical states (scene-keyed batch keys are the
    widest keys, and every walk must stay inside each scene)."""
    scenes = batch_scenes(count=8, smallest=64, largest=1200, size_seed=9)
    base = run_batch(scenes, 30)
    other = run_batch(scenes, 30, tuning)
    assert base[3].pair_count == other[3].pair_count and base[3].contact_count == other[3].contact_count
    assert other[3].error_flags == 0
    assert np.array_equal(base[0].view(np.uint32), other[0].view(np.uint32))
    assert np.array_equal(base[1].view(np.uint32), other[1].view(np.uint32))


def test_tuning_round_trip_and_validation():
    scene = cuboid_pile(3, 3, 3, seed=0)
    with GpuPhysics.new_with_config(scene, CFG) as gpu:
        t = gpu.get_tuning()
        assert t["sort_tile"] == 2048 and t["sort_kind"] == 0 and t["traverse_kind"] == 0 and t["scan_blocks"] == 4096
        assert t["solve_kind"] == 0
        gpu.set_tuning(sort_tile=1024, scan_blocks=64)
        t = gpu.get_tuning()
        assert t["sort_tile"] == 1024 and t["scan_blocks"] == 64 and t["solve_blocks"] == 4096
        with pytest.raises(mh.DeviceError):
            gpu.set_tuning(sort_tile=3000)


def test_config3_batch_at_scale():
    """256 scenes of 64..4,096 cuboids (seeds 1000+s) in one handle."""
    scenes = batch_scenes()
    assert len(scenes) == 256 and sum(s.body_count for s in scenes) > 200_000
    cfg = GpuPhysicsConfig(friction=0.6, solver_iterations=8)
    with GpuPhysics.new_batch(scenes, cfg) as gpu:
        for t in range(1, 31):
            gpu.dispatch_tick(t)
        rb = gpu.read_last_tick()
        p, v = gpu.read_bodies()
        ranges = [gpu.scene_range(s) for s in (0, 97, 255)]
    assert rb.error_flags == 0 and rb.contact_count > 0
    for s, (first, count, _, _) in zip((0, 97, 255), ranges):
        sp, sv, _ = run_solo(scenes[s], 30)
        assert np.array_equal(p[first:first + count].view(np.uint32), sp.view(np.uint32)), s
