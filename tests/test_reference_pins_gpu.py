"""The reference's own behaviour pins for the contact path, through the C ABI.

Each scene is built the way the reference test builds it, through the
construction-graph lowering, and run on the GPU; the reference's thresholds
are asserted on the GPU state, and the state is also compared with the oracle
bit for bit.
"""
import math

import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd import construction as C
from tests._oracle import Oracle

pytestmark = pytest.mark.gpu


def run_both(scene, ticks, iterations=8, collisions=True, impulse=None):
    cfg = mh.GpuPhysicsConfig(collisions_enabled=collisions, solver_iterations=iterations)
    gpu = mh.GpuPhysics.new_with_config(scene, cfg)
    ora = Oracle(scene, iterations=iterations, collisions=collisions)
    if impulse is not None:
        gpu.apply_impulse(*impulse)
        ora.apply_impulse(*impulse)
    for tick in range(1, ticks + 1):
        gpu.dispatch_tick(tick)
        ora.step(tick)
    gp, gv = gpu.read_bodies()
    op, ov = ora.bodies()
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), "poses differ from the oracle"
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32)), "velocities differ from the oracle"
    return gpu, ora


def test_off_centre_external_impulse_changes_linear_and_angular_motion():
    # device.rs:4144-4176: a 1 m cube at y = 2 m struck 0.5 m above its centre
    # along +x moves along +x and turns about -z within one tick.
    g = C.ConstructionGraph()
    g.spawn(C.Cuboid((4, 4, 4), C.BuildPose((0, 8, 0))))
    out = g.compile()
    initial = out.compounds[0].root_translation
    gpu, _ = run_both(out.scene, 1, impulse=(0, [initial[0], initial[1] + 0.5, initial[2]], [500.0, 0.0, 0.0]))
    snap = gpu.read_snapshot_transforms(1)
    assert snap[0, 0] > initial[0] + 0.01
    assert snap[0, 6] < -0.01
    gpu.close()


def test_offset_ground_contact_applies_angular_impulse_about_compound_centre():
    # device.rs:4178-4221: two 0.5 m cubes rigidly linked, the lower one on the
    # ground 2 m to one side of the compound's centre of mass: the support acts
    # through the offset collider and turns the compound about +z.
    g = C.ConstructionGraph()
    lower = g.spawn(C.Cuboid((2, 2, 2), C.BuildPose((4, 1, 0))))
    upper = g.spawn(C.Cuboid((2, 2, 2), C.BuildPose((-4, 9, 0))))
    g.rigid_link(lower, upper)
    out = g.compile()
    assert len(out.compounds) == 1 and len(out.colliders) == 2
    assert abs(out.colliders[0].local_center[0]) > 0.5
    gpu, _ = run_both(out.scene, 4)
    snap = gpu.read_snapshot_transforms(1)
    assert snap[0, 6] > 1.0e-4, snap[0]
    gpu.close()


def test_flat_cylinder_face_stays_supported_above_ground():
    # device.rs:4545-4567 (run_ticks, device.rs:4513-4538: 16 iterations): a
    # full cylinder of 1 m diameter and 0.5 m length standing on its face, as
    # sixteen segment colliders with ground roles, rests on the analytic rim
    # contacts without sinking through its 0.25 m half-length.
    g = C.ConstructionGraph()
    g.spawn(C.Cylinder(1.0, 0.0, 0.5, C.BuildPose((0, 4, 0))))
    out = g.compile()
    gpu, ora = run_both(out.scene, 120, iterations=16)
    r = gpu.read_last_tick()
    assert r.error_flags == 0
    snap = gpu.read_snapshot_transforms(120 % 3)
    assert snap[0, 1] >= 0.245, snap[0]
    x, y, z, w = snap[0, 4:8].astype(np.float64)
    up_y = 1.0 - 2.0 * (x * x + z * z)  # (rotation * Y) . Y
    assert up_y > 0.98
    gpu.close()


@pytest.mark.parametrize("tilt_deg", [90.0, 45.0])
def test_rolling_cylinder_rests_on_its_analytic_rim_point(tilt_deg):
    # The single analytic contact of a tilted or lying full cylinder
    # (collision.wgsl:193-238, 622-664) holds it up on the GPU as in the oracle.
    g = C.ConstructionGraph()
    g.spawn(C.Cylinder(1.0, 0.0, 1.0, C.BuildPose((0, 8, 0))))
    scene = g.compile().scene
    a = math.radians(tilt_deg)
    scene.poses[0].rotation[:] = [math.sin(a / 2), 0.0, 0.0, math.cos(a / 2)]
    gpu, _ = run_both(scene, 90, iterations=16)
    assert gpu.read_last_tick().error_flags == 0
    gpu.close()
