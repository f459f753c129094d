"""The serial small-mechanism solvers on the GPU (device.rs:19-20, 1370-1416;
collision.wgsl:979-1048; articulated.wgsl:216-285): the reference's car pins
(device.rs:4224-4473) with their thresholds, and the GPU state against the
oracle -- bit for bit while every joint angle is still zero (forward
kinematics then evaluates sin and cos only at zero, where the device and host
libraries agree), within the articulated tolerance after."""
import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd import construction as C
from mechanic_amd.physics import GpuPhysicsConfig
from tests._car import (GpuSim, assert_drop, assert_struck, assert_wall, car_fixture, drop, struck_wheel,
                        wall_impact)
from tests._oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def car():
    return car_fixture()


def gpu_for(scene):
    cfg = GpuPhysicsConfig(solver_iterations=8, mechanism_self_collisions=False)
    return mh.GpuPhysics.new_with_config(scene, cfg)


def oracle_for(scene):
    return Oracle(scene, iterations=8, self_collisions=False)


def same_bits(gpu, ora):
    gp, gv = gpu.read_bodies()
    op, ov = ora.bodies()
    return np.array_equal(gp.view(np.uint32), op.view(np.uint32)) and np.array_equal(gv.view(np.uint32),
                                                                                      ov.view(np.uint32))


def test_car_matches_the_oracle_through_its_first_contacts(car):
    gpu, ora = gpu_for(car.compiled.scene), oracle_for(car.compiled.scene)
    first_contact = None
    for t in range(1, 61):
        gpu.dispatch_tick(t)
        ora.step(t)
        r = gpu.read_last_tick()
        assert r.active_contact_count == int(ora.diag()[5]), t
        if first_contact is None:
            assert same_bits(gpu, ora), f"tick {t}"
            if r.active_contact_count > 0:
                first_contact = t
        else:
            gp, _ = gpu.read_bodies()
            op, _ = ora.bodies()
            np.testing.assert_allclose(gp[:, :3], op[:, :3], atol=5e-4, err_msg=f"tick {t}")
    print(f"[serial] first contact tick {first_contact}", flush=True)
    assert first_contact is not None and first_contact < 30
    assert gpu.read_last_tick().error_flags == 0
    gpu.close()


def test_more_than_64_active_contacts_sweep_in_the_sorted_order(car):
    """The car beside 96 loose boxes resting on the ground: every tick has over
    64 active contacts, and the serial sweeps over them match the oracle bit
    for bit until the car lands."""
    g = C.ConstructionGraph()
    boxes = []
    for i in range(12):
        for j in range(8):
            boxes.append(g.spawn(C.Cuboid((2, 2, 2), C.BuildPose((14 + 3 * i, 1, -12 + 3 * j)))))
    base = car_fixture()
    # Rebuild the car in the same graph after the boxes.
    chassis = g.spawn(C.Cuboid((8, 2, 12), C.BuildPose((0, 5, 0))))
    for z_units in (-4, 4):
        anchor_z = -1.0 if z_units < 0 else 1.0
        for x_units in (-3, 3):
            knuckle = g.spawn(C.Cuboid((2, 2, 2), C.BuildPose((x_units, 3, z_units))))
            wheel = g.spawn(C.Cylinder(1.0, 0.0, 0.5, C.BuildPose((-5 if x_units < 0 else 5, 3, z_units),
                                                                  C.GridRotation(0, 0, 1))))
            g.add_bearing(chassis, knuckle, (-0.75 if x_units < 0 else 0.75, 1.0, anchor_z), (0.0, -1.0, 0.0))
            g.add_bearing(knuckle, wheel, (-1.0 if x_units < 0 else 1.0, 0.75, anchor_z),
                          (-1.0, 0.0, 0.0) if x_units < 0 else (1.0, 0.0, 0.0))
    out = g.compile()
    assert len(out.bearings) == len(base.compiled.bearings) == 8
    gpu, ora = gpu_for(out.scene), oracle_for(out.scene)
    for t in range(1, 11):
        gpu.dispatch_tick(t)
        ora.step(t)
        r = gpu.read_last_tick()
        assert r.active_contact_count == int(ora.diag()[5]) > 64, t
        assert same_bits(gpu, ora), f"tick {t}"
    gpu.close()


def test_car_drop_settles_without_drift_or_ground_penetration(car):
    gpu = gpu_for(car.compiled.scene)
    m = drop(GpuSim(gpu), car)
    print(f"[serial] drop {m}", flush=True)
    assert_drop(m)
    gpu.close()


def test_struck_wheel_recovers_without_crossing_ground(car):
    gpu = gpu_for(car.compiled.scene)
    m = struck_wheel(GpuSim(gpu), car)
    print(f"[serial] struck {m}", flush=True)
    assert_struck(m)
    gpu.close()


def test_wall_impact_remains_bounded_and_decays(car):
    gpu = gpu_for(car.compiled.scene)
    m = wall_impact(GpuSim(gpu), car)
    print(f"[serial] wall {m}", flush=True)
    assert_wall(m)
    gpu.close()
