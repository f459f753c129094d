"""The reference's articulated car (device.rs:3953-4049) built through the
construction-graph lowering, and the scenario runners of its car pins
(device.rs:4224-4473), written once for any simulator that offers tick,
poses, impulse and diagnostics."""
import math
from dataclasses import dataclass
from typing import List

import numpy as np

from mechanic_amd import construction as C

ANCHOR_TOLERANCE_METERS = 1.0e-5  # lib.rs:38
AXIS_TOLERANCE_DEGREES = 1.0e-3   # lib.rs:41


@dataclass
class Car:
    compiled: object
    chassis: int
    wheels: List[int]
    dynamic: List[int]


def car_fixture() -> Car:
    """Chassis on four steering knuckles, each carrying a 1 m wheel on a
    bearing: eight bearings, ten bodies, a dynamic root, and a wall welded to
    the ground (device.rs:3953-4049)."""
    g = C.ConstructionGraph()
    chassis = g.spawn(C.Cuboid((8, 2, 12), C.BuildPose((0, 5, 0))))
    knuckles, wheels = [], []
    for z_units in (-4, 4):
        anchor_z = -1.0 if z_units < 0 else 1.0
        for x_units in (-3, 3):
            steering_x = -0.75 if x_units < 0 else 0.75
            knuckle = g.spawn(C.Cuboid((2, 2, 2), C.BuildPose((x_units, 3, z_units))))
            wheel = g.spawn(C.Cylinder(1.0, 0.0, 0.5, C.BuildPose((-5 if x_units < 0 else 5, 3, z_units),
                                                                  C.GridRotation(0, 0, 1))))
            knuckles.append(knuckle)
            wheels.append(wheel)
            g.add_bearing(chassis, knuckle, (steering_x, 1.0, anchor_z), (0.0, -1.0, 0.0))
            g.add_bearing(knuckle, wheel, (-1.0 if x_units < 0 else 1.0, 0.75, anchor_z),
                          (-1.0, 0.0, 0.0) if x_units < 0 else (1.0, 0.0, 0.0))
    wall = g.spawn(C.Cuboid((16, 8, 2), C.BuildPose((0, 4, -10))))
    g.weld(wall)
    out = g.compile()
    body = out.part_to_compound
    wheel_bodies = [body[w] for w in wheels]
    dynamic = sorted({body[chassis], *(body[k] for k in knuckles), *wheel_bodies})
    return Car(out, body[chassis], wheel_bodies, dynamic)


def _qconj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]])


def _qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def up_dot(rotation) -> float:
    """(rotation * Y) . Y."""
    x, y, z, w = np.asarray(rotation, np.float64)
    return 1.0 - 2.0 * (x * x + z * z)


def snapshot_speed(current, previous) -> float:
    """device.rs:4107-4109."""
    return float(np.linalg.norm(np.asarray(current[:3], np.float64) - np.asarray(previous[:3], np.float64)) * 60.0)


def bearing_speed(current, previous, bearing) -> float:
    """device.rs:4111-4120 with relative_bearing_rotation (device.rs:3357-3364)."""
    def relative(poses):
        return _qmul(_qconj(poses[bearing.compound_a, 4:8].astype(np.float64)),
                     poses[bearing.compound_b, 4:8].astype(np.float64))
    delta = _qmul(_qconj(relative(previous)), relative(current))
    return 2.0 * math.atan2(float(np.linalg.norm(delta[:3])), abs(float(delta[3]))) * 60.0


class OracleSim:
    def __init__(self, oracle):
        self.o = oracle

    def tick(self, t):
        self.o.step(t)

    def poses(self):
        return self.o.bodies()[0]

    def impulse(self, body, point, impulse):
        self.o.apply_impulse(body, point, impulse)

    def flags_and_residuals(self):
        d = self.o.diag()
        return int(d[0]), d[3] * 1.0e-6, d[4] * 1.0e-6


class GpuSim:
    def __init__(self, gpu):
        self.g = gpu

    def tick(self, t):
        self.g.dispatch_tick(t)

    def poses(self):
        return self.g.read_bodies()[0]

    def impulse(self, body, point, impulse):
        self.g.apply_impulse(body, point, impulse)

    def flags_and_residuals(self):
        r = self.g.read_last_tick()
        return int(r.error_flags), r.anchor_residual_meters, r.axis_residual_degrees


def drop(sim, car: Car) -> dict:
    """articulated_car_drop_settles_without_drift_or_ground_penetration
    (device.rs:4224-4308): 1,200 ticks sampled at 10..300 step 10 and 330..1200
    step 30."""
    initial = np.asarray(car.compiled.compounds[car.chassis].root_translation, np.float64)
    samples = list(range(10, 301, 10)) + list(range(330, 1201, 30))
    out = {"min_chassis_y": math.inf, "min_up": math.inf, "min_wheel_y": math.inf}
    done, previous = 0, None
    for s in samples:
        for t in range(done + 1, s + 1):
            if t == s == 1200:
                previous = sim.poses()
            sim.tick(t)
        done = s
        p = sim.poses()
        out["min_chassis_y"] = min(out["min_chassis_y"], float(p[car.chassis, 1]))
        out["min_up"] = min(out["min_up"], up_dot(p[car.chassis, 4:8]))
        out["min_wheel_y"] = min(out["min_wheel_y"], float(p[car.wheels, 1].min()))
    current = sim.poses()
    out["flags"], out["anchor_m"], out["axis_deg"] = sim.flags_and_residuals()
    out["drift"] = math.hypot(current[car.chassis, 0] - initial[0], current[car.chassis, 2] - initial[2])
    out["speed"] = max(snapshot_speed(current[b], previous[b]) for b in car.dynamic)
    out["bearing_speed"] = max(bearing_speed(current, previous, b) for b in car.compiled.bearings)
    return out


def struck_wheel(sim, car: Car) -> dict:
    """struck_articulated_wheel_recovers_without_crossing_ground
    (device.rs:4310-4373): settle 300 ticks, strike wheel 0 downward with
    3 x its mass, sample 301..360 and 370..900 step 10."""
    for t in range(1, 301):
        sim.tick(t)
    wheel = car.wheels[0]
    p = sim.poses()
    mass = car.compiled.compounds[wheel].mass_properties.mass
    sim.impulse(wheel, p[wheel, :3].astype(np.float64), np.array([0.0, -mass * 3.0, 0.0]))
    minimum = math.inf
    done = 300
    for s in list(range(301, 361)) + list(range(370, 901, 10)):
        for t in range(done + 1, s + 1):
            sim.tick(t)
        done = s
        minimum = min(minimum, float(sim.poses()[wheel, 1]))
    final = sim.poses()
    flags, _, _ = sim.flags_and_residuals()
    return {"min_wheel_y": minimum, "final_wheel_y": float(final[wheel, 1]), "flags": flags,
            "up": up_dot(final[car.chassis, 4:8])}


def wall_impact(sim, car: Car) -> dict:
    """articulated_car_wall_impact_remains_bounded_and_decays
    (device.rs:4375-4473): push the chassis toward the wall with 0.35 x its
    mass in m/s, run 1,200 ticks sampled every 30."""
    chassis_position = np.asarray(car.compiled.compounds[car.chassis].root_translation, np.float64)
    mass = car.compiled.compounds[car.chassis].mass_properties.mass
    sim.impulse(car.chassis, chassis_position, np.array([0.0, 0.0, -mass * 0.35]))
    out = {"max_speed": 0.0, "max_bearing_speed": 0.0, "min_chassis_y": math.inf, "min_wheel_y": math.inf,
           "finite": True}
    for s in range(30, 1201, 30):
        for t in range(s - 29, s + 1):
            if t == s:
                previous = sim.poses()
            sim.tick(t)
        current = sim.poses()
        out["finite"] &= bool(np.all(np.isfinite(current[car.dynamic, :3])))
        out["max_speed"] = max(out["max_speed"], max(snapshot_speed(current[b], previous[b]) for b in car.dynamic))
        out["max_bearing_speed"] = max(out["max_bearing_speed"],
                                       max(bearing_speed(current, previous, b) for b in car.compiled.bearings))
        out["min_chassis_y"] = min(out["min_chassis_y"], float(current[car.chassis, 1]))
        out["min_wheel_y"] = min(out["min_wheel_y"], float(current[car.wheels, 1].min()))
    out["flags"], out["anchor_m"], out["axis_deg"] = sim.flags_and_residuals()
    out["final_speed"] = max(snapshot_speed(current[b], previous[b]) for b in car.dynamic)
    out["final_bearing_speed"] = max(bearing_speed(current, previous, b) for b in car.compiled.bearings)
    return out


def assert_drop(m):
    assert m["min_chassis_y"] >= 0.74, m
    assert m["min_up"] > 0.9, m
    assert m["min_wheel_y"] >= 0.495, m
    assert m["flags"] == 0, m
    assert m["anchor_m"] <= ANCHOR_TOLERANCE_METERS and m["axis_deg"] <= AXIS_TOLERANCE_DEGREES, m
    assert m["drift"] < 0.05, m
    assert m["speed"] < 0.02 and m["bearing_speed"] < 0.02, m


def assert_struck(m):
    assert m["flags"] == 0, m
    assert m["min_wheel_y"] >= 0.44, m
    assert m["final_wheel_y"] >= 0.495, m
    assert m["up"] > 0.9, m


def assert_wall(m):
    assert m["finite"], m
    assert m["min_chassis_y"] >= 0.74 and m["min_wheel_y"] >= 0.49, m
    assert m["flags"] == 0, m
    assert m["anchor_m"] <= ANCHOR_TOLERANCE_METERS and m["axis_deg"] <= AXIS_TOLERANCE_DEGREES, m
    assert m["max_speed"] < 2.0 and m["max_bearing_speed"] < 10.0, m
    assert m["final_speed"] < 0.02 and m["final_bearing_speed"] < 0.02, m
