"""The reference's mechanism behaviour pins (device.rs:3112-3922, 4476-4592,
4594-4847), built through the construction-graph lowering the way the
reference tests build them, with scenario runners written once for any
simulator that offers tick, poses, impulses, coordinate and drive writes and
diagnostics (the GPU through the C ABI, or the oracle).

Each runner returns the quantities the reference test asserts on; the
assertions and their thresholds live beside the scene (`check_*`), so the GPU
and the oracle are held to the same reference figures.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence

import numpy as np

from mechanic_amd import _lib
from mechanic_amd import construction as C

HALF_PI = math.pi / 2.0


# ---- simulators ---------------------------------------------------------------

class GpuSim:
    """A GPU handle. The reference's run helpers use GpuPhysicsConfig with
    collisions on or off, self-collisions on, and 8 or 16 iterations."""

    def __init__(self, compiled, collisions=True, iterations=8):
        import mechanic_amd as mh

        cfg = mh.GpuPhysicsConfig(collisions_enabled=collisions, mechanism_self_collisions=True,
                                  solver_iterations=iterations)
        self.g = mh.GpuPhysics.new_with_config(compiled.scene, cfg)
        self.done = 0

    def run_to(self, tick):
        if tick > self.done:
            self.g.dispatch_ticks(self.done + 1, tick - self.done)
            self.done = tick

    def poses(self):
        # read_snapshot_transforms(tick % 3), as the reference tests read it.
        return self.g.read_snapshot_transforms(self.done % 3)

    def flags(self):
        return int(self.g.read_last_tick().error_flags)

    def active_contacts(self):
        return int(self.g.read_last_tick().active_contact_count)

    def impulse(self, body, point, impulse):
        self.g.apply_impulse(body, np.asarray(point, np.float32), np.asarray(impulse, np.float32))

    def init_coordinates(self, rows):
        self.g.initialize_mechanism_coordinates([_lib.MechanismCoordinate(a, w) for a, w in rows])

    def write_drives(self, rows):
        self.g.write_mechanism_drives(rows)

    def close(self):
        self.g.close()


class OracleSim:
    def __init__(self, compiled, collisions=True, iterations=8):
        from tests._oracle import Oracle

        self.o = Oracle(compiled.scene, iterations=iterations, collisions=collisions, self_collisions=True)
        self.done = 0

    def run_to(self, tick):
        for t in range(self.done + 1, tick + 1):
            self.o.step(t)
        self.done = max(self.done, tick)

    def poses(self):
        return self.o.bodies()[0]

    def flags(self):
        return int(self.o.diag()[0])

    def active_contacts(self):
        return int(self.o.diag()[5])

    def impulse(self, body, point, impulse):
        self.o.apply_impulse(body, point, impulse)

    def init_coordinates(self, rows):
        self.o.set_coordinates(np.asarray(rows, np.float32))

    def write_drives(self, rows):
        self.o.set_drives(rows)

    def close(self):
        pass


# ---- quaternions and bearing measurements ---------------------------------------

def qconj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]], np.float64)


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def qrotate(q, v):
    p = qmul(qmul(q, np.array([v[0], v[1], v[2], 0.0])), qconj(q))
    return p[:3]


def relative_rotation(poses, bearing):
    """relative_bearing_rotation (device.rs:3357-3364): conj(q_a) * q_b."""
    return qmul(qconj(poses[bearing.compound_a, 4:8].astype(np.float64)),
                poses[bearing.compound_b, 4:8].astype(np.float64))


def bearing_angle(poses, bearing):
    """2 atan2(rel.xyz . local_axis_a, rel.w) (device.rs:3391-3397, 3684-3693);
    compounds compile with identity rotation, so the local axis is the axis."""
    rel = relative_rotation(poses, bearing)
    axis = np.asarray(bearing.axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    return 2.0 * math.atan2(float(rel[:3] @ axis), float(rel[3]))


def bearing_speed(current, previous, bearing):
    """2 atan2(|delta.xyz|, |delta.w|) * 60 over one tick (device.rs:3398-3399)."""
    delta = qmul(qconj(relative_rotation(previous, bearing)), relative_rotation(current, bearing))
    return 2.0 * math.atan2(float(np.linalg.norm(delta[:3])), abs(float(delta[3]))) * 60.0


def angle_between(q1, q2):
    """Quat::angle_between: 2 acos(|q1 . q2|)."""
    d = min(1.0, abs(float(np.dot(np.asarray(q1, np.float64), np.asarray(q2, np.float64)))))
    return 2.0 * math.acos(d)


def current_and_previous(sim, tick):
    sim.run_to(tick - 1)
    previous = sim.poses()
    sim.run_to(tick)
    return sim.poses(), previous


def half_grid(v) -> C.BuildPose:
    """BuildPose::from_half_grid (geometry.rs:160-176): eighth-metre centre units."""
    units = tuple((c - c % 2) // 2 for c in v)
    half = tuple(c % 2 for c in v)
    return C.BuildPose(units, C.GridRotation(), half)


def cube(units, size=4) -> C.Cuboid:
    return C.Cuboid((size, size, size), C.BuildPose(tuple(units)))


# ---- scenes (device.rs test helpers) ---------------------------------------------

def pendulum(grounded: bool):
    """pendulum_creation (device.rs:3112-3148): a root cube hinged to a
    two-cube welded arm about X at (0.5, 0.5, 0)."""
    g = C.ConstructionGraph()
    root, arm_a, arm_b = g.spawn(cube((0, 2, 0))), g.spawn(cube((4, 2, 0))), g.spawn(cube((4, 2, 4)))
    g.weld(arm_a, arm_b)
    if grounded:
        g.weld(root)
    g.add_bearing(root, arm_a, (0.5, 0.5, 0.0), (1.0, 0.0, 0.0))
    return g.compile()


def tall_pendulum(second_link: bool):
    """tall_pendulum_creation (device.rs:3150-3227): a grounded seven-cube
    support, a four-cube arm hinged about X at its top and, optionally, a
    three-cube child hinged about Z at the arm's end."""
    g = C.ConstructionGraph()
    support = [g.spawn(cube((0, 2 * r + 1, 0), 2)) for r in range(7)]
    arm = [g.spawn(cube((2, 13, 2 * c), 2)) for c in range(4)]
    child = [g.spawn(cube((2, 2 * r + 13, 8), 2)) for r in range(3)] if second_link else None
    g.weld(support[0])
    for a, b in zip(support, support[1:]):
        g.weld(a, b)
    for a, b in zip(arm, arm[1:]):
        g.weld(a, b)
    g.add_bearing(support[6], arm[0], (0.25, 3.25, 0.0), (1.0, 0.0, 0.0))
    if child:
        for a, b in zip(child, child[1:]):
            g.weld(a, b)
        g.add_bearing(arm[3], child[0], (0.5, 3.25, 1.75), (0.0, 0.0, 1.0))
    return g.compile()


def branching_pendulum(arm_count: int, hanging: bool):
    """branching_pendulum_creation (device.rs:3229-3301): a grounded root, a
    bar on a vertical bearing above it, and one or two arms hinged to the
    bar's ends."""
    g = C.ConstructionGraph()
    root = g.spawn(C.Cuboid((8, 4, 8), C.BuildPose((0, 2, 0))))
    g.weld(root)
    bar = g.spawn(C.Cuboid((24, 2, 2), half_grid((0, 10, 0))))
    g.add_bearing(root, bar, (0.0, 1.0, 0.0), (0.0, 1.0, 0.0))
    y = 7 if hanging else 16
    arms = [((26, y, 1), (3.0, 1.25, 0.0), (1.0, 0.0, 0.0)), ((-26, y, -1), (-3.0, 1.25, 0.0), (-1.0, 0.0, 0.0))]
    for center, anchor, axis in arms[:arm_count]:
        arm = g.spawn(C.Cuboid((2, 8, 2), half_grid(center)))
        g.add_bearing(bar, arm, anchor, axis)
    return g.compile()


def coaxial_rotor():
    """swinging_arm_with_coaxial_rotor_creation (device.rs:3303-3355): a
    grounded block, a long arm hinged about X and a hollow rotor hinged about
    -Y at the arm's far end."""
    g = C.ConstructionGraph()
    root = g.spawn(C.Cuboid((8, 8, 8), C.BuildPose((0, 4, 0))))
    g.weld(root)
    arm = g.spawn(C.Cuboid((2, 2, 24), half_grid((10, 10, 24))))
    g.add_bearing(root, arm, (1.0, 1.25, 0.25), (1.0, 0.0, 0.0))
    rotor = g.spawn(C.Cylinder(0.25, 0.20, 0.75, half_grid((10, 5, 46))))
    g.add_bearing(arm, rotor, (1.25, 1.0, 5.75), (0.0, -1.0, 0.0))
    return g.compile()


def driven_arm(axis: str, limits: C.DriveLimits, target, loaded: bool = False):
    """driven_arm (device.rs:3583-3670): a grounded base and a hinged arm
    (optionally with an outrigger welded off the axis) wired to one control
    block. Returns the graph (for live reprogramming) and its compilation."""
    g = C.ConstructionGraph()
    if axis == "x":
        arm_units, anchor, axis_v = (4, 2, 0), (0.5, 0.5, 0.0), (1.0, 0.0, 0.0)
    else:
        arm_units, anchor, axis_v = (0, 6, 0), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0)
    base = g.spawn(cube((0, 2, 0)))
    arm = g.spawn(cube(arm_units))
    outrigger = g.spawn(cube((arm_units[0], arm_units[1], arm_units[2] + 4))) if loaded else None
    g.weld(base)
    if outrigger is not None:
        g.weld(arm, outrigger)
    bearing = g.add_bearing(base, arm, anchor, axis_v)
    controller = g.spawn(C.Controller(C.BuildPose((0, 40, 0))))
    g.wire(C.DriveWire(controller, bearing, target, limits))
    return g, g.compile()


def cylinder_drop(inner_diameter: float, drop_x_half_units: int):
    """cylinder_drop_creation (device.rs:4594-4641): a 1 m cylinder (solid or
    with a 0.6 m bore) rigidly linked to a grounded block, and a 0.25 m cube
    dropped from 5 m onto its axis or its ring."""
    g = C.ConstructionGraph()
    support = g.spawn(cube((8, 2, 0)))
    g.weld(support)
    cylinder = g.spawn(C.Cylinder(1.0, inner_diameter, 1.0, C.BuildPose((0, 12, 0))))
    g.rigid_link(support, cylinder)
    g.spawn(C.Cuboid((1, 1, 1), half_grid((drop_x_half_units, 40, 0))))
    return g.compile()


def tower_with_arm():
    """contact_supported_unwelded_tower_drives_welded_arm (device.rs:4670-4721):
    four unwelded stacked cubes, a three-cube welded arm hinged about X to the
    top one."""
    g = C.ConstructionGraph()
    parts = [g.spawn(cube((0, y, 0))) for y in (2, 6, 10, 14)] + [g.spawn(cube((4, 14, z))) for z in (0, 4, 8)]
    g.weld(parts[4], parts[5])
    g.weld(parts[5], parts[6])
    g.add_bearing(parts[3], parts[4], (0.5, 3.5, 0.0), (1.0, 0.0, 0.0))
    return g.compile()


def double_pendulum():
    """double_pendulum_transfers_motion_through_both_bearings
    (device.rs:4741-4781): a grounded cube and two two-cube welded links in
    series about X."""
    g = C.ConstructionGraph()
    s = [g.spawn(cube(u)) for u in ((0, 2, 0), (4, 2, 0), (4, 2, 4), (8, 2, 0), (8, 2, 4))]
    g.weld(s[1], s[2])
    g.weld(s[3], s[4])
    g.weld(s[0])
    g.add_bearing(s[0], s[1], (0.5, 0.5, 0.0), (1.0, 0.0, 0.0))
    g.add_bearing(s[1], s[3], (1.5, 0.5, 0.0), (1.0, 0.0, 0.0))
    return g.compile()


def balanced_child():
    """balanced_child_contacts_move_root_without_spurious_joint_motion
    (device.rs:4795-4834): a free root cube hinged about X to a welded
    three-cube child that rests on the ground."""
    g = C.ConstructionGraph()
    s = [g.spawn(cube(u)) for u in ((0, 6, 0), (4, 6, 0), (4, 2, 0), (4, 2, 4))]
    g.weld(s[1], s[2])
    g.weld(s[2], s[3])
    g.add_bearing(s[0], s[1], (0.5, 1.5, 0.0), (1.0, 0.0, 0.0))
    return g.compile()


# ---- scenarios and their reference thresholds -------------------------------------

Make = Callable[..., object]  # (compiled, collisions=, iterations=) -> sim


@dataclass
class LongPendulum:
    angles: List[float]
    speeds: List[float]
    flags: int
    active: int


def run_long_pendulum(make: Make, compiled, collisions: bool) -> LongPendulum:
    """run_long_pendulum (device.rs:3366-3404): 1,200 ticks, default config."""
    sim = make(compiled, collisions=collisions)
    current, previous = current_and_previous(sim, 1200)
    out = LongPendulum([bearing_angle(current, b) for b in compiled.bearings],
                       [bearing_speed(current, previous, b) for b in compiled.bearings], sim.flags(),
                       sim.active_contacts())
    sim.close()
    return out


def check_tall_pendulums(make: Make):
    """tall_single_and_double_pendulums_dissipate_energy (device.rs:3406-3444)."""
    single = tall_pendulum(False)
    assert len(single.colliders) == 11
    r = run_long_pendulum(make, single, True)
    assert r.flags == 0 and r.active == 0, r
    assert abs(r.angles[0] - HALF_PI) < 0.02, r
    assert r.speeds[0] < 0.2, r
    double = tall_pendulum(True)
    assert len(double.colliders) == 14
    free = run_long_pendulum(make, double, False)
    assert free.flags == 0, free
    assert abs(free.angles[0] - HALF_PI) < 0.02 and abs(free.angles[1] + HALF_PI) < 0.02, free
    assert all(s < 1.0e-4 for s in free.speeds), free
    contact = run_long_pendulum(make, double, True)
    assert contact.flags == 0, contact
    assert all(s < 1.0e-4 for s in contact.speeds), contact
    return {"single": r, "double_free": free, "double_contact": contact}


def early_late_speeds(make: Make, compiled, collisions: bool, coordinates=None):
    """Bearing speeds at ticks 1,200 and 2,400 (device.rs:3476-3519, 3543-3580)."""
    sim = make(compiled, collisions=collisions)
    if coordinates is not None:
        sim.init_coordinates(coordinates)
    cur, prev = current_and_previous(sim, 1200)
    early = [bearing_speed(cur, prev, b) for b in compiled.bearings]
    cur, prev = current_and_previous(sim, 2400)
    late = [bearing_speed(cur, prev, b) for b in compiled.bearings]
    flags = sim.flags()
    sim.close()
    return early, late, flags


def check_branching_pendulums(make: Make):
    """branching_pendulum_reaches_rest_instead_of_retaining_spin (device.rs:3446-3520)."""
    out = {}
    for arm_count, hanging in ((1, False), (2, False), (1, True), (2, True)):
        compiled = branching_pendulum(arm_count, hanging)
        coords = [(0.0, 0.0), (0.35, 0.0)] if (arm_count == 1 and hanging) else None
        early, late, flags = early_late_speeds(make, compiled, False, coords)
        assert flags == 0, (arm_count, hanging, flags)
        assert all(s < 1.0e-4 for s in late) and all(l <= e + 1.0e-5 for l, e in zip(late, early)), \
            f"{arm_count}-arm hanging={hanging} bearing speeds grew from {early} to {late}"
        out[(arm_count, hanging)] = (early, late)
    return out


def check_coaxial_rotor(make: Make):
    """moving_coaxial_rotor_does_not_gain_perpetual_spin (device.rs:3522-3580)."""
    compiled = coaxial_rotor()
    early, late, flags = early_late_speeds(make, compiled, True, [(0.35, 0.0), (0.0, 0.0)])
    assert flags == 0, flags
    assert all(s < 0.02 for s in late) and all(l <= e + 1.0e-5 for l, e in zip(late, early)), \
        f"bearing state gained speed from {early} to {late}"
    return early, late


def run_driven_arm(make: Make, compiled, ticks: int, drives=None):
    """run_driven_arm (device.rs:3695-3718) with joint_angle (device.rs:3683-3693):
    collisions off, default config."""
    sim = make(compiled, collisions=False)
    if drives is not None:
        sim.write_drives(drives)
    sim.run_to(ticks)
    angle = bearing_angle(sim.poses(), compiled.bearings[0])
    flags = sim.flags()
    sim.close()
    return angle, flags


def limits(max_speed, torque, stops=None) -> C.DriveLimits:
    return C.DriveLimits(max_speed, torque, stops)


DRIVE_CASES = {
    # name: (axis, limits, target, loaded, ticks, expected angle, tolerance), device.rs:3720-3860
    "speed": ("x", (3.0, math.inf, None), ("speed", 1.0), 60, 1.0, 0.05),
    "negative_speed": ("x", (3.0, math.inf, None), ("speed", -1.0), 60, -1.0, 0.05),
    "max_speed_cap": ("x", (0.5, math.inf, None), ("speed", 3.0), 60, 0.5, 0.05),
    "angle_target": ("x", (3.0, 400.0, None), ("angle", 0.8), 240, 0.8, 0.02),
    "travel_limit": ("x", (3.0, math.inf, (-0.2, 0.2)), ("speed", 1.0), 120, 0.2, 0.02),
}


def check_drive_case(make: Make, name: str):
    axis, lim, target, ticks, expected, tolerance = DRIVE_CASES[name]
    _, compiled = driven_arm(axis, limits(*lim), target)
    angle, flags = run_driven_arm(make, compiled, ticks)
    assert flags == 0, (name, flags)
    assert abs(angle - expected) < tolerance, f"{name}: angle {angle}, expected {expected} +- {tolerance}"
    return angle


def check_weak_drive_stalls(make: Make):
    """weak_drive_stalls_lifting_a_gravity_loaded_arm (device.rs:3795-3819)."""
    target = ("speed", -1.0)
    _, strong = driven_arm("x", limits(3.0, math.inf), target, loaded=True)
    _, weak = driven_arm("x", limits(3.0, 0.5), target, loaded=True)
    strong_angle, strong_flags = run_driven_arm(make, strong, 60)
    weak_angle, weak_flags = run_driven_arm(make, weak, 60)
    assert strong_flags == 0 and weak_flags == 0
    assert abs(strong_angle + 1.0) < 0.05, strong_angle
    assert weak_angle > strong_angle + 0.5, (weak_angle, strong_angle)
    return strong_angle, weak_angle


def check_vertical_drive_escapes_sleep_clamp(make: Make):
    """vertical_axis_drive_is_not_zeroed_by_the_sleep_clamp (device.rs:3842-3860)."""
    _, compiled = driven_arm("y", limits(3.0, math.inf), ("speed", 0.002))
    angle, flags = run_driven_arm(make, compiled, 600)
    assert flags == 0
    assert angle > 0.015, angle
    return angle


def check_live_reprogram(make: Make, count_error: type):
    """reprogramming_a_wire_changes_the_drive_without_reloading_the_scene
    (device.rs:3862-3922): an empty drive write is refused with the drive
    count; a new program written to the loaded scene drives -2 rad/s."""
    graph, compiled = driven_arm("x", limits(3.0, math.inf), ("speed", 1.0))
    sim = make(compiled, collisions=False)
    try:
        sim.write_drives([])
    except count_error as e:
        refused = e
    else:
        raise AssertionError("an empty drive write was accepted")
    graph.wires[0] = C.DriveWire(graph.wires[0].controller, graph.wires[0].bearing, ("speed", -2.0),
                                 limits(3.0, math.inf))
    rows = [d.row() for d in graph.compile().coordinate_drives]
    sim.write_drives(rows)
    sim.run_to(60)
    angle = bearing_angle(sim.poses(), compiled.bearings[0])
    flags = sim.flags()
    sim.close()
    assert flags == 0
    assert abs(angle + 2.0) < 0.1, angle
    return refused, angle


def check_external_impulse_drives_bearing(make: Make):
    """external_impulse_drives_a_bearing_coordinate (device.rs:4476-4515)."""
    compiled = pendulum(True)
    child = compiled.bearings[0].compound_b

    def run(impulse):
        sim = make(compiled, collisions=False)
        if impulse is not None:
            sim.impulse(child, np.asarray(compiled.compounds[child].root_translation, np.float32), impulse)
        sim.run_to(10)
        rotation = sim.poses()[child, 4:8].copy()
        sim.close()
        return rotation

    baseline, struck = run(None), run(np.array([0.0, -500.0, 0.0], np.float32))
    angle = angle_between(baseline, struck)
    assert angle > 0.01, angle
    return angle


def run_ticks(make: Make, compiled, ticks: int, collisions: bool):
    """run_ticks (device.rs:4517-4543): self-collisions on, 16 iterations."""
    sim = make(compiled, collisions=collisions, iterations=16)
    sim.run_to(ticks)
    poses, flags = sim.poses(), sim.flags()
    sim.close()
    return poses, flags


def check_grounded_offset_pendulum(make: Make):
    """grounded_offset_pendulum_swings_without_detaching (device.rs:4569-4579)."""
    compiled = pendulum(True)
    poses, flags = run_ticks(make, compiled, 30, False)
    assert flags == 0
    assert abs(poses[1, 4]) > 1.0e-3, poses[1]
    root = np.asarray(compiled.compounds[0].root_translation, np.float64)
    assert np.all(np.abs(poses[0, :3] - root) <= 1.0e-5), (poses[0], root)
    return poses


def check_freely_falling_hinge(make: Make):
    """freely_falling_hinge_has_no_gravity_induced_relative_rotation (device.rs:4581-4592)."""
    compiled = pendulum(False)
    poses, flags = run_ticks(make, compiled, 30, False)
    assert flags == 0
    assert poses[0, 1] < compiled.compounds[0].root_translation[1] - 0.5, poses[0]
    assert np.all(np.abs(poses[1, 4:7]) < 1.0e-4), poses[1]
    return poses


def check_cylinder_bore(make: Make):
    """gpu_cylinder_bore_is_passable_and_annular_material_blocks_motion (device.rs:4643-4668)."""
    out = {}
    for name, inner, x in (("hollow", 0.60, 0), ("solid", 0.0, 0), ("annular", 0.60, 3)):
        poses, flags = run_ticks(make, cylinder_drop(inner, x), 60, True)
        assert flags == 0, (name, flags)
        out[name] = float(poses[1, 1])
    assert out["hollow"] < 2.5, out
    assert out["solid"] > 3.4, out
    assert out["annular"] > 3.35, out
    return out


def check_tower_drives_welded_arm(make: Make):
    """contact_supported_unwelded_tower_drives_welded_arm (device.rs:4670-4739)."""
    compiled = tower_with_arm()
    poses, flags = run_ticks(make, compiled, 90, True)
    assert flags == 0
    b = compiled.bearings[0]
    a_i, b_i = b.compound_a, b.compound_b
    anchor_a = poses[a_i, :3].astype(np.float64) + qrotate(poses[a_i, 4:8].astype(np.float64), b.local_anchor_a)
    anchor_b = poses[b_i, :3].astype(np.float64) + qrotate(poses[b_i, 4:8].astype(np.float64), b.local_anchor_b)
    assert np.all(np.abs(anchor_a - anchor_b) <= 1.0e-5), (anchor_a, anchor_b)
    assert abs(poses[b_i, 4]) > 1.0e-3, poses[b_i]
    assert poses[a_i, 1] > 0.5, poses[a_i]
    return poses


def check_double_pendulum(make: Make):
    """double_pendulum_transfers_motion_through_both_bearings (device.rs:4741-4793)."""
    poses, flags = run_ticks(make, double_pendulum(), 30, False)
    assert flags == 0
    first, second = poses[1, 4:8].astype(np.float64), poses[2, 4:8].astype(np.float64)
    assert abs(first[0]) > 1.0e-3, first
    assert abs(qmul(qconj(first), second)[0]) > 1.0e-4, (first, second)
    return poses


def check_balanced_child(make: Make):
    """balanced_child_contacts_move_root_without_spurious_joint_motion (device.rs:4795-4847)."""
    compiled = balanced_child()
    free, _ = run_ticks(make, compiled, 8, False)
    contact, flags = run_ticks(make, compiled, 8, True)
    assert flags == 0
    root, child = compiled.bearings[0].compound_a, compiled.bearings[0].compound_b
    assert contact[root, 1] > free[root, 1] + 1.0e-4, (contact[root], free[root])
    rel = qmul(qconj(contact[root, 4:8].astype(np.float64)), contact[child, 4:8].astype(np.float64))
    assert abs(rel[0]) < 1.0e-4, rel
    return free, contact


SCENES = {
    "pendulum_grounded": lambda: pendulum(True),
    "pendulum_free": lambda: pendulum(False),
    "tall_single": lambda: tall_pendulum(False),
    "tall_double": lambda: tall_pendulum(True),
    "branching_2_hanging": lambda: branching_pendulum(2, True),
    "coaxial_rotor": coaxial_rotor,
    "driven_arm_loaded": lambda: driven_arm("x", limits(3.0, 0.5), ("speed", -1.0), loaded=True)[1],
    "cylinder_annular": lambda: cylinder_drop(0.60, 3),
    "tower_with_arm": tower_with_arm,
    "double_pendulum": double_pendulum,
    "balanced_child": balanced_child,
}
