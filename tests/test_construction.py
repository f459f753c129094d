"""Construction-graph lowering against the reference's compile known answers
(crates/mechanic-core/src/compile.rs:1011-1773, same inputs and tolerances)
and the full-cylinder ground data of device.rs:3008-3045."""
import math

import numpy as np
import pytest

from mechanic_amd import construction as C


def cube(units):
    return C.Cuboid((4, 4, 4), C.BuildPose(units))


def spawn(g, units):
    return g.spawn(cube(units))


def grounded(g, part):
    g.weld(part, C.GROUND)


def test_hollow_cylinder_exact_mass_inertia_and_sixteen_colliders():
    # compile.rs:1011-1038
    g = C.ConstructionGraph()
    g.spawn(C.Cylinder(1.0, 0.5, 2.0))
    out = g.compile()
    p = out.compounds[0].mass_properties
    outer, inner = 0.5, 0.25
    mass = C.CUBOID_DENSITY_KG_M3 * math.pi * (outer * outer - inner * inner) * 2.0
    axial = mass * (outer * outer + inner * inner) * 0.5
    transverse = mass * (3.0 * (outer * outer + inner * inner) + 4.0) / 12.0
    assert len(out.colliders) == 16
    assert abs(p.mass - mass) < 1e-3
    assert np.allclose(p.center_of_mass, 0.0, atol=1e-6)
    assert abs(p.inertia[0, 0] - transverse) < 1e-3
    assert abs(p.inertia[1, 1] - axial) < 1e-3
    assert abs(p.inertia[2, 2] - transverse) < 1e-3
    for c in out.colliders:
        assert abs(c.half_extents[1] - 1.0) < 1e-6
        assert np.linalg.norm(c.local_center) >= inner - 1e-6


def test_cylinder_sector_exact_offset_mass_properties():
    # compile.rs:1041-1095
    g = C.ConstructionGraph()
    g.spawn(C.Cylinder(1.0, 0.5, 2.0, sweep_degrees=90))
    out = g.compile()
    p = out.compounds[0].mass_properties
    outer, inner, length, sweep = 0.5, 0.25, 2.0, math.pi / 2
    mass = C.CUBOID_DENSITY_KG_M3 * sweep * (outer * outer - inner * inner) * length * 0.5
    cx = 4.0 * math.sin(sweep * 0.5) * (outer ** 3 - inner ** 3) / (3.0 * sweep * (outer * outer - inner * inner))
    r2 = outer * outer + inner * inner
    par = r2 * (sweep + math.sin(sweep)) / (4.0 * sweep)
    perp = r2 * (sweep - math.sin(sweep)) / (4.0 * sweep)
    assert len(out.colliders) == 16
    assert abs(p.mass - mass) < 1e-3
    assert np.allclose(p.center_of_mass, [cx, 0.0, 0.0], atol=1e-6)
    assert abs(p.inertia[0, 0] - mass * (length * length / 12.0 + perp)) < 1e-3
    assert abs(p.inertia[1, 1] - mass * (par + perp - cx * cx)) < 1e-3
    assert all((c.local_center + p.center_of_mass)[0] > 0.0 for c in out.colliders)


def test_welds_compile_to_one_compound_with_parallel_axis_inertia():
    # compile.rs:1117-1141
    g = C.ConstructionGraph()
    a = spawn(g, (0, 0, 0))
    b = spawn(g, (4, 0, 0))
    g.weld(a, b)
    out = g.compile()
    p = out.compounds[0].mass_properties
    assert len(out.compounds) == 1 and len(out.colliders) == 2
    assert abs(p.mass - 1000.0) < 1e-3
    assert np.allclose(p.center_of_mass, [0.5, 0.0, 0.0], atol=1e-6)
    assert abs(p.inertia[0, 0] - 166.66667) < 1e-3
    assert abs(p.inertia[1, 1] - 416.6667) < 1e-3
    assert abs(p.inertia[2, 2] - 416.6667) < 1e-3
    # The rows the GPU receives: colliders at -0.5 / +0.5 m from the root.
    cols = out.scene.collider_array()
    assert np.allclose(cols[:, 0], [-0.5, 0.5]) and np.all(cols[:, 12].view(np.uint32) == 0)


def test_welding_to_ground_makes_only_that_group_static():
    # compile.rs:1203-1219
    g = C.ConstructionGraph()
    a = spawn(g, (0, 2, 0))
    spawn(g, (8, 2, 0))
    grounded(g, a)
    out = g.compile()
    assert len(out.compounds) == 2
    assert out.compounds[0].is_static and abs(out.compounds[0].mass_properties.inverse_mass) < 1e-7
    assert not out.compounds[1].is_static
    assert out.scene.inverse_masses[0] == 0.0 and out.scene.inverse_masses[1] > 0.0


def test_control_block_is_one_quarter_metre_cuboid():
    # compile.rs:1537-1550
    g = C.ConstructionGraph()
    g.spawn(C.Controller())
    out = g.compile()
    assert len(out.colliders) == 1
    assert np.array_equal(out.colliders[0].half_extents, [0.125, 0.125, 0.125])
    assert abs(out.compounds[0].mass_properties.mass - C.CUBOID_DENSITY_KG_M3 * 0.25 ** 3) < 1e-4


def test_bearing_collapsed_by_later_weld_is_rejected():
    # compile.rs:1222-1251
    g = C.ConstructionGraph()
    a = spawn(g, (0, 0, 0))
    b = spawn(g, (4, 0, 0))
    g.add_bearing(a, b, (0.5, 0.0, 0.0), (1, 0, 0))
    g.weld(a, b)
    with pytest.raises(C.TopologyError) as e:
        g.compile()
    assert e.value.kind == "SelfBearing"


def square_loop(g):
    a, b, c, d = (spawn(g, u) for u in ((0, 0, 0), (4, 0, 0), (4, 4, 0), (0, 4, 0)))
    return [g.add_bearing(a, b, (0.5, 0.0, 0.0), (1, 0, 0)), g.add_bearing(b, c, (1.0, 0.5, 0.0), (0, 1, 0)),
            g.add_bearing(c, d, (0.5, 1.0, 0.0), (-1, 0, 0)), g.add_bearing(d, a, (0.0, 0.5, 0.0), (0, -1, 0))]


def test_closed_square_has_one_hard_closure_edge():
    # compile.rs:1280-1340
    g = C.ConstructionGraph()
    square_loop(g)
    out = g.compile()
    assert len(out.tree_bearings) == 3 and len(out.closure_bearings) == 1
    assert list(out.mechanisms.body_components) == [0, 0, 0, 0]
    assert len(out.suppression) == 4


def test_floating_branch_has_one_canonical_root_and_stable_traversals():
    # compile.rs:1342-1375
    g = C.ConstructionGraph()
    root = spawn(g, (0, 0, 0))
    x_child = spawn(g, (4, 0, 0))
    y_child = spawn(g, (0, 4, 0))
    g.add_bearing(root, x_child, (0.5, 0.0, 0.0), (1, 0, 0))
    g.add_bearing(root, y_child, (0.0, 0.5, 0.0), (0, 1, 0))
    m = g.compile().mechanisms
    assert list(m.root_flags) == [1, 0, 0]
    assert list(m.parents[1:]) == [0, 0]
    assert [int(m.mechanism_bodies[b].traversal[2]) for b in (1, 2)] == [1, 2]


def test_grounded_body_is_root_even_when_not_the_first_compound():
    # compile.rs:1377-1398
    g = C.ConstructionGraph()
    child = spawn(g, (0, 6, 0))
    root = spawn(g, (0, 2, 0))
    grounded(g, root)
    g.add_bearing(child, root, (0.0, 1.0, 0.0), (0, -1, 0))
    m = g.compile().mechanisms
    assert list(m.root_flags) == [0, 1]
    assert int(m.parents[0]) == 1
    assert int(m.mechanism_bodies[0].metadata[2]) == 1  # traversed from B to A


def test_multiple_ground_anchors_remain_fixed_roots_with_a_closure_edge():
    # compile.rs:1400-1433
    g = C.ConstructionGraph()
    left, middle, right = spawn(g, (0, 2, 0)), spawn(g, (4, 2, 0)), spawn(g, (8, 2, 0))
    grounded(g, left)
    grounded(g, right)
    g.add_bearing(left, middle, (0.5, 0.5, 0.0), (1, 0, 0))
    g.add_bearing(middle, right, (1.5, 0.5, 0.0), (1, 0, 0))
    out = g.compile()
    assert len(out.tree_bearings) == 1 and len(out.closure_bearings) == 1
    assert list(out.mechanisms.root_flags) == [1, 0, 1]
    assert int(out.mechanisms.parents[1]) == 0


def test_driven_bearing_is_preferred_as_a_tree_edge_and_never_a_closure():
    # compile.rs:1553-1580
    g = C.ConstructionGraph()
    bearings = square_loop(g)
    g.wire(C.DriveWire(controller=g.spawn(C.Controller(C.BuildPose((0, 40, 0)))), bearing=bearings[3]))
    out = g.compile()
    assert bearings[3] in out.tree_bearings and bearings[3] not in out.closure_bearings
    # A driven bearing that no forest can take as a tree edge is rejected.
    g = C.ConstructionGraph()
    left, right = spawn(g, (0, 2, 0)), spawn(g, (4, 2, 0))
    grounded(g, left)
    grounded(g, right)
    b = g.add_bearing(left, right, (0.5, 0.5, 0.0), (1, 0, 0))
    g.wire(C.DriveWire(controller=0, bearing=b))
    with pytest.raises(C.TopologyError) as e:
        g.compile()
    assert e.value.kind == "DrivenClosureBearing"


def test_coordinate_axis_inertia_matches_a_hand_computed_arm():
    # compile.rs:1582-1608: a 1 m cube centred on the hinge axis, own x inertia only.
    g = C.ConstructionGraph()
    base = spawn(g, (0, 2, 0))
    grounded(g, base)
    arm = spawn(g, (4, 2, 0))
    g.add_bearing(base, arm, (0.5, 0.5, 0.0), (1, 0, 0))
    inertia = g.compile().coordinate_axis_inertia[0]
    assert abs(inertia - C.CUBOID_DENSITY_KG_M3 * 2.0 / 12.0) < 1e-2


def chain_of_two(g):
    base = spawn(g, (0, 2, 0))
    grounded(g, base)
    first, second = spawn(g, (0, 6, 0)), spawn(g, (0, 10, 0))
    lower = g.add_bearing(base, first, (0.0, 1.0, 0.0), (0, 1, 0))
    upper = g.add_bearing(first, second, (0.0, 2.0, 0.0), (0, 1, 0))
    return lower, upper


def test_coordinate_axis_inertia_includes_the_whole_child_subtree():
    # compile.rs:1610-1646
    g = C.ConstructionGraph()
    chain_of_two(g)
    inertia = g.compile().coordinate_axis_inertia
    assert len(inertia) == 2 and inertia[0] > inertia[1]


def test_duplicate_bearing_rows_share_one_coordinate():
    # compile.rs:1648-1681
    g = C.ConstructionGraph()
    base = spawn(g, (0, 2, 0))
    grounded(g, base)
    arm = spawn(g, (4, 2, 0))
    first = g.add_bearing(base, arm, (0.5, 0.5, 0.0), (1, 0, 0))
    duplicate = g.add_bearing(base, arm, (0.5, 0.5, 0.0), (1, 0, 0))
    out = g.compile()
    assert len(out.tree_bearings) == 1
    assert out.bearing_coordinates[first] == out.bearing_coordinates[duplicate]


def test_closure_bearings_have_no_coordinate_to_drive():
    # compile.rs:1683-1700
    g = C.ConstructionGraph()
    bearings = square_loop(g)
    out = g.compile()
    for b in out.closure_bearings:
        assert b not in out.bearing_coordinates
    for b in out.tree_bearings:
        assert b in out.bearing_coordinates
    assert len(out.bearing_coordinates) == len(bearings) - len(out.closure_bearings)


def test_coordinate_drives_apply_per_wire_reverse_and_leave_undriven_rows_passive():
    # compile.rs:1702-1750
    g = C.ConstructionGraph()
    lower, _ = chain_of_two(g)
    ctrl = g.spawn(C.Controller(C.BuildPose((0, 40, 0))))
    g.wire(C.DriveWire(ctrl, lower, ("speed", 3.0), C.DriveLimits(4.0, 20.0, (-1.0, 1.0)), reversed=True))
    out = g.compile()
    assert len(out.coordinate_drives) == 2
    k = out.tree_bearings.index(lower)
    row = out.coordinate_drives[k]
    assert row.mode == 1 and abs(row.target_speed + 3.0) < 1e-7
    assert abs(row.min_angle + 1.0) < 1e-7 and abs(row.max_angle - 1.0) < 1e-7
    assert abs(row.max_acceleration - 20.0 / out.coordinate_axis_inertia[k]) < 1e-3
    passive = out.coordinate_drives[1 - k]
    assert passive.mode == 0 and passive.max_acceleration == 0.0 and math.isinf(passive.max_angle)
    # The uploaded drive rows carry the same values.
    up = out.scene.drives[k]
    assert up.mode == 1 and abs(up.target_speed + 3.0) < 1e-7 and abs(up.max_speed - 4.0) < 1e-7


def test_unlimited_torque_compiles_to_an_unbounded_acceleration_budget():
    # compile.rs:1752-1772
    g = C.ConstructionGraph()
    base = spawn(g, (0, 2, 0))
    grounded(g, base)
    arm = spawn(g, (4, 2, 0))
    b = g.add_bearing(base, arm, (0.5, 0.5, 0.0), (1, 0, 0))
    g.wire(C.DriveWire(g.spawn(C.Controller(C.BuildPose((0, 40, 0)))), b))
    d = g.compile().coordinate_drives[0]
    assert math.isinf(d.max_acceleration) and math.isinf(d.min_angle) and d.min_angle < 0


def test_full_cylinder_segments_carry_ground_roles_and_radii():
    # device.rs:3015-3045: a full cylinder's sixteen segments get roles 1..16,
    # the segment-centre radius (local_center.w) and the outer radius (half_extents.w).
    g = C.ConstructionGraph()
    g.spawn(C.Cylinder(1.0, 0.0, 0.5, C.BuildPose((0, 4, 0))))
    g.spawn(C.Cylinder(1.0, 0.5, 2.0, C.BuildPose((12, 4, 0)), sweep_degrees=180))
    cols = g.compile().scene.collider_array()
    roles = cols[:, 15].view(np.uint32)
    assert list(roles[:16]) == list(range(1, 17)) and not roles[16:].any()
    assert np.allclose(cols[:16, 3], 0.25, atol=1e-6)          # (outer + inner) / 2
    assert np.allclose(cols[:16, 11], 0.25 + 0.25, atol=1e-6)  # + half radial thickness
    # Each segment's local x axis is its radial direction (compile.rs:905-915).
    angles = np.arctan2(cols[:16, 2], cols[:16, 0])
    assert np.allclose(np.sort(np.mod(angles, 2 * np.pi)), np.arange(16) * (2 * np.pi / 16), atol=1e-6)


def test_grid_rotation_matches_xyz_euler_quarter_turns():
    # geometry.rs:126-131: qx(a) * qy(b) * qz(c).
    q = C.GridRotation(1, 0, 0).quaternion()
    assert np.allclose(q, [math.sin(math.pi / 4), 0, 0, math.cos(math.pi / 4)])
    r = C._quat_matrix(C.GridRotation(1, 1, 0).quaternion())
    rx = C._quat_matrix(C.GridRotation(1, 0, 0).quaternion())
    ry = C._quat_matrix(C.GridRotation(0, 1, 0).quaternion())
    assert np.allclose(r, rx @ ry)
    assert np.allclose(C.BuildPose((1, 2, 3), half=(1, 0, 0)).translation(), [0.375, 0.5, 0.75])
