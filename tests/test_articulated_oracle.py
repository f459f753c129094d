"""Pins the oracle's articulated path (and the host-side lowering) by means
independent of its code: a dense FP64 least-squares solve of the closure Newton
step, the linear contraction of the damped Newton loop, the closed-form hinge
composition of every published pose, a driven hinge about a vertical axis whose
angle is known in closed form, angle limits, and the spanning-forest rules of
the compile step (crates/mechanic-core/src/compile.rs:418-585)."""
import math

import numpy as np
import pytest

from mechanic_amd import articulated as A
from tests._oracle import Oracle

DT = 1.0 / 60.0


def quat_mul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + bw * ax + (ay * bz - az * by), aw * by + bw * ay + (az * bx - ax * bz),
                     aw * bz + bw * az + (ax * by - ay * bx), aw * bw - (ax * bx + ay * by + az * bz)])


def quat_rot(q, v):
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return R @ np.asarray(v, np.float64)


def axis_rot(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    return np.concatenate([a * math.sin(angle / 2), [math.cos(angle / 2)]])


def rows(scene):
    br = np.ctypeslib.as_array(scene.bearings).view(np.float32).reshape(-1, 20)
    mb = np.ctypeslib.as_array(scene.mechanism_bodies).view(np.float32).reshape(-1, 16)
    return br, br.view(np.uint32), mb, mb.view(np.uint32)


def test_compile_spanning_forest_rules():
    """Driven bearings join the forest first; an edge inside a tree becomes a
    closure; traversal is breadth-first from the static body; every bearing
    suppresses its pair."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [1, 1, 0]], np.float32)
    static = np.array([True, False, False, False])
    z = (0, 0, 1)
    drive = A.Drive(mode=1, max_acceleration=10, max_speed=1, target_speed=1)
    specs = [A.BearingSpec(0, 1, (0.5, 0, 0), z), A.BearingSpec(1, 2, (1.5, 0, 0), z),
             A.BearingSpec(0, 3, (0.5, 0.5, 0), z), A.BearingSpec(3, 1, (1, 0.5, 0), z, drive)]
    m = A.compile_mechanisms(pos, static, specs)
    # Rows 3 (driven) first, then 0 and 1; row 2 would close the loop 0-1-3.
    assert m.tree_rows == [0, 1, 3]
    assert m.closure_count == 1
    br = np.ctypeslib.as_array(m.bearings).view(np.uint32).reshape(-1, 20)
    assert list(br[:, 18]) == [0, 1, 0xFFFFFFFF, 2]  # coordinates in row order, closure none
    assert list(br[:, 19]) == [0, 0, 1, 0]
    assert list(m.preorder) == [0, 1, 2, 3]
    assert list(m.parents) == [0, 0, 1, 1]
    assert list(m.depth) == [0, 1, 2, 2]
    assert m.suppressed_pairs == [(0, 1), (0, 3), (1, 2), (1, 3)]
    mb = np.ctypeslib.as_array(m.mechanism_bodies).view(np.uint32).reshape(-1, 16)
    assert list(mb[3, :4]) == [1, 3, 1, 0]  # parent 1 through row 3 traversed B -> A
    np.testing.assert_allclose(np.ctypeslib.as_array(m.mechanism_bodies).view(np.float32).reshape(-1, 16)[3, 8:11],
                               pos[3] - pos[1])
    with pytest.raises(ValueError):
        A.compile_mechanisms(pos, np.array([True, True, False, False]),
                             [A.BearingSpec(0, 1, (0.5, 0, 0), z, drive)])


def _fd_jacobian(o, theta, h=2.0 ** -16):
    """Central differences; h and theta are exact in float32 (the coordinate rows' type)."""
    q = len(theta)
    base = np.zeros((q, 2))
    cols = []
    for i in range(q):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        base[:, 0] = tp
        o.set_coordinates(base)
        o.fk_newton(0)
        rp = o.closure_residuals().ravel()
        base[:, 0] = tm
        o.set_coordinates(base)
        o.fk_newton(0)
        rm = o.closure_residuals().ravel()
        cols.append((rp - rm) / (2 * h))
    return np.stack(cols, axis=1)


def test_closure_newton_step_matches_dense_solve():
    """One Newton step moves the coordinates by 0.35 x, where x is the
    Jacobi-preconditioned minimum-norm solution of J^T J x = -J^T r (what PCG
    converges to on the four-bar's rank-deficient normal equations)."""
    scene = A.four_bar()
    o = Oracle(scene, fp64=True, collisions=False)
    # Coordinates are float32 in the ABI; pick values exact in float32.
    theta = np.array([0.125, -0.0625, 0.03125])
    J = _fd_jacobian(o, theta)
    o.set_coordinates(np.stack([theta, np.zeros(3)], 1))
    o.fk_newton(0)
    r = o.closure_residuals().ravel()
    o.set_coordinates(np.stack([theta, np.zeros(3)], 1))
    o.fk_newton(1)
    got = o.coordinates()[:, 0].astype(np.float64) - theta
    Amat = J.T @ J
    g = J.T @ r
    M = np.diag(Amat)
    s = 1.0 / np.sqrt(M)
    x = s * (np.linalg.pinv((Amat * s[:, None]) * s[None, :]) @ (-(g * s)))
    expect = np.clip(0.35 * x, -0.2, 0.2)
    np.testing.assert_allclose(got, expect, rtol=2e-5, atol=1e-7)
    assert np.abs(x).max() > 1e-3  # a real correction, not a no-op


def test_newton_loop_contracts_linearly():
    """12 damped Newton steps on a small violation shrink the closure residual
    by (1 - 0.35)^12 in the linear regime."""
    o = Oracle(A.four_bar(), fp64=True, collisions=False)
    o.set_coordinates(np.array([[0.01, 0.0], [-0.004, 0.0], [0.002, 0.0]]))
    o.fk_newton(0)
    r0 = np.linalg.norm(o.closure_residuals())
    o.set_coordinates(np.array([[0.01, 0.0], [-0.004, 0.0], [0.002, 0.0]]))
    o.fk_newton(12)
    r12 = np.linalg.norm(o.closure_residuals())
    assert r12 / r0 == pytest.approx(0.65 ** 12, rel=0.02)


def _check_hinge_composition(scene, poses, coords, tol):
    """Every non-root body's pose equals its parent's composed with the hinge
    rotation about the joint axis through the joint anchor (bind rotation
    identity)."""
    br, bru, mb, mbu = rows(scene)
    for b in range(scene.body_count):
        if mbu[b, 3] != 0:
            continue
        p, row, direction = mbu[b, 0], mbu[b, 1], mbu[b, 2]
        anchor_a, anchor_b, axis = br[row, 0:3], br[row, 4:7], br[row, 8:11]
        angle = float(coords[bru[row, 18], 0])
        assert direction == 0
        rel = axis_rot(axis, angle)
        rp = poses[p, 4:8].astype(np.float64)
        expect_rot = quat_mul(rp, rel)
        expect_rot /= np.linalg.norm(expect_rot)
        expect_pos = poses[p, :3] + quat_rot(rp, anchor_a - quat_rot(rel, anchor_b))
        got_rot = poses[b, 4:8]
        if np.dot(got_rot, expect_rot) < 0:
            expect_rot = -expect_rot
        np.testing.assert_allclose(poses[b, :3], expect_pos, atol=tol)
        np.testing.assert_allclose(got_rot, expect_rot, atol=tol)
        # The joint anchors coincide in the world.
        wa = poses[p, :3] + quat_rot(rp, anchor_a)
        wb = poses[b, :3] + quat_rot(poses[b, 4:8], anchor_b)
        np.testing.assert_allclose(wa, wb, atol=tol)


def test_published_poses_compose_hinges():
    scene = A.hinge_chain(links=6, length=0.4)
    o = Oracle(scene, iterations=8, collisions=False)
    for t in range(1, 41):
        o.step(t)
    poses, _ = o.bodies()
    coords = o.coordinates()
    assert np.abs(coords[:, 0]).max() > 0.05  # the chain has swung
    _check_hinge_composition(scene, poses, coords, 2e-5)
    assert o.diag()[0] == 0


def test_driven_vertical_hinge_closed_form():
    """A link hinged about the vertical axis and driven at 1 rad/s with an
    unlimited torque budget: gravity exerts no torque about the axis, so the
    angle after k ticks is k/60 rad and the pose is the bind pose turned about
    the anchor."""
    drive = A.Drive(mode=1, max_acceleration=float("inf"), max_speed=1.0, target_speed=1.0)
    scene = A.hinge_chain(links=1, length=0.6, axis=(0.0, 1.0, 0.0), drive=drive)
    o = Oracle(scene, iterations=8, collisions=False, fp64=True)
    k = 45
    for t in range(1, k + 1):
        o.step(t)
    coords = o.coordinates()
    assert coords[0, 0] == pytest.approx(k * DT, rel=1e-6)
    assert coords[0, 1] == pytest.approx(1.0, rel=1e-6)
    p, v = o.bodies_f64()
    anchor = np.array([0.0, 3.0, 0.0])
    bind = np.array([0.3, 3.0, 0.0])
    expect = anchor + quat_rot(axis_rot((0, 1, 0), k * DT), bind - anchor)
    np.testing.assert_allclose(p[1, :3], expect, atol=1e-9)
    # Reconstructed velocity: omega x r about the fixed anchor.
    np.testing.assert_allclose(v[1, 4:7], [0.0, 1.0, 0.0], atol=1e-9)
    np.testing.assert_allclose(v[1, :3], np.cross([0.0, 1.0, 0.0], expect - anchor), atol=1e-9)


def test_drive_stops_at_the_angle_limit():
    drive = A.Drive(mode=1, max_acceleration=100.0, max_speed=2.0, target_speed=2.0, min_angle=-0.5, max_angle=0.5)
    scene = A.hinge_chain(links=1, length=0.6, axis=(0.0, 1.0, 0.0), drive=drive)
    o = Oracle(scene, iterations=8, collisions=False)
    for t in range(1, 40):
        o.step(t)
    c = o.coordinates()
    assert c[0, 0] == np.float32(0.5) and c[0, 1] == 0.0


def test_angle_drive_settles_on_target():
    drive = A.Drive(mode=2, max_acceleration=20.0, max_speed=2.0, target_angle=0.7)
    scene = A.hinge_chain(links=1, length=0.6, axis=(0.0, 1.0, 0.0), drive=drive)
    o = Oracle(scene, iterations=8, collisions=False)
    for t in range(1, 241):
        o.step(t)
    c = o.coordinates()
    assert abs(c[0, 0] - 0.7) < 2e-3 and abs(c[0, 1]) < 0.05


def test_pendulum_settles_under_its_hinge():
    scene = A.hinge_chain(links=1, length=0.6)
    o = Oracle(scene, iterations=8, collisions=False)
    for t in range(1, 1201):
        o.step(t)
    p, v = o.bodies()
    assert abs(p[1, 0] - 0.0) < 5e-3 and p[1, 1] < 3.0 - 0.29
    assert np.abs(v[1]).max() < 1e-2
    assert o.diag()[0] == 0


def test_four_bar_keeps_its_loop_closed():
    """Started tilted (the parallelogram keeps its loop closed under the joint
    angles (t, -t, t)), the four-bar falls under gravity with its closure held."""
    o = Oracle(A.four_bar(), iterations=8, collisions=False)
    o.set_coordinates(np.array([[0.2, 0.0], [-0.2, 0.0], [0.2, 0.0]]))
    for t in range(1, 121):
        o.step(t)
        d = o.diag()
        # Anchor residual in micrometres stays within a small band.
        assert d[3] < 2000, (t, d)
    c = o.coordinates()[:, 0]
    assert abs(c[0] - 0.2) > 0.05  # it fell
    np.testing.assert_allclose(c, [c[0], -c[0], c[0]], atol=2e-3)  # still a parallelogram


def _uneven_four_bar():
    """A four-bar whose links differ in mass (12, 216 and 2.4 kg of cuboid)."""
    h = 3.0
    centers = [(0.5, h, 0.0), (0.0, h + 0.5, 0.0), (0.5, h + 1.0, 0.0), (1.0, h + 0.5, 0.0)]
    half = [(0.5, 0.05, 0.05), (0.15, 0.5, 0.15), (0.5, 0.3, 0.3), (0.05, 0.5, 0.05)]
    scene = A._cuboids(centers, half, [True, False, False, False])
    z = (0.0, 0.0, 1.0)
    specs = [A.BearingSpec(0, 1, (0.0, h, 0.0), z), A.BearingSpec(1, 2, (0.0, h + 1.0, 0.0), z),
             A.BearingSpec(2, 3, (1.0, h + 1.0, 0.0), z), A.BearingSpec(3, 0, (1.0, h, 0.0), z)]
    return A.attach(scene, A.compile_mechanisms(np.asarray(centers, np.float32), np.array([True, False, False, False]),
                                                specs))


def _closure_anchor_speed(scene, poses, velocities):
    """|v_a - v_b| of the closure bearing's two anchor points."""
    row = next(scene.bearings[i] for i in range(scene.bearing_count) if scene.bearings[i].metadata[3])
    out = []
    for body, local in ((row.metadata[0], row.local_anchor_a), (row.metadata[1], row.local_anchor_b)):
        x, y, z, w = poses[body, 4:8]
        u = np.array([x, y, z])
        v = np.array(list(local)[:3])
        t = 2.0 * np.cross(u, v)
        arm = v + w * t + np.cross(u, t)
        out.append(velocities[body, :3] + np.cross(velocities[body, 4:7], arm))
    return float(np.linalg.norm(out[0] - out[1]))


def test_velocity_projection_moves_the_closure_anchors_together():
    """The reference reconstructs body velocities along the spanning tree from
    coordinate speeds the loop does not constrain (articulated.wgsl:423-456),
    so the closure bearing's two anchor points separate at up to ~0.5 m/s on an
    uneven four-bar; with the speeds projected onto J v = 0 first they move
    together to rounding (FP64 oracle, 60 ticks)."""
    scene = _uneven_four_bar()
    worst = {}
    for projection in (False, True):
        o = Oracle(scene, iterations=8, collisions=False, fp64=True)
        o.set_closure(24, 64, 1e-6, True, projection)
        o.set_coordinates(np.array([[0.4, 0.0], [-0.4, 0.0], [0.4, 0.0]], np.float32))
        w = 0.0
        for t in range(1, 61):
            o.step(t)
            p, v = o.bodies_f64()
            w = max(w, _closure_anchor_speed(scene, p, v))
        worst[projection] = w
        assert o.diag()[0] == 0
    assert worst[True] < 1e-9, worst
    assert worst[False] > 0.1, worst


def test_loops_stay_closed_only_with_the_velocity_projection():
    """A reduced articulated scene (1,000 cuboids, 1,000 cylinders, 450 tree
    bearings, 50 four-bar loops, 20% of the bearings driven), 240 ticks in the
    FP32 oracle: the reconstructed loop-violating velocities pump energy into
    the loops, which open (the non-convergence flag) under the reference's
    closure solve from the first ticks and under the per-component solve
    without the velocity projection within 240 ticks; with it no flag is
    raised and every closure stays inside the 10 um tolerance."""
    scene = A.articulated_pile(nx=10, ny=10, nz=10, cylinders=1000, tree_bearings=450, closures=50)
    assert scene.bearing_count - scene.coordinate_count == 50
    first_flag = {}
    for mode, setting in (("reference", None), ("components", False), ("projected", True)):
        o = Oracle(scene, friction=0.6, self_collisions=False, threads=4)
        if setting is not None:
            o.set_closure(24, 64, 1e-6, True, setting)
        first_flag[mode] = None
        for t in range(1, 241):
            o.step(t)
            if o.diag()[0] and first_flag[mode] is None:
                first_flag[mode] = t
                if mode != "projected":
                    break
    assert first_flag["projected"] is None, first_flag
    assert first_flag["reference"] is not None and first_flag["components"] is not None, first_flag


@pytest.mark.parametrize("case", ["nan_angle", "inf_angle", "scaled_root_rotation", "clean"])
def test_validation_flags_non_finite_state(case):
    """validate_articulated_state (articulated.wgsl:458-484) in the oracle: a NaN
    or infinite angle or a root rotation with |q|^2 = 4 raises
    INVALID_NUMERIC (2) after one tick; a clean chain raises nothing."""
    import mechanic_amd as mh
    scene = A.hinge_chain(links=3, length=0.4)
    coords = np.full((scene.coordinate_count, 2), [0.1, 0.0], np.float32)
    if case == "nan_angle":
        coords[1, 0] = np.nan
    elif case == "inf_angle":
        coords[2, 0] = np.inf
    o = Oracle(scene, collisions=False)
    o.set_coordinates(coords)
    if case == "scaled_root_rotation":
        p, v = o.bodies()
        p[0, 4:8] *= 2.0
        o.set_bodies(p, v)
    o.step(1)
    flags = int(o.diag()[0])
    assert bool(flags & mh.INVALID_NUMERIC_FLAG) == (case != "clean"), flags
