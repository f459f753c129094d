"""GPU tests of the articulated path against the oracle and closed forms:
hinge chains, a closed four-bar, drives and limits, a small articulated pile
with a dynamic root (bearing projection inside the contact solve), the row-count
errors of write_mechanism_drives / initialize_mechanism_coordinates, and the
100,000-body articulated scene."""
import math

import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd import articulated as A
from mechanic_amd.physics import GpuPhysics, GpuPhysicsConfig, GpuPhysicsError
from tests._oracle import Oracle
from tests.test_articulated_oracle import DT, _check_hinge_composition, axis_rot, quat_rot

pytestmark = pytest.mark.gpu

# FK evaluates sin/cos with the device's math library and the oracle with the
# host's, so articulated states agree to a tolerance rather than bit for bit.
POSE_TOL = 2e-5


# Config 2's closure solve: one PCG per closure component to a 1e-6 tolerance
# (at most 64 iterations), 24 Newton steps a tick, and the coordinate speeds
# projected onto the loops' velocity constraint before the tree reconstruction.
CONFIG2_CLOSURES = dict(closure_components=True, closure_pcg_tolerance=1e-6, closure_pcg_iterations=64,
                        closure_newton_steps=24, closure_velocity_projection=True)


def run_both(scene, ticks, collisions=False, friction=0.6, iterations=8, coordinates=None, self_collisions=True,
             closures=None):
    closures = closures or {}
    cfg = GpuPhysicsConfig(collisions_enabled=collisions, friction=friction, solver_iterations=iterations,
                           mechanism_self_collisions=self_collisions, **closures)
    gpu = GpuPhysics.new_with_config(scene, cfg)
    ora = Oracle(scene, iterations=iterations, friction=friction, collisions=collisions,
                 self_collisions=self_collisions, threads=4)
    if closures:
        ora.set_closure(closures.get("closure_newton_steps", 0), closures.get("closure_pcg_iterations", 0),
                        closures.get("closure_pcg_tolerance", 0.0), closures.get("closure_components", False),
                        closures.get("closure_velocity_projection", False))
    if coordinates is not None:
        gpu.initialize_mechanism_coordinates([mh._lib.MechanismCoordinate(float(a), float(v)) for a, v in coordinates])
        ora.set_coordinates(np.asarray(coordinates, np.float32))
    for t in range(1, ticks + 1):
        gpu.dispatch_tick(t)
        ora.step(t)
    gp, gv = gpu.read_bodies()
    gc = gpu.read_coordinates()
    flags = gpu.read_last_tick().error_flags
    gpu.close()
    op, ov = ora.bodies()
    return (gp, gv, gc, flags), (op, ov, ora.coordinates(), int(ora.diag()[0]))


def test_hinge_chain_one_tick_matches_oracle():
    scene = A.hinge_chain(links=6, length=0.4)
    (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, 1)
    assert gf == of == 0
    np.testing.assert_allclose(gp, op, rtol=0, atol=POSE_TOL)
    np.testing.assert_allclose(gv, ov, rtol=0, atol=POSE_TOL)
    np.testing.assert_allclose(gc, oc, rtol=0, atol=POSE_TOL)


def test_chain_longer_than_one_publication_group_matches_oracle():
    # 300 links make one tree of more than 256 bodies, which the per-group
    # publication cannot hold: the single-workgroup publication runs instead.
    scene = A.hinge_chain(links=300, length=0.05)
    (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, 2)
    assert gf == of == 0
    np.testing.assert_allclose(gp, op, rtol=0, atol=POSE_TOL)
    np.testing.assert_allclose(gv, ov, rtol=0, atol=POSE_TOL)
    np.testing.assert_allclose(gc, oc, rtol=0, atol=POSE_TOL)


def fan_of_chains(chains=12, links=6, length=0.3, height=3.0):
    """One static beam along z with `chains` hinge chains hanging off it, each
    `links` cuboids along +x, hinged about z, 0.4 m apart along the beam."""
    centers, half, static, specs = [(0.0, height, 0.2 * (chains - 1))], [(0.05, 0.05, 0.2 * chains)], [True], []
    for c in range(chains):
        z = 0.4 * c
        first = len(centers)
        for k in range(links):
            centers.append(((k + 0.5) * length, height, z))
            half.append((length * 0.5, 0.05, 0.05))
            static.append(False)
            specs.append(A.BearingSpec(0 if k == 0 else first + k - 1, first + k, (k * length, height, z),
                                       (0.0, 0.0, 1.0)))
    scene = A._cuboids(centers, half, static)
    return A.attach(scene, A.compile_mechanisms(np.asarray(centers, np.float32), np.asarray(static), specs))


def test_chains_sharing_a_static_body_across_bearing_groups_match_oracle():
    # 72 rows on one static beam: the bearing islands (one per chain, the beam
    # joins nothing) fill two groups of at most 51 rows, both holding the beam.
    scene = fan_of_chains()
    (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, 3)
    assert gf == of == 0
    np.testing.assert_allclose(gp, op, rtol=0, atol=POSE_TOL)
    np.testing.assert_allclose(gv, ov, rtol=0, atol=POSE_TOL)
    np.testing.assert_allclose(gc, oc, rtol=0, atol=POSE_TOL)
    assert np.all(gv[0] == 0.0)  # the beam never moves


def test_hinge_chain_swings_like_the_oracle():
    scene = A.hinge_chain(links=6, length=0.4)
    (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, 60)
    assert gf == of == 0
    np.testing.assert_allclose(gc[:, 0], oc[:, 0], atol=2e-4)
    np.testing.assert_allclose(gp[:, :3], op[:, :3], atol=5e-4)
    _check_hinge_composition(scene, gp, gc, POSE_TOL)


def test_four_bar_matches_oracle_and_stays_closed():
    scene = A.four_bar()
    start = [(0.2, 0.0), (-0.2, 0.0), (0.2, 0.0)]
    (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, 60, coordinates=start)
    np.testing.assert_allclose(gc[:, 0], oc[:, 0], atol=2e-4)
    np.testing.assert_allclose(gp[:, :3], op[:, :3], atol=5e-4)
    c = gc[:, 0]
    assert abs(c[0] - 0.2) > 0.02
    np.testing.assert_allclose(c, [c[0], -c[0], c[0]], atol=2e-3)


def test_driven_vertical_hinge_closed_form():
    drive = A.Drive(mode=1, max_acceleration=float("inf"), max_speed=1.0, target_speed=1.0)
    scene = A.hinge_chain(links=1, length=0.6, axis=(0.0, 1.0, 0.0), drive=drive)
    k = 45
    with GpuPhysics.new_with_config(scene, GpuPhysicsConfig(collisions_enabled=False)) as gpu:
        for t in range(1, k + 1):
            gpu.dispatch_tick(t)
        p, v = gpu.read_bodies()
        c = gpu.read_coordinates()
        assert gpu.read_last_tick().error_flags == 0
    assert c[0, 0] == pytest.approx(k * DT, rel=1e-5)
    anchor = np.array([0.0, 3.0, 0.0])
    expect = anchor + quat_rot(axis_rot((0, 1, 0), k * DT), np.array([0.3, 3.0, 0.0]) - anchor)
    np.testing.assert_allclose(p[1, :3], expect, atol=2e-6)
    np.testing.assert_allclose(v[1, 4:7], [0.0, 1.0, 0.0], atol=2e-6)


def test_drive_limits_and_angle_mode():
    drive = A.Drive(mode=1, max_acceleration=100.0, max_speed=2.0, target_speed=2.0, min_angle=-0.5, max_angle=0.5)
    scene = A.hinge_chain(links=1, length=0.6, axis=(0.0, 1.0, 0.0), drive=drive)
    with GpuPhysics.new_with_config(scene, GpuPhysicsConfig(collisions_enabled=False)) as gpu:
        for t in range(1, 40):
            gpu.dispatch_tick(t)
        c = gpu.read_coordinates()
        assert c[0, 0] == np.float32(0.5) and c[0, 1] == 0.0
        # Reprogram the drive to seek an angle (write_mechanism_drives).
        gpu.write_mechanism_drives([A.Drive(mode=2, max_acceleration=20.0, max_speed=2.0, target_angle=-0.3).row()])
        for t in range(40, 280):
            gpu.dispatch_tick(t)
        c = gpu.read_coordinates()
    assert abs(c[0, 0] + 0.3) < 2e-3


def test_row_count_errors():
    scene = A.hinge_chain(links=3)
    with GpuPhysics.new_with_config(scene, GpuPhysicsConfig(collisions_enabled=False)) as gpu:
        with pytest.raises(GpuPhysicsError) as e:
            gpu.write_mechanism_drives([A.PASSIVE.row()] * 2)
        assert e.value.kind == "DriveStateCount" and e.value.provided == 2 and e.value.required == 3
        with pytest.raises(GpuPhysicsError) as e:
            gpu.initialize_mechanism_coordinates([mh._lib.MechanismCoordinate(0.0, 0.0)] * 4)
        assert e.value.kind == "CoordinateStateCount" and e.value.required == 3


def small_articulated_pile():
    return A.articulated_pile(seed=5, nx=4, ny=2, nz=4, cylinders=150, tree_bearings=80, closures=6)


def test_small_articulated_pile_one_tick_matches_oracle():
    """A dynamic root exists (free cuboids), so a bearing projection follows
    every contact pass."""
    scene = small_articulated_pile()
    assert scene.bearing_count > 64
    (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, 1, collisions=True, self_collisions=False)
    assert gf == of
    np.testing.assert_allclose(gp, op, rtol=0, atol=POSE_TOL)
    np.testing.assert_allclose(gv, ov, rtol=0, atol=1e-4)
    np.testing.assert_allclose(gc, oc, rtol=0, atol=1e-4)


def test_small_articulated_pile_rollout_statistics():
    scene = small_articulated_pile()
    (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, 60, collisions=True, self_collisions=False)
    assert gf & mh.INVALID_NUMERIC_FLAG == 0 and of & mh.INVALID_NUMERIC_FLAG == 0
    # Trajectories drift apart at float rounding; compare distributions.
    assert abs(gp[:, 1].mean() - op[:, 1].mean()) < 0.02
    assert abs(np.abs(gv[:, :3]).mean() - np.abs(ov[:, :3]).mean()) < 0.05 * max(np.abs(ov[:, :3]).mean(), 0.1)


def test_reference_closure_solve_beyond_lds_and_one_product_slot():
    """The reference's global Newton step past its two fast paths: more than
    4,608 coordinates (joint frames and PCG vectors read from global memory,
    not LDS) and more than 512 closures (threads take a second closure-side
    product slot, reloading its end points and path ranges), against the
    oracle over the first ticks."""
    scene = A.articulated_pile(seed=9, nx=2, ny=2, nz=2, cylinders=6000, tree_bearings=5000, closures=600)
    assert scene.coordinate_count > 4608 and scene.bearing_count - scene.coordinate_count > 512
    for ticks in (1, 3):
        (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, ticks)
        print(f"[articulated] {scene.coordinate_count} coordinates, ticks {ticks}: flags {gf} / {of}", flush=True)
        assert gf == of
        np.testing.assert_allclose(gp, op, rtol=0, atol=POSE_TOL * ticks)
        np.testing.assert_allclose(gc, oc, rtol=0, atol=1e-4)


@pytest.mark.parametrize("case", ["nan_angle", "inf_angle", "scaled_root_rotation"])
def test_articulated_validation_raises_invalid_numeric(case):
    """validate_articulated_state (articulated.wgsl:458-484): a non-finite
    coordinate angle, or a body rotation whose squared length leaves
    (0.5, 1.5), raises INVALID_NUMERIC on the GPU exactly as on the oracle; the
    untouched scene raises nothing."""
    scene = A.hinge_chain(links=3, length=0.4)
    coords = [(0.1, 0.0)] * scene.coordinate_count
    if case == "nan_angle":
        coords[1] = (float("nan"), 0.0)
    elif case == "inf_angle":
        coords[2] = (float("inf"), 0.0)
    flags = []
    for poisoned in (False, True):
        cfg = GpuPhysicsConfig(collisions_enabled=False)
        with GpuPhysics.new_with_config(scene, cfg) as gpu:
            ora = Oracle(scene, collisions=False)
            rows = coords if poisoned else [(0.1, 0.0)] * scene.coordinate_count
            gpu.initialize_mechanism_coordinates([mh._lib.MechanismCoordinate(a, v) for a, v in rows])
            ora.set_coordinates(np.asarray(rows, np.float32))
            if poisoned and case == "scaled_root_rotation":
                p, v = gpu.read_bodies()
                p[0, 4:8] *= 2.0  # the static root: |q|^2 = 4
                gpu.write_bodies(p, v)
                ora.set_bodies(p, v)
            gpu.dispatch_tick(1)
            ora.step(1)
            flags.append((gpu.read_last_tick().error_flags, int(ora.diag()[0])))
    assert flags[0] == (0, 0)
    g, o = flags[1]
    assert g & mh.INVALID_NUMERIC_FLAG and g == o, flags


def test_graph_replay_matches_direct_dispatch():
    scene = small_articulated_pile()
    out = []
    for graph in (True, False):
        cfg = GpuPhysicsConfig(friction=0.6, use_graph=graph, mechanism_self_collisions=False)
        with GpuPhysics.new_with_config(scene, cfg) as gpu:
            for t in range(1, 21):
                gpu.dispatch_tick(t)
            p, v = gpu.read_bodies()
            out.append((p.view(np.uint32).copy(), v.view(np.uint32).copy(), gpu.read_coordinates().view(np.uint32)))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_articulated_scene_at_scale():
    """The 100,000-body articulated scene: 90,000 cuboids, 10,000 cylinders,
    5,000 bearings (500 closures), 20% of the bearings driven. With the
    reference's closure solve (one global PCG, 8 iterations, 12 Newton steps,
    the tree reconstructing velocities from unprojected speeds) the loops gain
    energy and open within the first 30 ticks, as in the FP64 oracle, and the
    non-convergence flag is raised; with config 2's solve (a PCG per component
    to 1e-6, at most 64 iterations, 24 Newton steps, speeds projected onto the
    loops' velocity constraint) every closure stays inside the 10 um / 0.001
    degree tolerance over 150 ticks and no flag is raised."""
    scene = A.articulated_pile()
    assert scene.body_count == 100_000 and scene.bearing_count == 5000
    for closures, expect_flag, ticks in (({}, True, 30), (CONFIG2_CLOSURES, False, 150)):
        cfg = GpuPhysicsConfig(friction=0.6, mechanism_self_collisions=False, **closures)
        with GpuPhysics.new_with_config(scene, cfg) as gpu:
            worst = 0.0
            flags = 0
            for t in range(1, ticks + 1):
                gpu.dispatch_tick(t)
                if t % 5 == 0:
                    r = gpu.read_last_tick()
                    worst = max(worst, r.anchor_residual_meters)
                    flags |= r.error_flags
            rb = gpu.read_last_tick()
            p, v = gpu.read_bodies()
            c = gpu.read_coordinates()
        print(f"[articulated 100k] closures {closures or 'reference'}: flags {flags}, "
              f"worst anchor residual {worst * 1e6:.1f} um over {ticks} ticks", flush=True)
        assert bool(flags & mh.CONSTRAINT_NON_CONVERGENCE_FLAG) == expect_flag
        if not expect_flag:
            assert flags == 0
            # The reported maximum covers the tree bearings too, whose FK-published
            # anchors round at FP32 steps of ~8 um this far (~100 m) from the origin.
            assert worst <= 50e-6
        assert np.all(np.isfinite(p)) and np.all(np.isfinite(c))
        assert rb.contact_count > 0
        _check_hinge_composition(scene, p, c, 5e-5)


@pytest.mark.parametrize("ticks", [1, 30])
def test_closure_components_match_oracle(ticks):
    """The per-component closure solve against the oracle's restatement of it,
    on the small articulated pile (six loops on zigzag chains, a dynamic root)."""
    scene = small_articulated_pile()
    (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, ticks, collisions=True, self_collisions=False,
                                                  closures=CONFIG2_CLOSURES)
    assert gf == of
    if ticks == 1:
        np.testing.assert_allclose(gp, op, rtol=0, atol=POSE_TOL)
        np.testing.assert_allclose(gc, oc, rtol=0, atol=1e-4)
    else:
        assert gf & mh.CONSTRAINT_NON_CONVERGENCE_FLAG == 0
        assert abs(gp[:, 1].mean() - op[:, 1].mean()) < 0.02


def test_four_bar_closure_components_match_oracle_and_close():
    """A closure ending on a root (the static base) in the per-component solve."""
    scene = A.four_bar()
    start = [(0.2, 0.0), (-0.2, 0.0), (0.2, 0.0)]
    (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, 60, coordinates=start, closures=CONFIG2_CLOSURES)
    np.testing.assert_allclose(gc[:, 0], oc[:, 0], atol=2e-4)
    np.testing.assert_allclose(gp[:, :3], op[:, :3], atol=5e-4)
    assert gf == 0


def test_closure_component_too_large_is_refused():
    """A component beyond the per-component solve's LDS (a 300-link chain
    closed back to its base) is refused at creation, and runs globally."""
    scene = A.hinge_chain(links=300, length=0.05)
    specs_scene = A.hinge_chain(links=300, length=0.05)
    chain = A.compile_mechanisms(specs_scene.pose_array()[:, :3], np.array([True] + [False] * 300),
                                 [A.BearingSpec(k, k + 1, (k * 0.05, 3.0, 0.0), (0.0, 0.0, 1.0)) for k in range(300)]
                                 + [A.BearingSpec(300, 0, (300 * 0.05 - 0.025, 3.0, 0.0), (0.0, 0.0, 1.0))])
    scene = A.attach(specs_scene, chain)
    with pytest.raises(Exception) as e:
        GpuPhysics.new_with_config(scene, GpuPhysicsConfig(collisions_enabled=False, **CONFIG2_CLOSURES))
    assert "closure component" in str(e.value) or "unsupported" in str(e.value)
    with GpuPhysics.new_with_config(scene, GpuPhysicsConfig(collisions_enabled=False)) as gpu:
        gpu.dispatch_tick(1)
        assert gpu.read_last_tick().error_flags & mh.INVALID_NUMERIC_FLAG == 0


def test_velocity_projection_moves_the_closure_anchors_together():
    """The velocity-level closure on the GPU: on an uneven four-bar the closure
    bearing's anchor points move together (FP32), and the state follows the
    oracle's."""
    from tests.test_articulated_oracle import _closure_anchor_speed, _uneven_four_bar
    scene = _uneven_four_bar()
    start = [(0.4, 0.0), (-0.4, 0.0), (0.4, 0.0)]
    cfg = GpuPhysicsConfig(collisions_enabled=False, **CONFIG2_CLOSURES)
    worst = 0.0
    with GpuPhysics.new_with_config(scene, cfg) as gpu:
        gpu.initialize_mechanism_coordinates([mh._lib.MechanismCoordinate(a, v) for a, v in start])
        for t in range(1, 61):
            gpu.dispatch_tick(t)
            p, v = gpu.read_bodies()
            worst = max(worst, _closure_anchor_speed(scene, p.astype(np.float64), v.astype(np.float64)))
        assert gpu.read_last_tick().error_flags == 0
    assert worst < 1e-4, worst
    (gp, gv, gc, gf), (op, ov, oc, of) = run_both(scene, 60, coordinates=start, closures=CONFIG2_CLOSURES)
    np.testing.assert_allclose(gc[:, 0], oc[:, 0], atol=2e-4)
    np.testing.assert_allclose(gp[:, :3], op[:, :3], atol=5e-4)
