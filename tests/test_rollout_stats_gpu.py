"""Rollout statistics of the HIP step (FP32) against the FP64 oracle.

The configuration-0 drop (1,000 cuboids of 0.25 m in a 10x10x10 lattice, 0.27 m
spacing, 0.5 m above the ground, friction 0.6, 8 iterations) is stepped 300
ticks on the GPU and by the oracle in FP64. A falling pile is chaotic: the FP32
and FP64 trajectories part at rounding level within a few ticks, so what is
compared is per-tick statistics -- centre-of-mass height, kinetic energy,
contact count, mean and maximum penetration. (Against the FP32 oracle the same
run is bit-exact; test_parity_gpu checks that.)

With uniform yaw, 0.25 m cubes on a 0.27 m lattice overlap their neighbours by up
to 8 cm at tick 1 (a yawed cube spans up to 0.35 m), so the pile pushes apart
as it falls and the ten-high Jacobi stack then sinks into the ground plane; the
FP64 oracle shows the same, and the bounds below are on the agreement.

The four-bar's loop closure is compared the same way: the closure anchor
residual after each of 300 ticks, GPU (the validation pass's maximum) against
the FP64 oracle's final links.

The articulated configuration's closure solve (a PCG per closure component to
a 1e-6 tolerance, at most 64 iterations, 24 Newton steps, speeds projected on
the loops) is rolled out 300 ticks on a reduced articulated scene built by the
same rules (cylinder chains of 2-20 links with loops, 20% of the bearings
driven at +-2 rad/s within +-pi/2, cuboids beside them): the same statistics
plus the closure residuals and the FP64 oracle's PCG iterations per tick.
MH_ROLLOUT_FULL=1 runs it on the 100,000-body scene instead (the FP64 oracle
then takes several minutes)."""
import os

import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd import _lib
from mechanic_amd import articulated as A
from tests._oracle import Oracle

pytestmark = pytest.mark.gpu

TICKS = 300


def quat_matrix(q):
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], axis=1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], axis=1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1)], axis=1)


def tick_stats(scene, poses, velocities, penetrations):
    rows = np.ctypeslib.as_array(scene.inertias).view(np.float32).reshape(-1, 16).astype(np.float64)
    m = rows[:, 0]
    inertia = np.stack([rows[:, 4:7], rows[:, 8:11], rows[:, 12:15]], axis=1)
    p = np.asarray(poses, np.float64)
    v = np.asarray(velocities, np.float64)
    # Angular velocity in the body frame: R^T w.
    wb = np.einsum("nji,nj->ni", quat_matrix(p[:, 4:8]), v[:, 4:7])
    kinetic = 0.5 * (m * (v[:, :3] ** 2).sum(1)).sum() + 0.5 * np.einsum("ni,nij,nj->n", wb, inertia, wb).sum()
    pen = np.maximum(np.asarray(penetrations, np.float64), 0.0)
    return {"com_y": (m * p[:, 1]).sum() / m.sum(), "kinetic": kinetic, "contacts": len(pen),
            "pen_mean": pen.mean() if len(pen) else 0.0, "pen_max": pen.max() if len(pen) else 0.0}


def rollout_config0():
    scene = mh.named_pile("pile_1k")
    cfg = mh.GpuPhysicsConfig(friction=0.6, solver_iterations=8)
    gpu = mh.GpuPhysics.new_with_config(scene, cfg)
    ora = Oracle(scene, iterations=8, friction=0.6, fp64=True, threads=4)
    g_stats, o_stats = [], []
    for t in range(1, TICKS + 1):
        gpu.dispatch_tick(t)
        ora.step(t)
        n = gpu.read_last_tick().contact_count
        rows = np.frombuffer(gpu.read_buffer(_lib.BUFFER_CONTACTS, n * 64), np.float32).reshape(-1, 16)
        gp, gv = gpu.read_bodies()
        g_stats.append(tick_stats(scene, gp, gv, rows[:, 7]))
        op, ov = ora.bodies_f64()
        o_stats.append(tick_stats(scene, op, ov, ora.contacts()[:, 7]))
    flags = gpu.read_last_tick().error_flags
    gpu.close()
    return g_stats, o_stats, flags


def window(stats, key, lo, hi):
    return np.array([s[key] for s in stats[lo:hi]])


def test_config0_rollout_statistics_match_fp64_oracle():
    g, o, flags = rollout_config0()
    assert flags == 0
    # The drop: both fall, land and come to rest at the same height.
    for k in (1, 30, 60, 150, TICKS):
        gs, os_ = g[k - 1], o[k - 1]
        print(f"tick {k:3d}: com_y gpu {gs['com_y']:.4f} fp64 {os_['com_y']:.4f} | KE gpu {gs['kinetic']:.3f} "
              f"fp64 {os_['kinetic']:.3f} J | contacts gpu {gs['contacts']} fp64 {os_['contacts']} | "
              f"pen mean gpu {1e3 * gs['pen_mean']:.3f} fp64 {1e3 * os_['pen_mean']:.3f} mm, max gpu "
              f"{1e3 * gs['pen_max']:.3f} fp64 {1e3 * os_['pen_max']:.3f} mm")
    # The first ticks part only at rounding level.
    assert abs(g[9]["com_y"] - o[9]["com_y"]) < 1e-4
    assert abs(g[9]["kinetic"] - o[9]["kinetic"]) < 1e-3 * o[9]["kinetic"]
    assert g[9]["contacts"] == o[9]["contacts"]
    # The last 100 ticks, as means: height within 2 cm, kinetic energy within 30%,
    # contact count within 5%, mean penetration within 10%, maximum within 20%.
    lo, hi = TICKS - 100, TICKS
    mean = lambda s, k: window(s, k, lo, hi).mean()  # noqa: E731
    assert abs(mean(g, "com_y") - mean(o, "com_y")) < 0.02
    assert abs(mean(g, "kinetic") - mean(o, "kinetic")) < 0.3 * mean(o, "kinetic")
    assert abs(mean(g, "contacts") - mean(o, "contacts")) < 0.05 * mean(o, "contacts")
    assert abs(mean(g, "pen_mean") - mean(o, "pen_mean")) < 0.1 * mean(o, "pen_mean")
    assert abs(mean(g, "pen_max") - mean(o, "pen_max")) < 0.2 * mean(o, "pen_max")


def test_four_bar_closure_residual_matches_fp64_oracle():
    scene = A.four_bar()
    start = [(0.2, 0.0), (-0.2, 0.0), (0.2, 0.0)]
    cfg = mh.GpuPhysicsConfig(collisions_enabled=False)
    gpu = mh.GpuPhysics.new_with_config(scene, cfg)
    gpu.initialize_mechanism_coordinates([_lib.MechanismCoordinate(float(a), float(b)) for a, b in start])
    ora = Oracle(scene, collisions=False, fp64=True)
    ora.set_coordinates(np.asarray(start, np.float32))
    g_res, o_res = [], []
    for t in range(1, TICKS + 1):
        gpu.dispatch_tick(t)
        ora.step(t)
        g_res.append(1e6 * gpu.read_last_tick().anchor_residual_meters)
        o_res.append(1e6 * np.linalg.norm(ora.closure_residuals()[:, :3], axis=1).max())
    flags = gpu.read_last_tick().error_flags
    gpu.close()
    g_res, o_res = np.array(g_res, np.float64), np.array(o_res)
    print(f"closure anchor residual (um): gpu mean {g_res.mean():.3f} max {g_res.max():.3f}; "
          f"fp64 mean {o_res.mean():.3f} max {o_res.max():.3f}")
    # Below the reference's 10 um tolerance throughout, in both.
    assert flags & mh.CONSTRAINT_NON_CONVERGENCE_FLAG == 0
    assert g_res.max() < 10.0 and o_res.max() < 10.0


# The articulated configuration's closure solve.
CONFIG2_CLOSURES = dict(closure_components=True, closure_pcg_tolerance=1e-6, closure_pcg_iterations=64,
                        closure_newton_steps=24, closure_velocity_projection=True)


def test_articulated_rollout_statistics_match_fp64_oracle():
    full = os.environ.get("MH_ROLLOUT_FULL") == "1"
    scene = A.articulated_pile() if full else \
        A.articulated_pile(seed=2, nx=12, ny=8, nz=12, cylinders=2000, tree_bearings=900, closures=100)
    cfg = mh.GpuPhysicsConfig(friction=0.6, solver_iterations=8, mechanism_self_collisions=False, **CONFIG2_CLOSURES)
    gpu = mh.GpuPhysics.new_with_config(scene, cfg)
    ora = Oracle(scene, iterations=8, friction=0.6, self_collisions=False, fp64=True, threads=16 if full else 8)
    c = CONFIG2_CLOSURES
    pcg = ora.set_closure(c["closure_newton_steps"], c["closure_pcg_iterations"], c["closure_pcg_tolerance"],
                          True, True)
    g_stats, o_stats, g_anchor, o_anchor, o_closure, o_pcg = [], [], [], [], [], []
    flags = 0
    for t in range(1, TICKS + 1):
        gpu.dispatch_tick(t)
        ora.step(t)
        r = gpu.read_last_tick()
        flags |= r.error_flags
        rows = np.frombuffer(gpu.read_buffer(_lib.BUFFER_CONTACTS, r.contact_count * 64), np.float32).reshape(-1, 16)
        gp, gv = gpu.read_bodies()
        g_stats.append(tick_stats(scene, gp, gv, rows[:, 7]))
        op, ov = ora.bodies_f64()
        o_stats.append(tick_stats(scene, op, ov, ora.contacts()[:, 7]))
        g_anchor.append(1e6 * r.anchor_residual_meters)
        od = ora.diag()
        flags_o = int(od[0])
        o_anchor.append(float(od[3]))
        o_closure.append(1e6 * np.linalg.norm(ora.closure_residuals()[:, :3], axis=1).max())
        total = ora.set_closure(c["closure_newton_steps"], c["closure_pcg_iterations"], c["closure_pcg_tolerance"],
                                True, True)
        o_pcg.append(total - pcg)
        pcg = total
        if full and t % 10 == 0:
            print(f"[articulated rollout] tick {t}", flush=True)
    gpu.close()
    print(f"[articulated rollout] {scene.body_count} bodies, {scene.bearing_count} bearings "
          f"({scene.bearing_count - scene.coordinate_count} closures)")
    for k in (1, 30, 100, 200, TICKS):
        gs, os_ = g_stats[k - 1], o_stats[k - 1]
        print(f"tick {k:3d}: com_y gpu {gs['com_y']:.4f} fp64 {os_['com_y']:.4f} | KE gpu {gs['kinetic']:.4g} "
              f"fp64 {os_['kinetic']:.4g} J | contacts gpu {gs['contacts']} fp64 {os_['contacts']} | pen mean gpu "
              f"{1e3 * gs['pen_mean']:.2f} fp64 {1e3 * os_['pen_mean']:.2f} mm, max gpu {1e3 * gs['pen_max']:.1f} "
              f"fp64 {1e3 * os_['pen_max']:.1f} mm | max anchor residual gpu {g_anchor[k - 1]:.0f} fp64 "
              f"{o_anchor[k - 1]:.0f} um (fp64 closures {o_closure[k - 1]:.3f} um) | fp64 PCG iterations "
              f"{o_pcg[k - 1]}", flush=True)
    g_anchor, o_closure, o_pcg = np.array(g_anchor), np.array(o_closure), np.array(o_pcg)
    print(f"closure residual over {TICKS} ticks: gpu (all bearings, whole um) mean {g_anchor.mean():.2f} max "
          f"{g_anchor.max():.0f}; fp64 closures mean {o_closure.mean():.3f} max {o_closure.max():.3f} um; "
          f"fp64 PCG iterations per tick mean {o_pcg.mean():.0f} max {o_pcg.max()}", flush=True)
    assert flags == 0 and flags_o == 0
    # The closures stay inside the reference's 10 um tolerance (the GPU figure
    # also covers tree bearings far from the origin, rounded to whole um).
    assert o_closure.max() <= 10.5 and g_anchor.max() <= 50.0
    # The first tick parts only at rounding level (a contact at the emission
    # threshold can fall either way in FP32 and FP64).
    assert abs(g_stats[0]["com_y"] - o_stats[0]["com_y"]) < 1e-4
    assert abs(g_stats[0]["contacts"] - o_stats[0]["contacts"]) <= 1e-4 * o_stats[0]["contacts"]
    # The last 100 ticks, as means.
    lo, hi = TICKS - 100, TICKS
    mean = lambda s, k: window(s, k, lo, hi).mean()  # noqa: E731
    assert abs(mean(g_stats, "com_y") - mean(o_stats, "com_y")) < 0.05 * abs(mean(o_stats, "com_y"))
    assert abs(mean(g_stats, "kinetic") - mean(o_stats, "kinetic")) < 0.3 * mean(o_stats, "kinetic")
    assert abs(mean(g_stats, "contacts") - mean(o_stats, "contacts")) < 0.05 * mean(o_stats, "contacts")
    assert abs(mean(g_stats, "pen_mean") - mean(o_stats, "pen_mean")) < 0.2 * mean(o_stats, "pen_mean")
    assert abs(mean(g_stats, "pen_max") - mean(o_stats, "pen_max")) < 0.3 * mean(o_stats, "pen_max")
