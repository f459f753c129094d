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
the FP64 oracle's final links."""
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
