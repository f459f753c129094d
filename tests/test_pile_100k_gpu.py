"""The 100,000-cuboid pile at full size: the settled state the benchmark times,
compared with the oracle bit for bit, and the learned warm start on it
(16 host threads run the oracle)."""
import os

import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd import learned as L
from mechanic_amd.physics import GpuPhysicsConfig
from tests._oracle import Oracle

pytestmark = pytest.mark.gpu

FRICTION, ITERATIONS = 0.6, 8


def host_threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 1


@pytest.fixture(scope="module")
def settled():
    """The pile after 300 ticks from the drop (friction 0.6, 8 passes)."""
    scene = mh.named_pile("pile_100k")
    cfg = GpuPhysicsConfig(friction=FRICTION, solver_iterations=ITERATIONS)
    with mh.GpuPhysics.new_with_config(scene, cfg) as gpu:
        gpu.dispatch_ticks(1, 300)
        poses, velocities = gpu.read_bodies()
        r = gpu.read_last_tick()
    print(f"[settled] tick 300: pairs {r.pair_count} contacts {r.contact_count} active {r.active_contact_count}",
          flush=True)
    return scene, poses, velocities


def test_settled_pile_100k_bit_exact(settled):
    """Ticks 301..303 from the settled state, restored into a fresh handle and
    into the oracle (both with an empty manifold cache, which the two later
    ticks fill and use): bodies, pair, contact and active-contact counts and
    flags bit for bit, at the scale the benchmark times (over a million pairs,
    the slot-overflow path, every cache rule)."""
    scene, poses, velocities = settled
    cfg = GpuPhysicsConfig(friction=FRICTION, solver_iterations=ITERATIONS)
    gpu = mh.GpuPhysics.new_with_config(scene, cfg)
    gpu.write_bodies(poses, velocities)
    ora = Oracle(scene, iterations=ITERATIONS, friction=FRICTION, threads=host_threads())
    ora.set_bodies(poses, velocities)
    for tick in (301, 302, 303):
        gpu.dispatch_tick(tick)
        ora.step(tick)
        r = gpu.read_last_tick()
        d = ora.diag()
        print(f"[settled] tick {tick}: pairs {r.pair_count} contacts {r.contact_count} "
              f"active {r.active_contact_count} flags {r.error_flags}", flush=True)
        assert (r.error_flags, r.pair_count, r.contact_count, r.active_contact_count) == \
            (int(d[0]), int(d[1]), int(d[2]), int(d[5])), tick
        gp, gv = gpu.read_bodies()
        op, ov = ora.bodies()
        assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), f"poses differ at tick {tick}"
        assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32)), f"velocities differ at tick {tick}"
    assert r.pair_count > 1_000_000 and r.active_contact_count > 600_000
    gpu.close()


def test_learned_warm_start_on_the_100k_pile():
    """The learned warm start at full size: trained on 2M contacts logged
    from the plain solver (32 passes) on piles of seeds 10-19, it runs the
    100,000-cuboid pile (seed 1) with 4 passes against the cache warm start
    with 8, over 300 ticks: the mean contact residual is no higher, and the
    mean and maximum penetration no more than 5% higher."""
    rng = np.random.default_rng(0)
    log_cfg = GpuPhysicsConfig(friction=FRICTION, solver_iterations=32)
    feats, targs = [], []
    for seed in range(10, 20):
        F, T = L.collect_rows(mh.named_pile("pile_100k", seed=seed), log_cfg, 120, range(40, 121, 20))
        pick = rng.choice(F.shape[0], size=min(200_000, F.shape[0]), replace=False)
        feats.append(F[pick])
        targs.append(T[pick])
    net = L.WarmStartNet(seed=3)
    losses = net.fit(np.concatenate(feats), np.concatenate(targs), epochs=20, batch=4096, learning_rate=1e-3)
    assert losses[-1] < losses[0]
    scene = mh.named_pile("pile_100k")
    plain = L.residual_stats(scene, GpuPhysicsConfig(friction=FRICTION, solver_iterations=8), 300)
    learned = L.residual_stats(scene, GpuPhysicsConfig(friction=FRICTION, solver_iterations=4), 300, net=net)
    summary = {k: (float(plain[k].mean()), float(learned[k].mean())) for k in plain}
    print(f"[learned 100k] plain 8 / learned 4 (means over ticks 1..300): {summary}", flush=True)
    for k in plain:
        assert np.all(np.isfinite(learned[k]))
    assert learned["rms"].mean() <= plain["rms"].mean()
    assert learned["mean_penetration"].mean() <= 1.05 * plain["mean_penetration"].mean()
    assert learned["max_penetration"].mean() <= 1.05 * plain["max_penetration"].mean()
