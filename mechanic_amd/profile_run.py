"""Steps one named workload with direct kernel launches, for profilers.

    python -m mechanic_amd.profile_run pile_100k --ticks 20 [--settle 300]
    python -m mechanic_amd.profile_run articulated --ticks 20 [--closures config2|reference]
    python -m mechanic_amd.profile_run batch --ticks 20

(batch: the 256 scenes of 64..4,096 cuboids in one handle, with the cached
launch shape for its size.)

Captured graph launches hide the individual kernels from some tracers, so this
driver dispatches every kernel directly (use_graph off). It prints one line
with the workload, the ticks run and the last tick's diagnostics.
"""
from __future__ import annotations

import argparse
import sys

import mechanic_amd as mh

# The articulated configuration's closure solve: one PCG per closure component
# to a 1e-6 tolerance (at most 64 iterations), 24 Newton steps a tick, speeds
# projected onto the loops' velocity constraint.
CONFIG2_CLOSURES = dict(closure_components=True, closure_pcg_tolerance=1e-6, closure_pcg_iterations=64,
                        closure_newton_steps=24, closure_velocity_projection=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=("pile_100k", "pile_1k", "articulated", "batch"))
    ap.add_argument("--ticks", type=int, default=20)
    ap.add_argument("--settle", type=int, default=0, help="untimed ticks first (pile: 300 for the settled state)")
    ap.add_argument("--closures", choices=("config2", "reference"), default="config2")
    args = ap.parse_args(argv)
    if args.workload == "articulated":
        from mechanic_amd import articulated as A

        scene = A.articulated_pile()
        extra = CONFIG2_CLOSURES if args.closures == "config2" else {}
        cfg = mh.GpuPhysicsConfig(friction=0.6, solver_iterations=8, mechanism_self_collisions=False,
                                  use_graph=False, **extra)
    elif args.workload == "batch":
        from mechanic_amd import autotune
        from mechanic_amd.scenes import batch_scenes

        scenes = batch_scenes()
        cfg = mh.GpuPhysicsConfig(friction=0.6, solver_iterations=8, use_graph=False)
    else:
        scene = mh.named_pile(args.workload)
        cfg = mh.GpuPhysicsConfig(friction=0.6, solver_iterations=8, use_graph=False)
    if args.workload == "batch":
        handle = mh.GpuPhysics.new_batch(scenes, cfg)
        tuning = autotune.cached_tuning(sum(sc.body_count for sc in scenes))
        if tuning:
            handle.set_tuning(**tuning)
    else:
        handle = mh.GpuPhysics.new_with_config(scene, cfg)
    with handle as gpu:
        last = args.settle + args.ticks
        for tick in range(1, last + 1):
            gpu.dispatch_tick(tick)
        r = gpu.read_last_tick()
    print(f"{args.workload}: ticks 1..{last}, pairs {r.pair_count}, contacts {r.contact_count}, active "
          f"{r.active_contact_count}, flags {r.error_flags}, anchor residual {r.anchor_residual_meters:.2e} m",
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
