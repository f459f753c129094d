"""Throughput of the MI355X step on the 100,000-cuboid pile.

Workload: 100,000 cuboids of 0.25 m edge (density 1000 kg/m^3) in a 50x40x50
jittered, yawed lattice with 0.27 m spacing dropped 0.5 m onto the ground
plane; dt = 1/60 s, gravity -9.81, friction 0.6, restitution 0, 8 solver
iterations. One step is one whole tick (integration, broadphase, narrowphase,
contact solve, snapshot) replayed as a captured hipGraph with inputs resident in
HBM. Also reports, as sub-objects of the same JSON line: the learned warm start
against the reference's cache warm start at equal contact residual, the
256-scene batch (autotuned launch shape against the default) and the articulated
100,000-body scene.

    python bench.py [--gpus 1] [--steps 600] [--warmup 60] [--no-cpu] [--no-learned]
                    [--no-batched] [--no-articulated]
"""
import argparse
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def algorithmic_bytes(n, c, s, p, k, a, r, iterations):
    """Compulsory HBM bytes per tick and kernel group: each input read once and
    each output written once (per-unit figures in DESIGN.md). r is the number of
    delta slots the tick's contact sides took (one per body per 64-contact group)."""
    return {
        "integrate": 280 * n,
        "morton": 136 * c + 8 * (s - c),
        "sort": 16 * s,
        "lbvh": 24 * (c - 1) + 136 * c,
        "traverse": 40 * c + 72 * (2 * c - 1) + 8 * p,
        "narrowphase": 200 * p + 64 * k,
        "prepare": 340 * k,
        "solve": 12 * a + 4 * r + iterations * (solve_kernel_bytes(n, a, r) + 24 * r + 112 * n),
    }


LAUNCHES = {"integrate": 1, "morton": 2, "lbvh": 2, "traverse": 1, "narrowphase": 1, "prepare": 1}


def solve_kernel_bytes(n, a, r):
    """Compulsory HBM bytes of one contact-solve pass kernel (k_solve): per active
    contact its 64-byte solver row, its accumulated impulse read and written
    (4 + 4) and its sides' group word (4); per delta slot the Q20 velocity deltas
    of one body summed over a 64-contact group (24); per body its 64-byte solver
    state read once."""
    return 76 * a + 24 * r + 64 * n


# What FETCH_SIZE and WRITE_SIZE (KiB per dispatch) read for a plain device copy
# of a known 256 MiB: FETCH_SIZE counts half the bytes read, WRITE_SIZE all the
# bytes written (rocprofv3 on MI355X; the copy's summary is kept with the profiles).
FETCH_KIB_SCALE = 2.0
WRITE_KIB_SCALE = 1.0


TRAFFIC_SETTLE = 5  # ticks that rebuild the contact cache after the state is restored


def traffic_probe(args):
    """Child of measure_traffic, under rocprofv3 --pmc restricted to k_solve: the
    headline run's final body state in a fresh handle, stepped TRAFFIC_SETTLE +
    profile_ticks ticks (long runs under counter collection crash the profiler)."""
    import numpy as np

    import mechanic_amd as mh

    scene = mh.named_pile("pile_100k")
    gpu = mh.GpuPhysics.new_with_config(scene, mh.GpuPhysicsConfig(friction=0.6, solver_iterations=8))
    gpu.set_tuning(**json.loads(args.tuning_json))
    state = np.load(args.state)
    gpu.write_bodies(state["poses"], state["velocities"])
    gpu.dispatch_ticks(1, TRAFFIC_SETTLE + args.profile_ticks)
    gpu.synchronize()
    gpu.close()


def measure_traffic(args, tuning, iterations, state):
    """HBM bytes per k_solve launch: FETCH_SIZE and WRITE_SIZE from two
    rocprofv3 --pmc runs (child processes) continuing the headline run's final
    state, averaged over the launches of the last profile_ticks ticks and scaled
    by the copy calibration."""
    import shutil
    import subprocess
    import tempfile

    from mechanic_amd.rocprof_summary import counter_tail_average

    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    kib = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
            cmd = ["rocprofv3", "--pmc", counter, "--kernel-include-regex", "k_solve", "-d", tmp, "-o", "run", "--",
                   sys.executable, os.path.abspath(__file__), "--traffic-probe", "--state", state,
                   "--profile-ticks", str(args.profile_ticks), "--tuning-json", json.dumps(tuning)]
            env = dict(os.environ, TMPDIR="/tmp")
            try:
                # rocprofv3 may end with a signal of its own after the program and
                # its results are complete: the results database decides.
                res = subprocess.run(cmd, cwd="/tmp", env=env, timeout=300, stdout=subprocess.DEVNULL,
                                     stderr=subprocess.PIPE, text=True)
            except (subprocess.SubprocessError, OSError) as e:
                return None, f"{counter} run failed: {type(e).__name__}"
            dbs = [os.path.join(r, f) for r, _, fs in os.walk(tmp) for f in fs if f.endswith("results.db")]
            value = counter_tail_average(dbs[0], "k_solve", counter, iterations * args.profile_ticks) if dbs else None
            if value is None:
                tail = " | ".join(l for l in res.stderr.strip().splitlines()[-3:] if l)[-300:]
                return None, f"{counter}: no k_solve samples (exit {res.returncode}: {tail})"
            kib[counter] = value
    fetch = kib["FETCH_SIZE"] * FETCH_KIB_SCALE * 1024.0
    write = kib["WRITE_SIZE"] * WRITE_KIB_SCALE * 1024.0
    return {"bytes": fetch + write, "read": fetch, "written": write}, None


def cpu_baseline(scene, friction, iterations, budget_s=15.0, max_ticks=40):
    from tests._oracle import Oracle

    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    threads = max(1, min(threads, 16))
    ora = Oracle(scene, iterations=iterations, friction=friction, threads=threads)
    t0 = time.perf_counter()
    ticks = 0
    while ticks < max_ticks and time.perf_counter() - t0 < budget_s:
        ticks += 1
        ora.step(ticks)
    dt = time.perf_counter() - t0
    return {"value": ticks / dt, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP32 C++ restatement, OpenMP {threads} threads, ticks 1..{ticks} of the same "
                      f"100k-cuboid pile ({dt:.1f} s)"}


def learned_warm_start(mh, scene, friction, plain_iterations, steps, value_plain, tuning=None):
    """Learned warm start against the reference's cache warm start on the same
    pile: trained on 2M logged contacts of seeds 10-19 (32 solver passes), the
    smallest pass count whose mean contact residual over ticks 1..300 does not
    exceed the plain 8-pass one, and its steps/s under the headline protocol."""
    import numpy as np
    from mechanic_amd import learned as L

    t0 = time.perf_counter()
    rng = np.random.default_rng(0)
    feats, targs = [], []
    log_cfg = mh.GpuPhysicsConfig(friction=friction, solver_iterations=32)
    for seed in range(10, 20):
        F, T = L.collect_rows(mh.named_pile("pile_100k", seed=seed), log_cfg, 120, range(40, 121, 20))
        pick = rng.choice(F.shape[0], size=min(200_000, F.shape[0]), replace=False)
        feats.append(F[pick])
        targs.append(T[pick])
        del F, T
    F = np.concatenate(feats)
    T = np.concatenate(targs)
    collect_s = time.perf_counter() - t0
    net = L.WarmStartNet(seed=3)
    t1 = time.perf_counter()
    losses = net.fit(F, T, epochs=20, batch=4096, learning_rate=1e-3)
    train_s = time.perf_counter() - t1

    ticks = 300
    base_cfg = mh.GpuPhysicsConfig(friction=friction, solver_iterations=plain_iterations)
    plain = L.residual_trace(scene, base_cfg, ticks)
    tried = {}
    chosen = None
    for k in (3, 4, 5, 6):
        r = L.residual_trace(scene, mh.GpuPhysicsConfig(friction=friction, solver_iterations=k), ticks, net=net)
        tried[k] = float(r.mean())
        if r.mean() <= plain.mean():
            chosen = k
            break
    out = {"plain_iterations": plain_iterations, "mean_residual_plain": round(float(plain.mean()), 5),
           "mean_residual_learned": {str(k): round(v, 5) for k, v in tried.items()},
           "residual": "rms over active contacts of |max(0, lambda + (bias - v_n)/d) - lambda| after the last "
                       f"pass, mean over ticks 1..{ticks}",
           "train_rows": int(F.shape[0]), "train_data": "seeds 10-19 of the same 100k pile, 32 solver passes, "
                                                         "200k rows sampled from ticks 40..120 per seed",
           "collect_s": round(collect_s, 2), "train_s": round(train_s, 3), "epochs": 20,
           "final_loss": round(float(losses[-1]), 5), "learned_iterations": chosen}
    if chosen is None:
        return out
    cfg = mh.GpuPhysicsConfig(friction=friction, solver_iterations=chosen, use_graph=True,
                              warm_start=mh._lib.WARM_START_LEARNED)
    gpu = mh.GpuPhysics.new_with_config(scene, cfg)
    if tuning:
        gpu.set_tuning(**tuning)
    gpu.set_mlp(net)
    gpu.dispatch_ticks(1, 1)
    gpu.synchronize()
    t2 = time.perf_counter()
    gpu.dispatch_ticks(2, steps)
    gpu.synchronize()
    v = steps / (time.perf_counter() - t2)
    flags = gpu.read_last_tick().error_flags
    gpu.close()
    out.update({"iteration_ratio": round(plain_iterations / chosen, 3), "steps_per_s": round(v, 2),
                "steps_per_s_plain": round(value_plain, 2), "speedup": round(v / value_plain, 3),
                "error_flags": flags})
    return out


def save_tuning(autotune, bodies, tuned, workload):
    """Adds the winner to the cache file named by MECHANIC_TUNING_OUT, if set."""
    out_path = os.environ.get("MECHANIC_TUNING_OUT")
    if not out_path:
        return
    entries = autotune.load_cache(out_path)
    entries[autotune.cache_key(autotune.device_arch(), bodies)] = {
        "tuning": tuned["tuning"], "ms": tuned["ms"], "default_ms": tuned["default_ms"], "workload": workload}
    autotune.save_cache(entries, out_path)


def timed_run(make, steps, tuning=None):
    """Steps/s of a fresh handle under the headline protocol (graph captured on
    tick 1, ticks 2..steps+1 timed), with its last-tick diagnostics."""
    gpu = make()
    if tuning:
        gpu.set_tuning(**tuning)
    gpu.dispatch_ticks(1, 1)
    gpu.synchronize()
    t0 = time.perf_counter()
    gpu.dispatch_ticks(2, steps)
    gpu.synchronize()
    value = steps / (time.perf_counter() - t0)
    rb = gpu.read_last_tick()
    gpu.close()
    return value, rb


def batched(mh, steps):
    """256 independent scenes of 64..4,096 cuboids (seeds 1000+s) in one handle,
    one graph launch per step; the launch shape autotuned against the default."""
    from mechanic_amd import autotune
    from mechanic_amd.scenes import batch_scenes

    scenes = batch_scenes()
    cfg = mh.GpuPhysicsConfig(friction=0.6, solver_iterations=8)
    make = lambda: mh.GpuPhysics.new_batch(scenes, cfg)  # noqa: E731
    t0 = time.perf_counter()
    tuned = autotune.autotune(make, warm=20, ticks=40)
    tune_s = time.perf_counter() - t0
    default_v, _ = timed_run(make, steps)
    tuned_v, rb = timed_run(make, steps, tuned["tuning"])
    bodies = sum(s.body_count for s in scenes)
    save_tuning(autotune, bodies, tuned, "256-scene batch")
    return {"workload": "256 scenes, sizes round(exp(U[ln 64, ln 4096])) cuboids, seeds 1000+s, friction 0.6, "
                        f"8 iterations, ticks 2..{steps + 1}",
            "scenes": len(scenes), "bodies": bodies, "steps": steps,
            "steps_per_s_default": round(default_v, 2), "steps_per_s_tuned": round(tuned_v, 2),
            "scene_steps_per_s_tuned": round(tuned_v * len(scenes), 1),
            "tuning": tuned["tuning"], "tuning_trials": len(tuned["trials"]), "tune_s": round(tune_s, 2),
            "error_flags": rb.error_flags, "contacts": rb.contact_count}


def articulated(mh, steps):
    """The articulated 100,000-body scene: steps/s and ms per phase."""
    from mechanic_amd import articulated as A

    scene = A.articulated_pile()
    cfg = mh.GpuPhysicsConfig(friction=0.6, solver_iterations=8, mechanism_self_collisions=False)
    make = lambda: mh.GpuPhysics.new_with_config(scene, cfg)  # noqa: E731
    value, rb = timed_run(make, steps)
    gpu = make()
    gpu.dispatch_ticks(1, steps)
    acc = {}
    n = 10
    for i in range(n):
        t = gpu.profile_tick(steps + 1 + i)
        for k, v in t.items():
            acc[k] = acc.get(k, 0.0) + v / n
    prb = gpu.read_last_tick()
    gpu.close()
    return {"workload": "90,000 cuboids + 10,000 cylinders (OBB), 4,500 tree bearings in chains of 2-20 links, "
                        "500 loop closures, 1,000 drives at +-2 rad/s, friction 0.6, 8 iterations, "
                        f"mechanism self-collisions off, ticks 2..{steps + 1}",
            "steps": steps, "steps_per_s": round(value, 2),
            "phase_ms": {k[:-3]: round(v, 4) for k, v in acc.items() if k.endswith("_ms")},
            "error_flags": rb.error_flags, "anchor_residual_m": round(prb.anchor_residual_meters, 6),
            "contacts": prb.contact_count}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--profile-ticks", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--tuning", choices=("auto", "cached", "default"), default="auto",
                    help="launch shape: autotune now, the committed cache, or the defaults")
    ap.add_argument("--no-articulated", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 --pmc traffic runs")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--tuning-json", default="{}", help=argparse.SUPPRESS)
    ap.add_argument("--state", default="", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import mechanic_amd as mh

    friction, iterations = 0.6, 8
    scene = mh.named_pile("pile_100k")
    cfg = mh.GpuPhysicsConfig(friction=friction, solver_iterations=iterations, use_graph=True)

    # Warm-up on a throwaway handle (code objects, clocks, graph path).
    warm = mh.GpuPhysics.new_with_config(scene, cfg)
    for t in range(1, args.warmup + 1):
        warm.dispatch_tick(t)
    warm.synchronize()
    warm.close()

    # Launch shape (results never depend on it): the autotuner's winner for this
    # scene, the committed cache entry for its size bucket, or the defaults.
    from mechanic_amd import autotune
    make = lambda: mh.GpuPhysics.new_with_config(scene, cfg)  # noqa: E731
    t_tune = time.perf_counter()
    if args.tuning == "auto":
        # Tune on the pile as it is while it is timed (after ~300 ticks), not
        # while it still falls: every trial restores that state.
        pre = make()
        pre.dispatch_ticks(1, 300)
        pre.synchronize()
        state = pre.read_bodies()
        pre.close()
        cached = autotune.cached_tuning(scene.body_count)
        tuned = autotune.autotune(make, warm=10, ticks=40, state=state, candidates=[cached] if cached else [])
        save_tuning(autotune, scene.body_count, tuned, "100k pile")
    elif args.tuning == "cached":
        tuned = {"tuning": autotune.cached_tuning(scene.body_count) or dict(autotune.DEFAULTS)}
    else:
        tuned = {"tuning": dict(autotune.DEFAULTS)}
    tune_s = time.perf_counter() - t_tune
    default_value = timed_run(make, args.steps)[0] if args.tuning == "auto" else None

    gpu = mh.GpuPhysics.new_with_config(scene, cfg)
    gpu.set_tuning(**tuned["tuning"])
    gpu.dispatch_ticks(1, 1)  # captures the tick graph
    gpu.synchronize()
    t0 = time.perf_counter()
    gpu.dispatch_ticks(2, args.steps)
    gpu.synchronize()
    wall = time.perf_counter() - t0
    steps = args.steps
    readback = gpu.read_last_tick()

    # Event-bracketed (uncaptured) ticks continuing the same run: ms per kernel group.
    groups = ["integrate_ms", "morton_ms", "sort_ms", "lbvh_ms", "traverse_ms", "narrowphase_ms", "prepare_ms",
              "solve_ms"]
    acc = {g: 0.0 for g in groups}
    kernel_ms = {"solve_kernel_ms": 0.0, "apply_kernel_ms": 0.0}
    total = 0.0
    counts = []
    first = steps + 2
    for i in range(args.profile_ticks):
        t = gpu.profile_tick(first + i)
        r = gpu.read_last_tick()
        slots = numpy.frombuffer(gpu.read_buffer(mh._lib.BUFFER_BODY_SLOTS, 4 * scene.body_count), numpy.uint32)
        counts.append((r.pair_count, r.contact_count, r.active_contact_count, int(slots.sum(dtype=numpy.int64))))
        for g in groups:
            acc[g] += t[g]
        for g in kernel_ms:
            kernel_ms[g] += t[g]
        total += t["total_ms"]
    npt = max(args.profile_ticks, 1)
    phase_ms = {g[:-3]: acc[g] / npt for g in groups}
    p = sum(c[0] for c in counts) / npt
    k = sum(c[1] for c in counts) / npt
    a = sum(c[2] for c in counts) / npt
    slots = sum(c[3] for c in counts) / npt
    n = scene.body_count
    c = len(scene.colliders)
    s = 1 << (c - 1).bit_length()
    algo = algorithmic_bytes(n, c, s, p, k, a, slots, iterations)
    # Roofline of the dominant kernel, k_solve (8 launches a tick): compulsory
    # bytes per launch over its mean launch span in the profiled ticks, stamped
    # on the device wall clock by the kernel's own workgroups (first start to
    # last end). HIP events bracket the kernel groups; around a single kernel
    # they add several microseconds of their own on this runtime.
    solve_ms = kernel_ms["solve_kernel_ms"] / npt
    apply_ms = kernel_ms["apply_kernel_ms"] / npt
    solve_bytes = solve_kernel_bytes(scene.body_count, a, slots)
    achieved = solve_bytes / (solve_ms * 1e-3) / 1e9 if solve_ms > 0 else 0.0
    per_phase_gbs = {g: (algo[g] / (phase_ms[g] * 1e-3) / 1e9 if phase_ms[g] > 0 else 0.0) for g in phase_ms}

    # Reference-style synchronous engine TPS: dispatch + diagnostics readback per tick.
    sync_ticks = 60
    t1 = time.perf_counter()
    base = first + args.profile_ticks
    for i in range(sync_ticks):
        gpu.dispatch_tick(base + i)
        gpu.read_last_tick()
    engine_tps = sync_ticks / (time.perf_counter() - t1)
    import tempfile

    import numpy as np

    state_fd, state_path = tempfile.mkstemp(suffix=".npz", dir="/tmp")
    os.close(state_fd)
    final_poses, final_velocities = gpu.read_bodies()
    np.savez(state_path, poses=final_poses, velocities=final_velocities)
    gpu.close()

    value = steps / wall
    learned = None if args.no_learned else learned_warm_start(mh, scene, friction, iterations, steps, value,
                                                                  tuned["tuning"])
    batch = None if args.no_batched else batched(mh, steps)
    art = None if args.no_articulated else articulated(mh, min(steps, 300))
    cpu = None if args.no_cpu else cpu_baseline(scene, friction, iterations)
    traffic, traffic_note = (None, "skipped (--no-traffic)") if args.no_traffic else \
        measure_traffic(args, tuned["tuning"], iterations, state_path)
    os.remove(state_path)

    out = {
        "metric": "steps/sec at 100k bodies + ms/phase",
        "value": round(value, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * wall / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 jittered cuboid lattice)",
        "config": {"workload": "pile_100k: 100,000 cuboids 0.25 m, 50x40x50 lattice 0.27 m spacing, "
                               "jitter 0.02 m, uniform yaw, 0.5 m drop, seed 1, friction 0.6, 8 iterations, "
                               f"ticks 2..{steps + 1} after the drop",
                   "bodies": n, "solver_iterations": iterations, "friction": friction, "parallelism": "1 GPU"},
        "phase_ms": {g: round(v, 4) for g, v in phase_ms.items()},
        "phase_gbs": {g: round(v, 1) for g, v in per_phase_gbs.items()},
        "profiled_tick_ms": round(total / npt, 4),
        "counts": {"pairs": round(p), "contacts": round(k), "active_contacts": round(a), "delta_slots": round(slots)},
        "error_flags": readback.error_flags,
        "engine_tps_sync": round(engine_tps, 2),
        "tuning": tuned["tuning"],
        "steps_per_s_default_tuning": None if default_value is None else round(default_value, 2),
        "tune_s": round(tune_s, 2),
        "roofline": {"bound": "hbm", "kernel": "k_solve", "launches_per_tick": iterations,
                     "algorithmic_bytes_per_launch": round(solve_bytes), "launch_us": round(solve_ms * 1e3, 2),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None if traffic is None else round(traffic["bytes"]),
                     "traffic_read": None if traffic is None else round(traffic["read"]),
                     "traffic_written": None if traffic is None else round(traffic["written"]),
                     "traffic_note": traffic_note or "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of k_solve, continuing "
                                                     "the final state, x2.0 / x1.0 KiB (copy calibration)"},
        "kernel_us": {"k_solve": round(solve_ms * 1e3, 2), "k_apply": round(apply_ms * 1e3, 2)},
        "cpu_baseline": cpu,
        "learned_warm_start": learned,
        "batched": batch,
        "articulated": art,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
