"""Throughput of the MI355X step on the 100,000-cuboid pile.

Workload: 100,000 cuboids of 0.25 m edge (density 1000 kg/m^3) in a 50x40x50
jittered, yawed lattice with 0.27 m spacing dropped 0.5 m onto the ground
plane; dt = 1/60 s, gravity -9.81, friction 0.6, restitution 0, 8 solver
iterations. One step is one whole tick (integration, broadphase, narrowphase,
contact solve, snapshot) replayed as a captured hipGraph with inputs resident in
HBM. The timed window is the settled pile: the K steps after the first
SETTLE_TICKS (300) ticks from the drop, where config 1's 600 steps spend most of
their time; the falling window (ticks 2..K+1) and the whole 600-step run are
reported beside it. Also reports, as sub-objects of the same JSON line: the
learned warm start against the reference's cache warm start at equal contact
residual and penetration, the 256-scene batch (autotuned launch shape against
the default), the articulated 100,000-body scene and the CPU oracle at one and
at all host threads.

    python bench.py [--gpus 1] [--steps 600] [--warmup 60] [--no-cpu] [--no-learned]
                    [--no-batched] [--no-articulated] [--no-traffic] [--no-graph]
"""
import argparse
import json
import os
import sys
import tempfile
import time

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
        "prepare": 328 * k,
        "solve": iterations * (solve_kernel_bytes(n, a, r, iterations) + 24 * r + 112 * n),
    }


LAUNCHES = {"integrate": 1, "morton": 2, "lbvh": 2, "traverse": 1, "narrowphase": 1, "prepare": 1}


def solve_kernel_bytes(n, a, r, iterations):
    """Compulsory HBM bytes of one contact-solve pass kernel (k_solve), as the
    mean over a tick's passes: per active contact its 52-byte solver row, its
    accumulated impulse read and written (4 + 4) and its sides' group word (4,
    written by the first pass and read by the others), plus its contact index
    (4, read by the last pass to persist the impulse); per delta slot the Q20
    velocity deltas of one body summed over a 64-contact group (24), plus the
    first pass's slot counter increment (4); per body its 64-byte solver state
    read once. The once-a-tick terms are spread over the passes."""
    return 64 * a + 24 * r + 64 * n + (4 * a + 4 * r) / iterations


# HBM traffic of a launch from the L2's memory-side request counters, one
# rocprofv3 --pmc pass each, with every request counted at its own size. On
# MI355X the derived FETCH_SIZE counts a 128-byte read as 64 bytes (its
# 128-byte term, TCC_BUBBLE, reads 0 on gfx950), so it reports half of a plain
# copy's reads; these counters read a 256 MiB copy's 268435456 bytes exactly
# (the calibration summary is kept with the profiles).
READ_COUNTERS = {"TCC_EA0_RDREQ_32B": 32, "TCC_EA0_RDREQ_64B": 64, "TCC_EA0_RDREQ_128B": 128}
WRITE_COUNTERS = ("TCC_EA0_WRREQ", "TCC_EA0_WRREQ_64B")  # all writes; the 64-byte ones (the rest are 32 B)


TRAFFIC_SETTLE = 5  # ticks that rebuild the contact cache after the state is restored
SETTLE_TICKS = 300  # ticks from the drop before the timed window


def traffic_probe(args):
    """Child of measure_traffic, under rocprofv3 --pmc restricted to k_solve: the
    settled state in a fresh handle, stepped TRAFFIC_SETTLE + profile_ticks
    ticks (long runs under counter collection crash the profiler). Writes the
    active-contact and delta-slot counts of the profiled ticks, so the parent
    forms the algorithmic bytes of exactly the launches the counters cover."""
    import numpy as np

    import mechanic_amd as mh

    scene = mh.named_pile("pile_100k")
    gpu = mh.GpuPhysics.new_with_config(scene, mh.GpuPhysicsConfig(friction=0.6, solver_iterations=8))
    gpu.set_tuning(**json.loads(args.tuning_json))
    state = np.load(args.state)
    gpu.write_bodies(state["poses"], state["velocities"])
    first = int(state["tick"]) + 1
    gpu.dispatch_ticks(first, TRAFFIC_SETTLE)
    counts = []
    for i in range(args.profile_ticks):
        gpu.dispatch_tick(first + TRAFFIC_SETTLE + i)
        r = gpu.read_last_tick()
        slots = np.frombuffer(gpu.read_buffer(mh._lib.BUFFER_BODY_SLOTS, 4 * scene.body_count), np.uint32)
        counts.append([r.active_contact_count, int(slots.sum(dtype=np.int64))])
    gpu.close()
    with open(args.counts_out, "w") as f:
        json.dump({"bodies": scene.body_count, "ticks": counts}, f)


def progress(message):
    """One line on stderr per finished phase (stdout carries only the JSON line)."""
    print(f"bench: {message}", file=sys.stderr, flush=True)


def measure_traffic(args, tuning, iterations, state):
    """HBM bytes per k_solve launch from two rocprofv3 --pmc runs (child
    processes) continuing the settled state: the read requests by size, then
    the write requests and their 64-byte share, averaged over the launches of
    the child's profile_ticks ticks, with the algorithmic bytes of the same
    launches. A child that ends with a non-zero status is a failed measurement."""
    import shutil
    import subprocess
    import tempfile

    from mechanic_amd.rocprof_summary import counter_tail_average

    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    mean = {}
    algorithmic = []
    for counters in (tuple(READ_COUNTERS), WRITE_COUNTERS):
        with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
            counts_path = os.path.join(tmp, "counts.json")
            cmd = ["rocprofv3", "--pmc", *counters, "--kernel-include-regex", "k_solve", "-d", tmp, "-o", "run",
                   "--", sys.executable, os.path.abspath(__file__), "--traffic-probe", "--state", state,
                   "--profile-ticks", str(args.profile_ticks), "--tuning-json", json.dumps(tuning),
                   "--counts-out", counts_path]
            env = dict(os.environ, TMPDIR="/tmp")
            try:
                res = subprocess.run(cmd, cwd="/tmp", env=env, timeout=300, stdout=subprocess.DEVNULL,
                                     stderr=subprocess.PIPE, text=True)
            except (subprocess.SubprocessError, OSError) as e:
                return None, f"{counters[0]} run failed: {type(e).__name__}"
            tail = " | ".join(l for l in res.stderr.strip().splitlines()[-3:] if l)[-300:]
            if res.returncode != 0 or not os.path.exists(counts_path):
                return None, f"{counters[0]} run exited with status {res.returncode}: {tail}"
            dbs = [os.path.join(r, f) for r, _, fs in os.walk(tmp) for f in fs if f.endswith("results.db")]
            for counter in counters:
                value = counter_tail_average(dbs[0], "k_solve", counter, iterations * args.profile_ticks) \
                    if dbs else None
                if value is None:
                    return None, f"{counter}: no k_solve samples ({tail})"
                mean[counter] = value
            with open(counts_path) as f:
                c = json.load(f)
            algorithmic += [solve_kernel_bytes(c["bodies"], a, r, iterations) for a, r in c["ticks"]]
    read = sum(mean[c] * size for c, size in READ_COUNTERS.items())
    all_writes, writes_64 = (mean[c] for c in WRITE_COUNTERS)
    written = 64.0 * writes_64 + 32.0 * (all_writes - writes_64)
    algo = sum(algorithmic) / len(algorithmic)
    return {"bytes": read + written, "read": read, "written": written, "algorithmic": algo}, None


def cpu_baseline(scene, friction, iterations, state):
    """The FP32 oracle (the C++ restatement of the reference kernels) on the
    host cores, from the same settled state as the timed window: all host
    threads (capped at the 16 a GPU's share of the host holds) for up to 8 s,
    then one thread for up to 12 s. Reports steps/s and ms per phase."""
    import numpy as np

    from tests._oracle import Oracle

    try:
        available = len(os.sched_getaffinity(0))
    except AttributeError:
        available = os.cpu_count() or 1
    threads = max(1, min(available, int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16))
    data = np.load(state)
    first = int(data["tick"]) + 1
    out = {}
    for label, t, budget, cap in (("all", threads, 8.0, 30), ("one", 1, 12.0, 12)):
        ora = Oracle(scene, iterations=iterations, friction=friction, threads=t)
        ora.set_bodies(data["poses"], data["velocities"])
        ora.step(first)  # rebuilds the manifold cache, untimed
        ora.phase_ms(reset=True)
        t0 = time.perf_counter()
        ticks = 0
        while ticks < cap and time.perf_counter() - t0 < budget:
            ora.step(first + 1 + ticks)
            ticks += 1
        dt = time.perf_counter() - t0
        phases = {k: round(v / ticks, 2) for k, v in ora.phase_ms().items() if v > 0.0}
        out[label] = {"threads": t, "ticks": ticks, "steps_per_s": round(ticks / dt, 3), "phase_ms": phases}
        del ora
    return {"value": out["all"]["steps_per_s"], "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP32 C++ restatement (-O3, no contraction), settled 100k pile restored at tick "
                      f"{first - 1}, one untimed tick, then {out['all']['ticks']} ticks on {threads} threads and "
                      f"{out['one']['ticks']} on one thread",
            "single_thread_steps_per_s": out["one"]["steps_per_s"],
            "phase_ms_per_tick": {"threads_" + str(threads): out["all"]["phase_ms"],
                                  "threads_1": out["one"]["phase_ms"]}}


def learned_warm_start(mh, scene, friction, plain_iterations, steps, value_plain, tuning=None, use_graph=True,
                       plain_state=None, plain_counts=None):
    """Learned warm start against the reference's cache warm start on the same
    pile: trained on 2M logged contacts of seeds 10-19 (32 solver passes), the
    smallest pass count whose mean contact residual over ticks 1..300 does not
    exceed the plain 8-pass one, with the mean and maximum penetration of both
    runs, and its steps/s under the headline protocol (settled window). A pile
    stepped with fewer passes settles into another state (fewer pairs), so the
    last tick's counts of both windows are reported, and the learned rate is
    also measured from the plain run's own settled state (`plain_state`)."""
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
    plain = L.residual_stats(scene, base_cfg, ticks)
    tried = {}
    chosen = None
    stats = None
    for k in (3, 4, 5, 6):
        r = L.residual_stats(scene, mh.GpuPhysicsConfig(friction=friction, solver_iterations=k), ticks, net=net)
        tried[k] = float(r["rms"].mean())
        if r["rms"].mean() <= plain["rms"].mean():
            chosen, stats = k, r
            break
    mean3 = lambda r: {key: round(float(v.mean()), 6) for key, v in r.items()}  # noqa: E731
    out = {"plain_iterations": plain_iterations, "mean_residual_plain": round(float(plain["rms"].mean()), 5),
           "mean_residual_learned": {str(k): round(v, 5) for k, v in tried.items()},
           "plain_means": mean3(plain), "learned_means": None if stats is None else mean3(stats),
           "residual": "rms over active contacts of |max(0, lambda + (bias - v_n)/d) - lambda| after the last "
                       f"pass; penetration mean and max over active contacts; all averaged over ticks 1..{ticks}",
           "train_rows": int(F.shape[0]), "train_data": "seeds 10-19 of the same 100k pile, 32 solver passes, "
                                                         "200k rows sampled from ticks 40..120 per seed",
           "collect_s": round(collect_s, 2), "train_s": round(train_s, 3), "epochs": 20,
           "final_loss": round(float(losses[-1]), 5), "learned_iterations": chosen}
    if chosen is None:
        return out
    def learned_rate(extra):
        cfg = mh.GpuPhysicsConfig(friction=friction, solver_iterations=chosen, use_graph=use_graph,
                                  warm_start=mh._lib.WARM_START_LEARNED, **extra)
        gpu = mh.GpuPhysics.new_with_config(scene, cfg)
        if tuning:
            gpu.set_tuning(**tuning)
        gpu.set_mlp(net)
        gpu.dispatch_ticks(1, SETTLE_TICKS)
        gpu.synchronize()
        before = gpu.read_learned_stats()["fallback_ticks"]
        t2 = time.perf_counter()
        gpu.dispatch_ticks(SETTLE_TICKS + 1, steps)
        gpu.synchronize()
        rate = steps / (time.perf_counter() - t2)
        rb = gpu.read_last_tick()
        fired = gpu.read_learned_stats()["fallback_ticks"] - before
        gpu.close()
        learned_rate.last = {"pairs": rb.pair_count, "active_contacts": rb.active_contact_count}
        return rate, rb.error_flags, fired

    v, flags, _ = learned_rate({})
    out.update({"iteration_ratio": round(plain_iterations / chosen, 3), "steps_per_s": round(v, 2),
                "steps_per_s_plain": round(value_plain, 2), "speedup": round(v / value_plain, 3),
                "error_flags": flags,
                "window_last_tick": {"learned": learned_rate.last, "plain": plain_counts}})
    if plain_state is not None:
        # The same workload as the headline's: the plain run's settled state
        # restored into a learned handle, 5 ticks to form its contacts, then
        # `steps` timed ticks.
        data = np.load(plain_state)
        cfg = mh.GpuPhysicsConfig(friction=friction, solver_iterations=chosen, use_graph=use_graph,
                                  warm_start=mh._lib.WARM_START_LEARNED)
        gpu = mh.GpuPhysics.new_with_config(scene, cfg)
        if tuning:
            gpu.set_tuning(**tuning)
        gpu.set_mlp(net)
        gpu.write_bodies(data["poses"], data["velocities"])
        first = int(data["tick"]) + 1
        gpu.dispatch_ticks(first, 5)
        gpu.synchronize()
        t3 = time.perf_counter()
        gpu.dispatch_ticks(first + 5, steps)
        gpu.synchronize()
        v_same = steps / (time.perf_counter() - t3)
        rb = gpu.read_last_tick()
        gpu.close()
        out["from_plain_state"] = {"steps_per_s": round(v_same, 2), "speedup": round(v_same / value_plain, 3),
                                   "ticks": f"{first + 5}..{first + 4 + steps} from the plain run's tick "
                                            f"{first - 1} state",
                                   "last_tick": {"pairs": rb.pair_count, "active_contacts": rb.active_contact_count},
                                   "error_flags": rb.error_flags}
    # The fallback: a tick whose residual after the learned passes exceeds the
    # plain run's mean runs `chosen` more passes (judged on the device).
    limit = float(plain["rms"].mean())
    fb = {"learned_residual_limit": limit, "learned_fallback_passes": chosen}
    fb_cfg = mh.GpuPhysicsConfig(friction=friction, solver_iterations=chosen, residual=True,
                                 warm_start=mh._lib.WARM_START_LEARNED, **fb)
    per_tick = {"rms": [], "mean_penetration": [], "max_penetration": []}
    fired = []
    with mh.GpuPhysics.new_with_config(scene, fb_cfg) as gpu:
        gpu.set_mlp(net)
        for tick in range(1, ticks + 1):
            gpu.dispatch_tick(tick)
            r = gpu.read_residual()
            for key in per_tick:
                per_tick[key].append(r[key])
            fired.append(gpu.read_learned_stats()["fallback_this_tick"])
        stats = gpu.read_learned_stats()
    fb_rate, fb_flags, fb_fired = learned_rate(fb)
    out["fallback"] = {
        "rule": f"a tick whose rms residual after the {chosen} learned passes exceeds {limit:.4f} (the plain "
                f"{plain_iterations}-pass mean over ticks 1..{ticks}) runs {chosen} more passes; a non-finite "
                "prediction takes the cache warm start",
        "fired_ticks": int(stats["fallback_ticks"]), "of_ticks": ticks,
        "fired_in_ticks_1_100": int(sum(fired[:100])), "fired_in_ticks_101_300": int(sum(fired[100:])),
        "nonfinite_predictions": int(stats["nonfinite_total"]),
        "means": {key: round(float(np.mean(v)), 6) for key, v in per_tick.items()},
        "steps_per_s_settled": round(fb_rate, 2), "fired_in_settled_window": int(fb_fired),
        "settled_window_ticks": f"{SETTLE_TICKS + 1}..{SETTLE_TICKS + steps}", "error_flags": fb_flags}
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


def timed_run(make, steps, tuning=None, settle=0):
    """Steps/s of a fresh handle under the headline protocol (graph captured on
    the first tick; with `settle`, ticks 1..settle run untimed first; the next
    `steps` ticks are timed), with its last-tick diagnostics."""
    gpu = make()
    if tuning:
        gpu.set_tuning(**tuning)
    gpu.dispatch_ticks(1, max(settle, 1))
    gpu.synchronize()
    t0 = time.perf_counter()
    gpu.dispatch_ticks(max(settle, 1) + 1, steps)
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


# Config 2's closure solve: one PCG per closure component to a 1e-6 tolerance
# (at most 64 iterations), 24 Newton steps a tick, and the coordinate speeds
# projected onto the loops' velocity constraint before the tree reconstruction.
CONFIG2_CLOSURES = dict(closure_components=True, closure_pcg_tolerance=1e-6, closure_pcg_iterations=64,
                        closure_newton_steps=24, closure_velocity_projection=True)


def articulated(mh, steps, use_graph=True):
    """The articulated 100,000-body scene: steps/s, ms per phase, the
    non-convergence flag and the worst closure anchor residual, with config 2's
    closure solve and with the reference's (one global PCG, 8 iterations, 12
    Newton steps)."""
    from mechanic_amd import articulated as A

    scene = A.articulated_pile()
    out = {"workload": "90,000 cuboids + 10,000 cylinders (OBB), 4,500 tree bearings in chains of 2-20 links, "
                       "500 loop closures (four-bars packed on the longest chains), 1,000 drives at +-2 rad/s "
                       "on the chains without loops, friction 0.6, 8 iterations, "
                       f"mechanism self-collisions off, ticks 2..{steps + 1}"}
    for label, closures in (("config2", CONFIG2_CLOSURES), ("reference_closures", {})):
        cfg = mh.GpuPhysicsConfig(friction=0.6, solver_iterations=8, mechanism_self_collisions=False,
                                  use_graph=use_graph, **closures)
        make = lambda: mh.GpuPhysics.new_with_config(scene, cfg)  # noqa: E731
        value, rb = timed_run(make, steps)
        gpu = make()
        worst = 0.0
        flags = 0
        for t in range(1, steps + 1):
            gpu.dispatch_tick(t)
            if t % 10 == 0 or t == steps:
                r = gpu.read_last_tick()
                worst = max(worst, r.anchor_residual_meters)
                flags |= r.error_flags
        acc = {}
        n = 10
        for i in range(n):
            t = gpu.profile_tick(steps + 1 + i)
            for k, v in t.items():
                acc[k] = acc.get(k, 0.0) + v / n
        prb = gpu.read_last_tick()
        gpu.close()
        out[label] = {"closure_solve": "per component, PCG tolerance 1e-6, max 64 iterations, 24 Newton steps, "
                                       "speeds projected onto the loops' velocity constraint"
                      if closures else "global PCG, 8 iterations, 12 Newton steps (the reference's)",
                      "steps_per_s": round(value, 2),
                      "phase_ms": {k[:-3]: round(v, 4) for k, v in acc.items() if k.endswith("_ms")},
                      "error_flags": flags, "worst_anchor_residual_m": round(worst, 7),
                      "contacts": prb.contact_count}
    out["steps"] = steps
    out["steps_per_s"] = out["config2"]["steps_per_s"]
    return out


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
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every tick's kernels directly instead of replaying the captured graph")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--tuning-json", default="{}", help=argparse.SUPPRESS)
    ap.add_argument("--state", default="", help=argparse.SUPPRESS)
    ap.add_argument("--counts-out", default="", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import mechanic_amd as mh

    friction, iterations = 0.6, 8
    steps = args.steps
    use_graph = not args.no_graph
    scene = mh.named_pile("pile_100k")
    cfg = mh.GpuPhysicsConfig(friction=friction, solver_iterations=iterations, use_graph=use_graph)
    make = lambda: mh.GpuPhysics.new_with_config(scene, cfg)  # noqa: E731

    # Warm-up on a throwaway handle (code objects, clocks, graph path).
    warm = make()
    for t in range(1, args.warmup + 1):
        warm.dispatch_tick(t)
    warm.synchronize()
    warm.close()

    # Launch shape (results never depend on it): the autotuner's winner for this
    # scene, the committed cache entry for its size bucket, or the defaults.
    from mechanic_amd import autotune
    t_tune = time.perf_counter()
    if args.tuning == "auto":
        # Tune on the settled pile, as it is in the timed window: every trial
        # restores the state after SETTLE_TICKS ticks.
        pre = make()
        pre.dispatch_ticks(1, SETTLE_TICKS)
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
    progress(f"launch shape chosen ({args.tuning}, {tune_s:.1f} s)")
    default_value = timed_run(make, steps, settle=SETTLE_TICKS)[0] if args.tuning == "auto" else None

    # The headline: ticks SETTLE_TICKS+1 .. SETTLE_TICKS+steps of one handle.
    gpu = make()
    gpu.set_tuning(**tuned["tuning"])
    gpu.dispatch_ticks(1, SETTLE_TICKS)  # the first tick captures the tick graph
    gpu.synchronize()
    t0 = time.perf_counter()
    gpu.dispatch_ticks(SETTLE_TICKS + 1, steps)
    gpu.synchronize()
    wall = time.perf_counter() - t0
    readback = gpu.read_last_tick()
    value = steps / wall
    tick = SETTLE_TICKS + steps

    # Reference-style synchronous engine TPS: dispatch + diagnostics readback per tick.
    sync_ticks = 60
    t1 = time.perf_counter()
    for i in range(sync_ticks):
        gpu.dispatch_tick(tick + 1 + i)
        gpu.read_last_tick()
    engine_tps = sync_ticks / (time.perf_counter() - t1)
    tick += sync_ticks

    # The state the profiled ticks start from: the traffic children and the CPU
    # baseline restore it.
    state_fd, state_path = tempfile.mkstemp(suffix=".npz", dir="/tmp")
    os.close(state_fd)
    poses0, velocities0 = gpu.read_bodies()
    np.savez(state_path, poses=poses0, velocities=velocities0, tick=np.int64(tick))

    # Event-bracketed (uncaptured) ticks continuing the same run: ms per kernel
    # group, and the device wall-clock spans of the solver-pass kernels.
    groups = ["integrate_ms", "morton_ms", "sort_ms", "lbvh_ms", "traverse_ms", "narrowphase_ms", "prepare_ms",
              "solve_ms"]
    acc = {g: 0.0 for g in groups}
    kernel_ms = {"solve_kernel_ms": 0.0, "apply_kernel_ms": 0.0, "solve_kernel_event_ms": 0.0}
    total = 0.0
    counts = []
    for i in range(args.profile_ticks):
        t = gpu.profile_tick(tick + 1 + i)
        r = gpu.read_last_tick()
        slots = np.frombuffer(gpu.read_buffer(mh._lib.BUFFER_BODY_SLOTS, 4 * scene.body_count), np.uint32)
        counts.append((r.pair_count, r.contact_count, r.active_contact_count, int(slots.sum(dtype=np.int64))))
        for g in groups:
            acc[g] += t[g]
        for g in kernel_ms:
            kernel_ms[g] += t[g]
        total += t["total_ms"]
    gpu.close()
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
    # bytes per launch over its mean launch duration in the profiled ticks,
    # timed by the start and stop HIP events each launch carries on its stream
    # (hipExtLaunchKernelGGL: the dispatch's own start and end, what rocprofv3's
    # dispatch records measure). The device wall-clock span of the kernel's
    # workgroups (first start to last end) is reported beside it.
    solve_ms = kernel_ms["solve_kernel_event_ms"] / npt
    span_ms = kernel_ms["solve_kernel_ms"] / npt
    apply_ms = kernel_ms["apply_kernel_ms"] / npt
    solve_bytes = solve_kernel_bytes(scene.body_count, a, slots, iterations)
    achieved = solve_bytes / (solve_ms * 1e-3) / 1e9 if solve_ms > 0 else 0.0
    per_phase_gbs = {g: (algo[g] / (phase_ms[g] * 1e-3) / 1e9 if phase_ms[g] > 0 else 0.0) for g in phase_ms}

    # The falling window (ticks 2..steps+1) and config 1 as specified: 600 steps
    # from the drop (ticks 2..601, the graph captured on tick 1).
    fall_value, _ = timed_run(make, steps, tuned["tuning"])
    run600_value, run600_rb = timed_run(make, 600, tuned["tuning"])

    progress(f"headline {value:.1f} steps/s; falling and 600-step windows done")
    learned = None if args.no_learned else learned_warm_start(
        mh, scene, friction, iterations, steps, value, tuned["tuning"], use_graph, plain_state=state_path,
        plain_counts={"pairs": readback.pair_count, "active_contacts": readback.active_contact_count})
    progress("learned warm start done")
    batch = None if args.no_batched else batched(mh, steps)
    progress("batched handles done")
    art = None if args.no_articulated else articulated(mh, min(steps, 300), use_graph)
    progress("articulated pile done")
    cpu = None if args.no_cpu else cpu_baseline(scene, friction, iterations, state_path)
    progress("cpu baseline done")
    traffic, traffic_note = (None, "skipped (--no-traffic)") if args.no_traffic else \
        measure_traffic(args, tuned["tuning"], iterations, state_path)
    progress("traffic counters done")
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
                               "jitter 0.02 m, uniform yaw, 0.5 m drop, seed 1, friction 0.6, 8 iterations; "
                               f"settled window: ticks {SETTLE_TICKS + 1}..{SETTLE_TICKS + steps} after the drop",
                   "bodies": n, "solver_iterations": iterations, "friction": friction, "parallelism": "1 GPU",
                   "launch": "captured hipGraph per tick" if use_graph else "direct launches"},
        "windows": {"settled_steps_per_s": round(value, 2),
                    "falling_steps_per_s": round(fall_value, 2), "falling_ticks": f"2..{steps + 1}",
                    "config1_600_steps_per_s": round(run600_value, 2), "config1_ticks": "2..601",
                    "config1_last_tick": {"pairs": run600_rb.pair_count, "contacts": run600_rb.contact_count,
                                          "active_contacts": run600_rb.active_contact_count,
                                          "error_flags": run600_rb.error_flags}},
        "phase_ms": {g: round(v, 4) for g, v in phase_ms.items()},
        "phase_gbs": {g: round(v, 1) for g, v in per_phase_gbs.items()},
        "profiled_tick_ms": round(total / npt, 4),
        "profiled_ticks": f"{tick + 1}..{tick + npt}",
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
                     "traffic_algorithmic_bytes": None if traffic is None else round(traffic["algorithmic"]),
                     "traffic_ratio": None if traffic is None else round(traffic["bytes"] / traffic["algorithmic"], 3),
                     "traffic_note": traffic_note or "rocprofv3 --pmc of k_solve's L2-to-memory requests by "
                                                     "size (TCC_EA0_RDREQ_32B/64B/128B, TCC_EA0_WRREQ and "
                                                     f"_64B) over the {npt} ticks after {TRAFFIC_SETTLE} ticks "
                                                     "from the profiled ticks' starting state; traffic_ratio "
                                                     "divides by the algorithmic bytes of the same launches"},
        "kernel_us": {"k_solve": round(solve_ms * 1e3, 2), "k_solve_span": round(span_ms * 1e3, 2),
                      "k_apply_span": round(apply_ms * 1e3, 2)},
        "cpu_baseline": cpu,
        "learned_warm_start": learned,
        "batched": batch,
        "articulated": art,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
