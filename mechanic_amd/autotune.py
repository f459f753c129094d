"""Autotuner for the launch shape of a handle's kernels (mh_tuning).

Coordinate descent over the tunable fields: starting from the defaults, each
field in turn is swept over its candidates with the others fixed, and the
fastest value kept. Every trial steps a fresh handle of the same scene (so all
trials start from the same state) through `warm` ticks -- which include the
graph capture -- and then times `ticks` ticks twice, keeping the faster
window. Results never depend on the tuning (tests/test_batch_gpu.py), only the
time does.

Winners are cached per (device architecture, scene-size bucket) in a JSON file,
the bucket being the power of two at or below the body count.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time
from typing import Callable, Dict, Optional

from . import _lib

SPACE = {
    "sort_kind": [0, 1],
    "traverse_kind": [0, 1, 2],
    "solve_kind": [0, 1],
    "sort_tile": [512, 1024, 2048],
    "sort_levels": [2, 3, 4],
    "narrow_blocks": [512, 1024, 2048, 4096],
    "prepare_blocks": [512, 1024, 2048, 4096],
    "solve_blocks": [1024, 2048, 4096, 8192],
    "scan_blocks": [1024, 2048, 4096, 8192],
}
DEFAULTS = {"sort_kind": 0, "traverse_kind": 0, "solve_kind": 0, "sort_tile": 2048, "sort_levels": 4,
            "narrow_blocks": 1024, "prepare_blocks": 1024, "solve_blocks": 4096, "scan_blocks": 4096}
CACHE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")


def device_arch() -> str:
    buf = ctypes.create_string_buffer(128)
    if _lib.load().mh_device_arch(buf, 128) != _lib.MH_OK:
        return "unknown"
    return buf.value.decode().split(":")[0]


def size_bucket(bodies: int) -> int:
    return 1 << int(math.floor(math.log2(max(bodies, 1))))


def cache_key(arch: str, bodies: int) -> str:
    return f"{arch}/{size_bucket(bodies)}"


def load_cache(path: Optional[str] = None) -> Dict[str, dict]:
    path = path or os.path.join(CACHE_DIR, "cache.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_cache(entries: Dict[str, dict], path: str) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as f:
        json.dump(entries, f, indent=1, sort_keys=True)


def cached_tuning(bodies: int, path: Optional[str] = None) -> Optional[dict]:
    entry = load_cache(path).get(cache_key(device_arch(), bodies))
    return None if entry is None else dict(entry["tuning"])


def measure(make_handle: Callable, tuning: dict, warm: int = 20, ticks: int = 40, state=None) -> float:
    """Milliseconds per tick of a fresh handle with `tuning`, started from
    `state` = (poses, velocities) when given (contacts re-form in `warm` ticks)."""
    gpu = make_handle()
    try:
        gpu.set_tuning(**tuning)
        if state is not None:
            gpu.write_bodies(state[0], state[1])
        gpu.dispatch_ticks(1, warm)
        gpu.synchronize()
        best = float("inf")
        first = warm + 1
        for _ in range(2):
            t0 = time.perf_counter()
            gpu.dispatch_ticks(first, ticks)
            gpu.synchronize()
            best = min(best, (time.perf_counter() - t0) * 1e3 / ticks)
            first += ticks
        return best
    finally:
        gpu.close()


def autotune(make_handle: Callable, warm: int = 20, ticks: int = 40, space: Optional[dict] = None, state=None,
             candidates=()) -> dict:
    """Coordinate descent from the defaults, every trial from `state` when given
    (the scene as it is when it is measured); then the winner, the defaults and
    any extra `candidates` are measured again over twice the ticks, in two
    alternating rounds, and the fastest is kept. Returns it, its time, the default's time and every trial."""
    space = space or SPACE
    current = dict(DEFAULTS)
    trials = []
    base_ms = measure(make_handle, current, warm, ticks, state)
    trials.append((dict(current), base_ms))
    best_ms = base_ms
    for field, values in space.items():
        for value in values:
            if value == current[field]:
                continue
            cand = dict(current, **{field: value})
            ms = measure(make_handle, cand, warm, ticks, state)
            trials.append((cand, ms))
            if ms < best_ms:
                best_ms, current = ms, cand
    finalists = []
    for cand in [current, dict(DEFAULTS)] + [dict(DEFAULTS, **c) for c in candidates if c]:
        if cand not in finalists:
            finalists.append(cand)
    # Two alternating rounds over fresh handles, so that drift between handles
    # and over time does not favour whichever finalist happens to run first.
    best = [float("inf")] * len(finalists)
    for _ in range(2):
        for i, cand in enumerate(finalists):
            best[i] = min(best[i], measure(make_handle, cand, warm, 2 * ticks, state))
    confirmed = list(zip(finalists, best))
    current, best_ms = min(confirmed, key=lambda x: x[1])
    default_ms = next(ms for c, ms in confirmed if c == dict(DEFAULTS))
    return {"tuning": current, "ms": best_ms, "default_ms": default_ms, "trials": trials + confirmed}
