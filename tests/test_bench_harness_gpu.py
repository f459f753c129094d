"""The reference's benchmark harness on the GPU (mechanic-bench/src/main.rs):
the small scenarios' gates, and the four-bar's outcome against the oracle."""
import json

import numpy as np
import pytest

from mechanic_amd import bench_harness as H
from tests._oracle import Oracle

pytestmark = pytest.mark.gpu


def run_main(capsys, *args):
    rc = H.main(list(args))
    out = capsys.readouterr().out.strip().splitlines()
    return rc, json.loads(out[-1])


def test_smoke_gate_passes(capsys):
    rc, line = run_main(capsys, "--scenario", "smoke")
    print(line)
    assert rc == 0 and line["gate_passed"] and line["error_flags"] == 0
    assert line["bodies"] == 1024 and line["measured_ticks"] == 60 and line["contacts"] > 0
    assert line["p95_gpu_tick_ms"] is not None and line["physics_tps"] >= 60.0


def test_invalid_loop_raises_the_non_convergence_flag(capsys):
    rc, line = run_main(capsys, "--scenario", "invalid_loop")
    assert rc == 0 and line["gate_passed"]
    assert line["error_flags"] & H.CONSTRAINT_NON_CONVERGENCE_FLAG
    assert line["anchor_residual_m"] > 0.1


def test_four_bar_follows_the_fp32_oracle(capsys):
    """The four-bar's closure ends at the edge of the axis tolerance: the
    Newton evaluation's FP32 acos (closure.wgsl) reads the loop as closed while
    the validation's atan2 (bearings.wgsl) measures 0.0012 degrees, so the
    reference algorithm in FP32 raises the non-convergence flag; the GPU
    reports what the FP32 oracle does, and the FP64 oracle ends inside the
    tolerance."""
    rc, line = run_main(capsys, "--scenario", "four_bar")
    scene = H.build_scenario("four_bar").scene
    flags = {}
    for fp64 in (False, True):
        o = Oracle(scene, collisions=False, fp64=fp64)
        o.set_coordinates(np.array([[0.001, 0.0], [0.0, 0.0], [0.0, 0.0]], np.float32))
        f = 0
        for t in range(1, 61):
            o.step(t)
            f |= int(o.diag()[0])
        flags[fp64] = f
    assert line["error_flags"] == flags[False]
    assert flags[True] == 0
    assert line["anchor_residual_m"] <= 1.1e-5
    assert rc == (0 if line["error_flags"] == 0 else 2)
