"""The reference's benchmark harness restated (mechanic-bench/src/main.rs):
scenario rows as its own tests check them (main.rs:556-580), the JSONL field
order of main.rs:207-272, and the percentile rule of main.rs:537-540."""
import json

import pytest

from mechanic_amd import bench_harness as H

# The keys of the reference's JSONL line, in order (main.rs:209-229).
REFERENCE_KEYS = [
    "type", "scenario", "adapter", "backend", "bodies", "colliders", "bearings", "warmup_ticks",
    "measured_ticks", "construction_ms", "mean_engine_tick_ms", "p95_engine_tick_ms", "p95_gpu_tick_ms",
    "kernel_pipeline_p95_ms", "physics_tps", "kernel_integration_p95_ms", "kernel_mechanism_p95_ms",
    "kernel_broadphase_p95_ms", "kernel_narrowphase_p95_ms", "kernel_contact_solver_p95_ms",
    "kernel_bearings_p95_ms", "kernel_snapshot_p95_ms", "pairs", "contacts", "active_contacts",
    "anchor_residual_m", "axis_residual_deg", "error_flags", "timing_source", "kernel_coverage_complete",
    "correctness_passed", "budget_passed", "gate_passed"]


def test_smoke_scene_has_expected_rows():
    out = H.build_scenario("smoke")
    assert out.scene.body_count == 1024 and len(out.scene.colliders) == 1024
    assert out.scene.bearing_count == 0


@pytest.mark.parametrize("scenario", ["four_bar", "invalid_loop"])
def test_four_bar_scenarios_have_one_closure(scenario):
    out = H.build_scenario(scenario)
    assert out.scene.body_count == 4
    assert len(out.tree_bearings) == 3 and len(out.closure_bearings) == 1


def test_invalid_loop_moves_the_closure_anchor():
    good, bad = H.build_scenario("four_bar").scene, H.build_scenario("invalid_loop").scene
    k = next(i for i in range(good.bearing_count) if good.bearings[i].metadata[3])
    delta = [b - a for a, b in zip(list(good.bearings[k].local_anchor_b)[:3], list(bad.bearings[k].local_anchor_b)[:3])]
    assert delta == pytest.approx([0.2, 0.13, 0.07], abs=1e-6)


def test_dense_lattice_positions():
    out = H.build_dense(2001)
    p = out.scene.pose_array()
    # index 0 at (0, 0.5, 0); index 51 at x = 1, y = 1.5; index 2000 at z = 1 (main.rs:416-419).
    assert tuple(p[0, :3]) == pytest.approx((0.0, 0.5, 0.0))
    assert tuple(p[51, :3]) == pytest.approx((1.0, 1.5, 0.0))
    assert tuple(p[2000, :3]) == pytest.approx((0.0, 0.5, 1.0))


def test_percentile_rule():
    assert H.percentile_95([1.0]) == 1.0
    samples = sorted(float(i) for i in range(1, 101))
    assert H.percentile_95(samples) == 95.0
    assert H.percentile_95(samples[:60]) == 57.0  # rank ceil(0.95 * 60) = 57


def test_jsonl_line_fields(monkeypatch):
    """The line the harness prints parses as JSON with the reference's keys in order."""
    line = ('{"type":"benchmark","scenario":"smoke","adapter":"gfx950","backend":"Hip","bodies":1,"colliders":1,'
            '"bearings":0,"warmup_ticks":0,"measured_ticks":60,"construction_ms":1.000,"mean_engine_tick_ms":1.000,'
            '"p95_engine_tick_ms":1.000,"p95_gpu_tick_ms":null,"kernel_pipeline_p95_ms":null,"physics_tps":1.00,'
            '"kernel_integration_p95_ms":null,"kernel_mechanism_p95_ms":null,"kernel_broadphase_p95_ms":null,'
            '"kernel_narrowphase_p95_ms":null,"kernel_contact_solver_p95_ms":null,"kernel_bearings_p95_ms":null,'
            '"kernel_snapshot_p95_ms":null,"pairs":0,"contacts":0,"active_contacts":0,"anchor_residual_m":0.00000000,'
            '"axis_residual_deg":0.00000000,"error_flags":0,"timing_source":"gpu_timestamp",'
            '"kernel_coverage_complete":true,"correctness_passed":true,"budget_passed":false,"gate_passed":false}')
    assert list(json.loads(line).keys()) == REFERENCE_KEYS
    import inspect
    src = inspect.getsource(H.run)
    # Every reference key appears in the harness's format string, in order.
    positions = [src.index(f'"{k}":') if f'"{k}":' in src else src.index(f'\\"{k}\\":') for k in REFERENCE_KEYS]
    assert positions == sorted(positions)
