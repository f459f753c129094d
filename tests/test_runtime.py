"""Lifecycle and clock: the reference's own runtime and scheduler tests
(crates/mechanic-gpu/src/runtime.rs:283-297, scheduler.rs:86-107)."""
import pytest

import mechanic_amd as mh
from mechanic_amd import runtime as rt


def test_snapshots_advance_only_for_running_successful_ticks():
    runtime = rt.PhysicsRuntime()
    runtime.load(mh.boxes([[0, 1, 0]], [0.5, 0.5, 0.5]))
    runtime.publish_tick(rt.TickStatistics())
    assert runtime.published_state().current_tick == 0
    runtime.start()
    runtime.publish_tick(rt.TickStatistics())
    runtime.publish_tick(rt.TickStatistics())
    p = runtime.published_state()
    assert (p.previous_tick, p.current_tick) == (1, 2)
    assert (p.previous_slot, p.current_slot) == (1, 2)
    assert runtime.status().state == rt.SimulationState.RUNNING


def test_one_second_always_produces_exactly_sixty_ticks():
    s = rt.FixedStepScheduler()
    ticks = s.advance(1_000_000_000)
    assert ticks.count == 60
    assert list(ticks) == list(range(1, 61))


def test_render_sized_fragments_do_not_drop_or_duplicate_ticks():
    s = rt.FixedStepScheduler()
    emitted = []
    for _ in range(144):
        emitted.extend(s.advance(6_944_444))
    emitted.extend(s.advance(64))
    assert emitted == list(range(1, 61))


def test_failure_transitions():
    runtime = rt.PhysicsRuntime()
    with pytest.raises(rt.RuntimeError_):
        runtime.start()  # nothing loaded
    runtime.load(mh.boxes([[0, 1, 0]], [0.5, 0.5, 0.5]))
    runtime.start()
    runtime.fail(rt.InvalidNumericState())
    assert runtime.failure() == rt.InvalidNumericState()
    with pytest.raises(rt.RuntimeError_):
        runtime.start()
    runtime.pause()
    assert runtime.status().state == rt.SimulationState.BUILDING and runtime.published_state() is None


def test_capacity_failure_on_load():
    class Big:
        body_count = mh.MAX_BODIES + 1

    runtime = rt.PhysicsRuntime()
    with pytest.raises(rt.RuntimeError_) as err:
        runtime.load(Big())
    assert err.value.failure == rt.CapacityOverflow(rt.CapacityKind.BODIES, mh.MAX_BODIES + 1, mh.MAX_BODIES)
    assert runtime.status().state == rt.SimulationState.FAILED


def test_flag_mapping():
    assert rt.failure_from_flags(0) is None
    assert rt.failure_from_flags(mh.INVALID_NUMERIC_FLAG) == rt.InvalidNumericState()
    assert isinstance(rt.failure_from_flags(mh.PAIR_OVERFLOW_FLAG), rt.CapacityOverflow)
    assert isinstance(rt.failure_from_flags(mh.CONSTRAINT_NON_CONVERGENCE_FLAG), rt.ConstraintNonConvergence)
