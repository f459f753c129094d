"""Host-side publication lifecycle and the fixed 60 Hz clock.

`PhysicsRuntime` mirrors crates/mechanic-gpu/src/runtime.rs:94-257 (status
machine, 3-slot snapshot publication, failure states) and `FixedStepScheduler`
mirrors crates/mechanic-gpu/src/scheduler.rs:41-76 (rational accumulator that
never drops or duplicates a tick). `PhysicsDriver` ties them to a `GpuPhysics`
handle the way the app's fixed-update system does (crates/mechanic-app/src/
main.rs:570-657): dispatch every scheduled tick, read diagnostics, stop on a
non-zero error flag, publish otherwise.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field, replace
from typing import Iterator, Optional, Union

from . import _lib

PHYSICS_TPS = 60  # lib.rs:30
_PHASE_DENOMINATOR = 1_000_000_000  # scheduler.rs:5


class CapacityKind(enum.Enum):
    """runtime.rs:6-14."""

    BODIES = "Bodies"
    BEARINGS = "Bearings"
    CONTACT_PAIRS = "ContactPairs"


@dataclass(frozen=True)
class InvalidTopology:
    message: str


@dataclass(frozen=True)
class ConstraintNonConvergence:
    bearing: Optional[int]
    anchor_residual_meters: float
    axis_residual_degrees: float


@dataclass(frozen=True)
class CapacityOverflow:
    kind: CapacityKind
    required: int
    capacity: int


@dataclass(frozen=True)
class InvalidNumericState:
    pass


@dataclass(frozen=True)
class DeviceFailure:
    message: str


FailureStatus = Union[InvalidTopology, ConstraintNonConvergence, CapacityOverflow, InvalidNumericState,
                      DeviceFailure]  # runtime.rs:17-43


class SimulationState(enum.Enum):
    BUILDING = "Building"
    PAUSED = "Paused"
    RUNNING = "Running"
    FAILED = "Failed"


@dataclass(frozen=True)
class SimulationStatus:
    """runtime.rs:46-56; `failure` is set only for FAILED."""

    state: SimulationState
    failure: Optional[FailureStatus] = None


@dataclass(frozen=True)
class PublishedGpuState:
    """runtime.rs:59-73."""

    generation: int
    previous_slot: int
    current_slot: int
    previous_tick: int
    current_tick: int
    body_count: int


@dataclass(frozen=True)
class TickStatistics:
    """runtime.rs:76-92."""

    tick_index: int = 0
    physics_tps: float = 0.0
    gpu_tick_ms: float = 0.0
    body_count: int = 0
    contact_count: int = 0
    anchor_residual_meters: float = 0.0
    axis_residual_degrees: float = 0.0


class RuntimeError_(Exception):
    """Raised with a FailureStatus by load/start, like the reference's Err."""

    def __init__(self, failure: FailureStatus):
        super().__init__(repr(failure))
        self.failure = failure


class PhysicsRuntime:
    """runtime.rs:95-257. `scene` stands for the CompiledCreation."""

    def __init__(self):
        self._status = SimulationStatus(SimulationState.BUILDING)
        self._scene = None
        self._statistics = TickStatistics()
        self._published: Optional[PublishedGpuState] = None
        self._generation = 0

    def load(self, scene) -> None:
        bodies = scene.body_count
        bearings = getattr(scene, "bearing_count", 0)
        if bodies > _lib.MAX_BODIES:
            self._fail_capacity(CapacityKind.BODIES, bodies, _lib.MAX_BODIES)
        if bearings > _lib.MAX_BEARINGS:
            self._fail_capacity(CapacityKind.BEARINGS, bearings, _lib.MAX_BEARINGS)
        self._generation = (self._generation + 1) & 0xFFFFFFFFFFFFFFFF
        self._scene = scene
        self._statistics = TickStatistics(body_count=bodies)
        self._published = PublishedGpuState(self._generation, 0, 0, 0, 0, bodies)
        self._status = SimulationStatus(SimulationState.PAUSED)

    def start(self) -> None:
        if self._status.state == SimulationState.PAUSED:
            self._status = SimulationStatus(SimulationState.RUNNING)
            return
        if self._status.state == SimulationState.FAILED:
            raise RuntimeError_(self._status.failure)
        raise RuntimeError_(InvalidTopology("no paused compiled scene is ready"))

    def pause(self) -> None:
        if self._status.state == SimulationState.FAILED:
            self._scene = None
            self._published = None
            self._status = SimulationStatus(SimulationState.BUILDING)
        else:
            self._status = SimulationStatus(
                SimulationState.PAUSED if self._scene is not None else SimulationState.BUILDING)

    def status(self) -> SimulationStatus:
        return self._status

    def failure(self) -> Optional[FailureStatus]:
        return self._status.failure if self._status.state == SimulationState.FAILED else None

    def tick_statistics(self) -> TickStatistics:
        return self._statistics

    def published_state(self) -> Optional[PublishedGpuState]:
        return self._published

    def creation(self):
        return self._scene

    def publish_tick(self, statistics: TickStatistics) -> None:
        """Advances the ring only while Running (runtime.rs:221-236)."""
        if self._status.state != SimulationState.RUNNING:
            return
        next_tick = (self._statistics.tick_index + 1) & 0xFFFFFFFFFFFFFFFF
        self._statistics = replace(statistics, tick_index=next_tick, body_count=self._statistics.body_count)
        if self._published is not None:
            p = self._published
            self._published = PublishedGpuState(p.generation, p.current_slot, next_tick % 3, p.current_tick,
                                                next_tick, p.body_count)

    def fail(self, failure: FailureStatus) -> None:
        self._status = SimulationStatus(SimulationState.FAILED, failure)

    def _fail_capacity(self, kind: CapacityKind, required: int, capacity: int) -> None:
        failure = CapacityOverflow(kind, required, capacity)
        self._status = SimulationStatus(SimulationState.FAILED, failure)
        raise RuntimeError_(failure)


@dataclass(frozen=True)
class ScheduledTicks:
    """scheduler.rs:7-31: `count` ticks starting at `first`."""

    first_tick: int
    count: int

    def first(self) -> Optional[int]:
        return None if self.count == 0 else self.first_tick

    def __iter__(self) -> Iterator[int]:
        return iter(range(self.first_tick, self.first_tick + self.count))


@dataclass
class FixedStepScheduler:
    """scheduler.rs:41-76: elapsed nanoseconds x 60 accumulate as a rational
    phase; every whole 10^9 is one tick."""

    phase: int = 0
    next_tick_index: int = 1

    def advance(self, elapsed_nanos: int) -> ScheduledTicks:
        self.phase += int(elapsed_nanos) * PHYSICS_TPS
        due, self.phase = divmod(self.phase, _PHASE_DENOMINATOR)
        first = self.next_tick_index
        self.next_tick_index += due
        return ScheduledTicks(first, due)

    def interpolation_alpha(self) -> float:
        return min(self.phase, 0xFFFFFFFF) / 1_000_000_000.0

    def next_tick(self) -> int:
        return self.next_tick_index


def failure_from_flags(error_flags: int, anchor_m: float = 0.0, axis_deg: float = 0.0) -> Optional[FailureStatus]:
    """Maps GpuDiagnostics.error_flags (abi.rs:3-13) to the failure the caller
    records; the app stops on any non-zero flag (main.rs:617-632)."""
    if error_flags == 0:
        return None
    if error_flags & _lib.INVALID_NUMERIC_FLAG:
        return InvalidNumericState()
    if error_flags & _lib.CONSTRAINT_NON_CONVERGENCE_FLAG:
        return ConstraintNonConvergence(None, anchor_m, axis_deg)
    if error_flags & (_lib.PAIR_OVERFLOW_FLAG | _lib.MANIFOLD_OVERFLOW_FLAG):
        return CapacityOverflow(CapacityKind.CONTACT_PAIRS, 0, _lib.MAX_CONTACT_PAIRS)
    return DeviceFailure(f"physics kernel reported flags {error_flags}")


@dataclass
class PhysicsDriver:
    """Runs scheduled ticks on a GpuPhysics handle and publishes them."""

    gpu: object
    runtime: PhysicsRuntime
    scheduler: FixedStepScheduler = field(default_factory=FixedStepScheduler)

    def update(self, elapsed_nanos: int) -> int:
        """Dispatches the ticks due after `elapsed_nanos`; returns how many ran."""
        ran = 0
        for tick in self.scheduler.advance(elapsed_nanos):
            if self.runtime.status().state != SimulationState.RUNNING:
                break
            self.gpu.dispatch_tick(tick)
            readback = self.gpu.read_last_tick()
            failure = failure_from_flags(readback.error_flags, readback.anchor_residual_meters,
                                         readback.axis_residual_degrees)
            if failure is not None:
                self.runtime.fail(failure)
                break
            self.runtime.publish_tick(TickStatistics(
                gpu_tick_ms=readback.gpu_tick_ms or 0.0, contact_count=readback.contact_count,
                anchor_residual_meters=readback.anchor_residual_meters,
                axis_residual_degrees=readback.axis_residual_degrees))
            ran += 1
        return ran
