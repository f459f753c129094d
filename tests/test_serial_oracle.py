"""The serial small-mechanism solvers of the oracle (device.rs:19-20,
1370-1416; collision.wgsl:979-1048; articulated.wgsl:216-285), pinned by the
reference's own articulated-car tests (device.rs:4224-4473) with their
thresholds. The car has eight bearings and a dynamic root, so it takes the
serial sweeps; run without them (the oracle's serial limit set to 0, which
plants the parallel Jacobi passes in their place) the same scenes sink into the
ground, so the pins fail without the serial path."""
import pytest

from tests._car import (OracleSim, assert_drop, assert_struck, assert_wall, car_fixture, drop, struck_wheel,
                        wall_impact)
from tests._oracle import Oracle


@pytest.fixture(scope="module")
def car():
    return car_fixture()


def oracle_for(car, serial_limit=64):
    o = Oracle(car.compiled.scene, iterations=8, self_collisions=False)
    o.set_serial_limit(serial_limit)
    return OracleSim(o)


def test_car_is_a_serial_mechanism(car):
    assert len(car.compiled.bearings) == 8 <= 64
    assert car.compiled.scene.body_count == 10
    assert not car.compiled.compounds[car.chassis].is_static


def test_car_drop_settles_without_drift_or_ground_penetration(car):
    assert_drop(drop(oracle_for(car), car))


def test_struck_wheel_recovers_without_crossing_ground(car):
    assert_struck(struck_wheel(oracle_for(car), car))


def test_wall_impact_remains_bounded_and_decays(car):
    assert_wall(wall_impact(oracle_for(car), car))


@pytest.mark.parametrize("scenario,check", [(drop, assert_drop), (struck_wheel, assert_struck),
                                            (wall_impact, assert_wall)])
def test_pins_fail_without_the_serial_sweeps(car, scenario, check):
    with pytest.raises(AssertionError):
        check(scenario(oracle_for(car, serial_limit=0), car))
