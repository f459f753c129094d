"""The reference's mechanism behaviour pins on the oracle (device.rs:3406-3922,
4476-4592, 4643-4847): pendulums that come to rest, drives, limits, the sleep
clamp, live reprogramming, impulses through a bearing, cylinder bores, a
contact-supported tower and a balanced child, each at the reference's own
thresholds. A plausible mistake in the oracle's articulated path (a dropped
damping factor, a wrong drive clamp or sign, a lost contact coupling) fails one
of them."""
import pytest

from tests import _mech_pins as M

SIM = M.OracleSim


@pytest.mark.parametrize("check", [
    M.check_tall_pendulums, M.check_branching_pendulums, M.check_coaxial_rotor, M.check_weak_drive_stalls,
    M.check_vertical_drive_escapes_sleep_clamp, M.check_external_impulse_drives_bearing,
    M.check_grounded_offset_pendulum, M.check_freely_falling_hinge, M.check_cylinder_bore,
    M.check_tower_drives_welded_arm, M.check_double_pendulum, M.check_balanced_child,
], ids=lambda f: f.__name__[6:])
def test_reference_pin(check):
    check(SIM)


@pytest.mark.parametrize("case", sorted(M.DRIVE_CASES))
def test_drive_pin(case):
    M.check_drive_case(SIM, case)


def test_live_reprogram_pin():
    M.check_live_reprogram(SIM, ValueError)
