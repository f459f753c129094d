"""The reference's mechanism behaviour pins on the GPU through the C ABI
(device.rs:3406-3922, 4476-4592, 4643-4847), at the reference's own
thresholds, and each pin scene against the oracle: bit for bit while every
joint angle is still exactly zero (forward kinematics then evaluates sin and
cos only at zero, where the device and host libraries agree), within the
articulated tolerance once the joints have moved."""
import numpy as np
import pytest

import mechanic_amd as mh
from tests import _mech_pins as M

pytestmark = pytest.mark.gpu

SIM = M.GpuSim


@pytest.mark.parametrize("check", [
    M.check_tall_pendulums, M.check_branching_pendulums, M.check_coaxial_rotor, M.check_weak_drive_stalls,
    M.check_vertical_drive_escapes_sleep_clamp, M.check_external_impulse_drives_bearing,
    M.check_grounded_offset_pendulum, M.check_freely_falling_hinge, M.check_cylinder_bore,
    M.check_tower_drives_welded_arm, M.check_double_pendulum, M.check_balanced_child,
], ids=lambda f: f.__name__[6:])
def test_reference_pin(check):
    print(f"[pins] {check.__name__}", flush=True)
    check(SIM)


@pytest.mark.parametrize("case", sorted(M.DRIVE_CASES))
def test_drive_pin(case):
    print(f"[pins] drive {case}", flush=True)
    M.check_drive_case(SIM, case)


def test_live_reprogram_pin():
    refused, _ = M.check_live_reprogram(SIM, mh.GpuPhysicsError)
    assert refused.kind == "DriveStateCount" and refused.provided == 0 and refused.required == 1


@pytest.mark.parametrize("name", sorted(M.SCENES))
def test_pin_scene_matches_oracle(name):
    print(f"[pins] scene {name}", flush=True)
    compiled = M.SCENES[name]()
    gpu, ora = M.GpuSim(compiled), M.OracleSim(compiled)
    for tick in (1, 30):
        gpu.run_to(tick)
        ora.run_to(tick)
        g, o = gpu.poses(), ora.poses()
        assert np.all(np.isfinite(g))
        tolerance = 2.0e-5 if tick == 1 else 1.0e-3
        np.testing.assert_allclose(g, o, atol=tolerance, err_msg=f"{name} tick {tick}")
        assert gpu.flags() == ora.flags(), (name, tick)
        assert gpu.active_contacts() == ora.active_contacts(), (name, tick)
    gpu.close()
