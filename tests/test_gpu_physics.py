"""The oracle's physics known-answer tests run on the HIP backend (backend parameterisation,
libs/sim/tests/backends.hpp:14-47), same bounds."""
import numpy as np
import pytest

import physics_cases as pc

pytestmark = pytest.mark.gpu


@pytest.fixture
def make(hip):
    return hip.HipSolver


def test_free_fall(make):
    drop, closed, lateral = pc.free_fall(make)
    assert np.all(np.abs(drop / 1.22625 - 1.0) < 0.005)
    assert np.all(np.abs(drop / closed - 1.0) < 1e-3)
    assert lateral < 1e-5


def test_pins_hold(make):
    assert pc.pins_hold(make) < 1e-7


def test_assembling_disables_gravity(make):
    assert pc.assembling_still(make) < 1e-6


@pytest.mark.parametrize("grain,expect", [((1.0, 0.0), 0.025), ((0.0, 1.0), 0.05)])
def test_strip_strain(make, grain, expect):
    frames, strain = pc.strip_strain(make, grain)
    assert frames > 0
    assert strain == pytest.approx(expect, rel=0.02)


def test_flat_sheet_stays_flat(make):
    assert pc.flat_stays_flat(make) < 1e-6


def test_rests_on_sphere(make):
    phi = pc.rests_on_sphere(make)
    assert np.all(np.abs(phi - pc.K_CONTACT) < 5e-4)


def test_friction_holds_and_slides(make):
    assert abs(pc.slope_slide(make, 0.5)) < 0.002
    assert pc.slope_slide(make, 0.2) < -0.05


def test_seams_close_at_closing_speed(make):
    start, mid, end = pc.seams_close(make)
    assert start == pytest.approx(0.1, abs=1e-6)
    assert 0.047 <= mid <= 0.053
    assert end < 0.001


def test_snapshot_restore_is_deterministic_in_contact(make):
    a, b = pc.snapshot_replay(make)
    np.testing.assert_array_equal(a, b)
