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


# The reference's bending KATs: they pin the hinge stiffness K = sum over sides of 1/2 B(phi) |e|^2 / (A1 + A2)
# (sim_model.cpp:20-27, bending.cpp:7-18) and the seam-hinge rule (sim_model.cpp:147-175) end to end.
CANTILEVERS = {
    "isotropic": (lambda: pc.scenes.test_stiff(), (1.0, 0.0), 100e-6),  # CantileverIsotropic
    "aniso_warp": (pc.test_aniso, (1.0, 0.0), 100e-6),                  # CantileverAnisoWarp
    "aniso_weft": (pc.test_aniso, (0.0, 1.0), 40e-6),                   # CantileverAnisoWeft
}


@pytest.mark.parametrize("case", list(CANTILEVERS))
def test_cantilever_rigidity_within_10pct(make, case):
    """test_cantilever.cpp:66-74: Peirce's flexural rigidity from the settled droop within 10 % of B."""
    fab, grain, B = CANTILEVERS[case]
    frames, ratio, theta = pc.cantilever_ratio(make, fab(), grain, B)
    print(f"cantilever {case}: G/B {ratio:.3f}, theta {theta:.1f} deg, still after {frames} frames")
    assert frames > 0
    assert abs(ratio - 1.0) <= 0.10


def test_seam_does_not_hinge(make):
    """test_seam_hinge.cpp:38-59: a strip sewn across its middle droops like the uncut strip (0.8 .. 1.25)."""
    single, sewn, f1, f2 = pc.seam_hinge_drops(make)
    print(f"tip drop single {single:.4f} sewn {sewn:.4f} (frames {f1}, {f2})")
    assert f1 > 0 and f2 > 0
    assert 0.8 <= sewn / single <= 1.25


def test_hanging_sheet_settles_symmetric(make):
    """test_drape_runner.cpp:108-125: drapes to rest, centre of mass at x = 0.25 m, bottom corners level (2 mm)."""
    settled, com_x, y_left, y_right = pc.hanging_settles(make)
    print(f"hanging sheet: settled {settled}, com x {com_x:.5f}, bottom corners {y_left:.5f} / {y_right:.5f}")
    assert settled
    assert abs(com_x - 0.25) <= 0.002
    assert abs(y_left - y_right) <= 0.002


def test_hinge_stiffness_closed_form_on_device_terms(hip):
    """The HIP build's hinge stiffness (host model, uploaded terms) against the same closed form."""
    for L, a, hA, b, hB, grain, bw, bf in [(0.01, 0.003, 0.008, 0.006, 0.012, 0.3, 2e-6, 1.5e-6),
                                           (0.01, 0.002, 0.007, 0.009, 0.005, 0.9, 1e-6, 50e-6)]:
        g = hip.HipSolver()
        g.build(pc.two_triangle_hinge(L, a, hA, b, hB, grain, bw, bf))
        assert g.terms()["hinge_k"][0] == pytest.approx(pc.hinge_k_closed_form(L, hA, hB, grain, bw, bf), rel=1e-12)
