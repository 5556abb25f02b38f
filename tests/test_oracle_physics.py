"""Physics known-answer tests of the FP64 oracle (the reference's CpuSolver KATs, libs/sim/tests)."""
import numpy as np
import pytest

import physics_cases as pc


@pytest.fixture
def make(oracle):
    return oracle.OracleSolver


def test_free_fall(make):
    drop, closed, lateral = pc.free_fall(make)
    assert np.all(np.abs(drop / 1.22625 - 1.0) < 0.005)  # the reference's bound (test_cpu_solver.cpp:86)
    assert np.all(np.abs(drop / closed - 1.0) < 1e-3)    # implicit Euler closed form
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


def _mass(sc):
    g = sc.garments[0]
    return np.maximum(g.mesh.vertex_area * g.piece_fabric[0]["weight"], 1e-12)


def test_stats_closed_forms(oracle):
    """SimModel::stats (sim_model.cpp:222-241) on planted states of the 32x32 sheet (2 pins, sphere r 0.2 m at
    the origin) and a small seamed tube: kinetic energy and rms speed of unpinned particles moving at one velocity
    (pinned ones move faster and one holds a NaN, which only the non-finite flag may see), penetration of a
    particle 6 mm inside the sphere (trilinear SDF error at a 4 mm cell < 0.1 mm), and the gap of one seam
    opened to a 3-4-5 triangle."""
    from drape_amd import scenes
    sc = scenes.c1_sheet()
    o = oracle.OracleSolver()
    o.build(sc)
    n = o.particle_count()
    pinned = [p for p, _ in sc.pins]
    v = np.tile(np.float32([0.3, -0.4, 0.0]), (n, 1))
    v[pinned[0]] = 5.0
    v[pinned[1]] = np.nan
    x = o.read_positions()
    x[100] = (0.0, 0.2 - 0.006, 0.0)
    o.write_state(x=x, v=v)
    st = o.stats()
    v2 = float(np.float32(0.3)) ** 2 + float(np.float32(-0.4)) ** 2
    m = _mass(sc)
    assert st["has_non_finite"]
    assert st["rms_speed"] == pytest.approx(np.sqrt(v2), rel=1e-14)
    assert st["kinetic_energy"] == pytest.approx(0.5 * v2 * (m.sum() - m[pinned].sum()), rel=1e-12)
    assert st["max_body_penetration"] == pytest.approx(0.006, abs=1e-4)
    assert st["max_seam_gap"] == 0.0
    v[pinned[1]] = 0.0
    o.write_state(v=v)
    assert not o.stats()["has_non_finite"]

    tube = scenes.c3_tube(nx=12, ny=10)
    o = oracle.OracleSolver()
    o.build(tube)
    pairs = tube.garments[0].mesh.seam_pairs
    x = o.read_positions()
    x[pairs[:, 1]] = x[pairs[:, 0]]
    x[pairs[3, 0]] += np.float32([0.003, 0.004, 0.0])
    o.write_state(x=x)
    st = o.stats()
    assert st["max_seam_gap"] == pytest.approx(0.005, rel=1e-5)
    assert st["max_body_penetration"] == 0.0
    assert st["kinetic_energy"] == 0.0 and st["rms_speed"] == 0.0


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


@pytest.mark.parametrize("L,a,hA,b,hB,grain,bw,bf", [
    (0.01, 0.003, 0.008, 0.006, 0.012, 0.3, 2e-6, 1.5e-6),     # testJersey-like, grain off the edge
    (0.004, 0.001, 0.004, 0.002, 0.002, 1.2, 100e-6, 40e-6),  # anisotropic, grain near the edge normal
    (0.02, 0.01, 0.01, 0.01, 0.01, 0.0, 1e-5, 1e-5),          # isotropic, grain along the edge
    (0.01, 0.002, 0.007, 0.009, 0.005, 0.9, 1e-6, 50e-6),     # weft much stiffer: the 1/20 floor on p binds
])
def test_hinge_stiffness_closed_form(oracle, L, a, hA, b, hB, grain, bw, bf):
    """The oracle's hinge stiffness on one two-triangle hinge against the closed form from the heights and the
    grain angle (not from cross products or grain dot products)."""
    o = oracle.OracleSolver()
    o.build(pc.two_triangle_hinge(L, a, hA, b, hB, grain, bw, bf))
    k = o.terms()["hinge_k"]
    assert len(k) == 1
    assert k[0] == pytest.approx(pc.hinge_k_closed_form(L, hA, hB, grain, bw, bf), rel=1e-12)
