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
