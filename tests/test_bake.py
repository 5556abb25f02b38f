"""SDF bake of closed triangle meshes (bakeCollisionField, libs/avatar/src/collision_bake.cpp:98-208): the oracle
pinned by the reference's own field tests (libs/avatar/tests/test_collision_field.cpp:34-100), then the device bake
against the oracle value for value."""
import time

import numpy as np
import pytest

from drape_amd import bodies


def _random_directions(rng, n):
    d = rng.normal(0, 1, (n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def _check_field_kats(bake, oracle):
    """test_collision_field.cpp: SphereDistanceAccuracy, SignInsideOutside, GradientPointsOutward, UnionIsMin,
    OutsideGridIsFarOutside, RejectsOpenMesh (the last on the given bake function)."""
    f = bake([bodies.icosphere((0, 0, 0), 0.1, 4)])
    rng = np.random.default_rng(7)
    p = (0.1 + rng.uniform(-0.015, 0.015, (2000, 1))) * _random_directions(rng, 2000)
    phi, _ = oracle.field_query(f, p)
    err = np.abs(phi - (np.linalg.norm(p, axis=1) - 0.1))
    assert err.max() <= 0.001, err.max()
    phi, _ = oracle.field_query(f, np.array([(0, 0, 0), (0.15, 0, 0)], float))
    assert phi[0] == pytest.approx(-0.02, abs=1e-6) and phi[1] == pytest.approx(0.02, abs=1e-6)
    d = _random_directions(np.random.default_rng(11), 200)
    _, g = oracle.field_query(f, 0.105 * d)
    assert np.abs(np.linalg.norm(g, axis=1) - 1).max() <= 1e-9
    assert np.degrees(np.arccos(np.clip(np.sum(g * d, axis=1), -1, 1))).max() <= 5.0
    u = bake([bodies.icosphere((0.08, 0, 0), 0.1, 4), bodies.icosphere((-0.08, 0, 0), 0.1, 4)])
    phi, _ = oracle.field_query(u, np.array([(0, 0, 0), (0.08, 0, 0.115), (0, 0, 0.085)], float))
    expected = np.linalg.norm(np.array([0, 0, 0.085]) - np.array([0.08, 0, 0])) - 0.1
    assert phi[0] < 0.0
    assert phi[1] == pytest.approx(0.015, abs=0.001)
    assert 0.0 < phi[2] <= expected + 1e-6 and phi[2] == pytest.approx(expected, abs=0.002)
    lo = np.asarray(f.origin)
    hi = lo + f.cell_size * (np.asarray(f.dims) - 1)
    probes = [(10.0, 10.0, 10.0)]
    for axis in range(3):
        below, above = 0.5 * (lo + hi), 0.5 * (lo + hi)
        below[axis] = lo[axis] - 1e-6
        above[axis] = hi[axis] + 1e-6
        probes += [tuple(below), tuple(above)]
    phi, g = oracle.field_query(f, np.array(probes))
    assert np.all(phi == f.band) and np.all(g[0] == 0)  # zero gradient far outside (:79)
    v, t = bodies.icosphere((0, 0, 0), 0.1, 2)
    with pytest.raises(ValueError):
        bake([(v, t[:-1])])


def _mannequin():
    """Capsule-union body: torso and two arms (the shape of the C3 body, as closed meshes)."""
    c = np.cos(np.pi / 4)
    return [bodies.capsule((0, -0.275, 0), (0, 0.275, 0), 0.15, 48, 12),
            bodies.capsule((0.15, 0.275, 0), (0.15 + 0.5 * c, 0.275 - 0.5 * c, 0), 0.05, 24, 6),
            bodies.capsule((-0.15, 0.275, 0), (-0.15 - 0.5 * c, 0.275 - 0.5 * c, 0), 0.05, 24, 6)]


def test_body_meshes_are_closed_and_outward():
    for v, t in [bodies.icosphere((0, 0, 0), 0.1, 3)] + _mannequin():
        assert bodies.is_closed_manifold(t)
        assert bodies.signed_volume(v, t) > 0
    v, t = bodies.icosphere((0, 0, 0), 0.1, 2)
    assert not bodies.is_closed_manifold(t[:-1])


def test_oracle_bake_reference_kats(oracle):
    _check_field_kats(oracle.bake_collision_field, oracle)


def test_oracle_bake_capsule_closed_form(oracle):
    """Distances to a capsule mesh (48 sides) within the faceting bound: r (1 - cos(pi / 48)) < 0.33 mm, plus
    the grid's trilinear error."""
    v, t = bodies.capsule((0, -0.2, 0), (0, 0.2, 0), 0.1, 48, 12)
    f = oracle.bake_collision_field([(v, t)], cell_size=0.004)
    rng = np.random.default_rng(3)
    p = rng.uniform(-0.15, 0.15, (3000, 3)) * np.array([1, 2.5, 1])
    y = np.clip(p[:, 1], -0.2, 0.2)
    exact = np.linalg.norm(p - np.stack([np.zeros_like(y), y, np.zeros_like(y)], axis=1), axis=1) - 0.1
    keep = np.abs(exact) < 0.015
    phi, _ = oracle.field_query(f, p[keep])
    assert keep.sum() > 200
    assert np.abs(phi - exact[keep]).max() <= 0.001


@pytest.mark.gpu
def test_device_bake_reference_kats(hip, oracle):
    _check_field_kats(hip.bake_collision_field, oracle)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sphere", "union", "mannequin"])
def test_device_bake_matches_oracle(hip, oracle, case):
    """The device bake equals the oracle's sequential bake sample for sample (same FP64 operations, same nearest
    triangle and sign, same flood fill)."""
    comps = {"sphere": [bodies.icosphere((0.01, -0.02, 0.003), 0.1, 3)],
             "union": [bodies.icosphere((0.07, 0, 0), 0.1, 3), bodies.icosphere((-0.07, 0.01, 0), 0.08, 3)],
             "mannequin": _mannequin()}[case]
    cell = 0.008 if case == "mannequin" else 0.004
    a = hip.bake_collision_field(comps, cell_size=cell)
    b = oracle.bake_collision_field(comps, cell_size=cell)
    assert a.dims == b.dims and a.origin == b.origin
    diff = np.abs(a.values - b.values)
    print(f"{case}: dims {a.dims}, max difference {diff.max():.3e}, samples inside {(a.values < 0).sum()}",
          flush=True)
    np.testing.assert_array_equal(a.values, b.values)


@pytest.mark.gpu
def test_device_bake_of_the_mannequin_at_4mm_under_two_seconds(hip):
    """The reference's bake budget (test_collision_perf.cpp:20-30: under 2 s) for a capsule mannequin at the
    default 4 mm cell."""
    comps = _mannequin()
    hip.bake_collision_field([bodies.icosphere((0, 0, 0), 0.05, 1)])  # load the kernels
    t0 = time.perf_counter()
    f = hip.bake_collision_field(comps)
    dt = time.perf_counter() - t0
    print(f"mannequin bake: dims {f.dims} ({np.prod(f.dims) / 1e6:.1f} M samples), "
          f"{sum(len(t) for _, t in comps)} triangles, {dt * 1e3:.1f} ms", flush=True)
    assert dt < 2.0
    assert f.values.min() == pytest.approx(-0.02) and f.values.max() == pytest.approx(0.02)
