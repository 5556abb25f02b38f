"""Vertex normals with seam welding (vertexNormals, libs/render/src/normals.cpp:5-24): the oracle pinned by the
reference's own render tests (libs/render/tests/test_render_math.cpp:20-57) and closed forms, then the device
kernels bit for bit against the oracle (same FP32 operations in the same order)."""
import numpy as np
import pytest

from drape_amd import scenes


def _flat_grid():
    """3x3 grid in the XZ plane, triangles counter-clockwise seen from +Y (test_render_math.cpp:21-31)."""
    x = np.array([(0.1 * i, 0.0, -0.1 * j) for j in range(3) for i in range(3)], np.float32)
    tris = []
    for j in range(2):
        for i in range(2):
            a = j * 3 + i
            b, c = a + 1, a + 3
            tris += [(a, b, c + 1), (a, c + 1, c)]
    return x, np.array(tris, np.uint32)


def _welded_quads():
    """Quad A in XZ (normal +Y), quad B in XY (normal +Z), B's edge welded to A's (test_render_math.cpp:38-43)."""
    x = np.array([(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 0, 0), (1, 0, 0), (1, -1, 0), (0, -1, 0)], np.float32)
    tris = np.array([(0, 2, 1), (0, 3, 2), (4, 6, 5), (4, 7, 6)], np.uint32)
    weld = np.array([(0, 4), (1, 5)], np.uint32)
    return x, tris, weld


def _octahedron():
    x = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], np.float32)
    tris = []
    for sx, i in ((1, 0), (-1, 1)):
        for sy, j in ((1, 2), (-1, 3)):
            for sz, k in ((1, 4), (-1, 5)):
                tris.append((i, j, k) if sx * sy * sz > 0 else (i, k, j))  # outward winding
    return x, np.array(tris, np.uint32)


def _random_case(seed, n=500, nt=900, nw=120):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    tris = rng.integers(0, n, (nt, 3)).astype(np.uint32)
    tris[:5, 1] = tris[:5, 0]  # repeated vertices (degenerate triangles)
    weld = rng.integers(0, n, (nw, 2)).astype(np.uint32)
    weld[10:20, 0] = weld[0:10, 1]  # chains: a vertex in several pairs, order matters
    weld[30] = (weld[30, 0], weld[30, 0])  # a vertex welded to itself
    return x, tris, weld


def _check_kats(normals):
    x, t = _flat_grid()
    n = normals(x, t, None)
    assert np.abs(n - np.float32([0, 1, 0])).max() <= 1e-6
    x, t, w = _welded_quads()
    n = normals(x, t, w)
    for a, b in w:
        assert np.linalg.norm(n[a] - n[b]) <= 1e-6
        assert np.arccos(np.clip(n[a] @ np.float32([0, 1, 0]), -1, 1)) == pytest.approx(np.pi / 4, abs=1e-5)
        assert np.arccos(np.clip(n[a] @ np.float32([0, 0, 1]), -1, 1)) == pytest.approx(np.pi / 4, abs=1e-5)
    assert np.linalg.norm(n[2] - np.float32([0, 1, 0])) <= 1e-6
    x = np.zeros((3, 3), np.float32)
    n = normals(x, np.array([(0, 1, 2)], np.uint32), None)
    np.testing.assert_array_equal(n, np.tile(np.float32([0, 1, 0]), (3, 1)))


def test_oracle_normals_reference_kats_and_closed_forms(oracle):
    _check_kats(oracle.vertex_normals)
    x, t = _octahedron()  # by symmetry each vertex normal is its own direction
    np.testing.assert_allclose(oracle.vertex_normals(x, t), x, atol=1e-7)
    # Weld order: (0, 1) then (1, 2) gives n0 = a0 + a1 and n1 = n2 = a0 + a1 + a2 (before normalising).
    x = np.array([(0, 0, 0), (1, 0, 0), (0, 0, -1), (2, 0, 0), (3, 0, 0), (2, 1, 0), (4, 0, 0), (4, 0, 1),
                  (4, 1, 0)], np.float32)
    t = np.array([(0, 1, 2), (3, 4, 5), (6, 7, 8)], np.uint32)  # +Y, +Z, -X facing triangles
    n = oracle.vertex_normals(x, t, np.array([(0, 3), (3, 6)], np.uint32))
    ey, ez, ex = np.float32([0, 1, 0]), np.float32([0, 0, 1]), np.float32([-1, 0, 0])
    np.testing.assert_allclose(n[0], (ey + ez) / np.sqrt(2), atol=1e-7)
    np.testing.assert_allclose(n[3], (ey + ez + ex) / np.sqrt(3), atol=1e-7)
    np.testing.assert_allclose(n[6], n[3], atol=0)


@pytest.mark.gpu
def test_device_normals_reference_kats(hip):
    _check_kats(hip.vertex_normals)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_device_normals_match_oracle_bit_for_bit(hip, oracle, seed):
    x, t, w = _random_case(seed)
    a = hip.vertex_normals(x, t, w)
    b = oracle.vertex_normals(x, t, w)
    print(f"seed {seed}: max difference {np.abs(a - b).max():.3e}", flush=True)
    np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
def test_solver_positions_and_normals_match_oracle(hip, oracle):
    """The solver's live positions and normals (every garment's triangles, seam pairs welded) after a few steps of
    the 30x20 seamed tube equal the oracle's normals of the same positions, bit for bit."""
    sc = scenes.c3_tube(nx=30, ny=20)
    g = hip.HipSolver()
    g.build(sc)
    for _ in range(3):
        g.step(1 / 60, 1, 5)
    pos, nrm = g.read_positions_normals()
    np.testing.assert_array_equal(pos, g.read_positions())
    mesh = sc.garments[0].mesh
    weld = np.asarray(mesh.seam_pairs, np.uint32)[:, :2]
    ref = oracle.vertex_normals(pos, mesh.triangles, weld)
    np.testing.assert_array_equal(nrm, ref)
    assert np.abs(np.linalg.norm(nrm, axis=1) - 1).max() <= 1e-6
    with pytest.raises(IndexError):  # a triangle index past the vertex count: std::out_of_range
        hip.vertex_normals(pos[:10], mesh.triangles, None)
