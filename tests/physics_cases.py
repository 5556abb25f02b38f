"""Physics known-answer scenes shared by the CPU (oracle) and GPU (HIP) test modules.

Each function takes a solver factory, so one scene runs on either backend — the reference's backend-parameterised
tests (libs/sim/tests/backends.hpp:14-47). Scenes are regular grids built here (the reference meshes its patterns
with CDT; the answers below do not depend on the mesh).
"""
import math

import numpy as np

from drape_amd import scenes

G = 9.81


def _flat(gid, nx, ny, spacing, fab, origin=(0.0, 0.0, 0.0), grain=(1.0, 0.0)):
    return scenes.flat_garment(gid, scenes.grid_mesh(nx, ny, spacing, grain), fab, origin)


def free_fall(make):
    """test_cpu_solver.cpp:78-93: 30 frames x 10 substeps of free fall; returns (drop, implicit-Euler closed form,
    max lateral drift)."""
    fab = dict(scenes.test_jersey(), damping=0.0)
    sc = scenes.Scene(garments=[_flat("sheet", 6, 6, 0.02, fab)])
    s = make()
    s.build(sc)
    before = s.read_positions().astype(np.float64)
    for _ in range(30):
        s.step(1 / 60, 10, 2)
    after = s.read_positions().astype(np.float64)
    h = 1 / 600
    n = 300
    closed = G * h * h * n * (n + 1) / 2  # implicit Euler from rest: x_k = x_0 - g h^2 k (k + 1) / 2
    drop = before[:, 1] - after[:, 1]
    lateral = np.max(np.abs(after[:, [0, 2]] - before[:, [0, 2]]))
    return drop, closed, lateral


def pins_hold(make):
    sc = scenes.Scene(garments=[_flat("sheet", 6, 6, 0.02, scenes.test_jersey())])
    sc.pins = [(0, tuple(sc.garments[0].initial_positions[0]))]
    s = make()
    s.build(sc)
    for _ in range(60):
        s.step(1 / 60, 10, 2)
    return np.linalg.norm(s.read_positions()[0].astype(np.float64) - np.asarray(sc.pins[0][1]))


def assembling_still(make):
    sc = scenes.Scene(garments=[_flat("sheet", 6, 6, 0.02, scenes.test_jersey())])
    s = make()
    s.build(sc)
    s.set_phase("sheet", 0)
    before = s.read_positions()
    for _ in range(30):
        s.step(1 / 60, 10, 2)
    return float(np.max(np.linalg.norm(s.read_positions() - before, axis=1)))


def run_until_still(s, dt, substeps, iterations, rms, hold, max_seconds):
    still = 0
    for f in range(1, int(math.ceil(max_seconds / dt)) + 1):
        s.step(dt, substeps, iterations)
        still = still + 1 if s.stats()["rms_speed"] < rms else 0
        if still >= hold:
            return f
    return -1


def strip_strain(make, grain=(1.0, 0.0)):
    """test_cpu_solver.cpp:104-145: 0.2 x 0.05 m strip (5 mm grid), pinned at x = 0, pulled along +x with 0.5 N
    shared by tributary width at x = 0.2, no gravity; strain = slope of x over rest x in [0.04, 0.16] - 1."""
    fab = dict(scenes.test_jersey(), damping=1.0)
    g = _flat("strip", 41, 11, 0.005, fab, grain=grain)
    m = g.mesh
    sc = scenes.Scene(garments=[g], settings=scenes.settings(gravity=(0.0, 0.0, 0.0)))
    sc.pins = [(int(v), tuple(g.initial_positions[v])) for v in np.flatnonzero(m.rest[:, 0] <= 1e-9)]
    loaded = np.flatnonzero(m.rest[:, 0] >= 0.2 - 1e-9)
    loaded = loaded[np.argsort(m.rest[loaded, 1])]
    ys = m.rest[loaded, 1]
    for i, v in enumerate(loaded):
        lo = ys[i] if i == 0 else 0.5 * (ys[i] + ys[i - 1])
        hi = ys[i] if i + 1 == len(loaded) else 0.5 * (ys[i] + ys[i + 1])
        sc.forces.append((int(v), (0.5 * (hi - lo) / 0.05, 0.0, 0.0)))
    s = make()
    s.build(sc)
    frames = run_until_still(s, 1 / 60, 1, 2, 1e-5, 60, 30.0)
    x = s.read_positions().astype(np.float64)
    sel = (m.rest[:, 0] >= 0.04) & (m.rest[:, 0] <= 0.16)
    slope = np.polyfit(m.rest[sel, 0], x[sel, 0], 1)[0]
    return frames, slope - 1.0


def flat_stays_flat(make):
    """test_bending.cpp:56-64."""
    sc = scenes.Scene(garments=[_flat("sheet", 21, 21, 0.01, scenes.test_stiff())],
                      settings=scenes.settings(gravity=(0.0, 0.0, 0.0)))
    s = make()
    s.build(sc)
    for _ in range(60):
        s.step(1 / 60, 2, 10)
    return float(np.max(np.abs(s.read_positions()[:, 1])))


SPHERE_R = 0.15
K_CONTACT = 0.003 + 0.0006 / 2  # collision offset + half testJersey thickness (sphere_scene.hpp:21)


def sphere_field():
    return scenes.sdf_field(scenes.sphere_distance((0, 0, 0), SPHERE_R), lo=(-0.35, -0.2, -0.35), hi=(0.35, 0.3, 0.35))


def centred_sheet(size, h, height):
    n = int(round(size / h)) + 1
    return _flat("sheet", n, n, h, scenes.test_jersey(), origin=(-size / 2, height, -size / 2))


def rests_on_sphere(make, frames=120):
    """test_body_collision.cpp:31-43: a small patch dropped 5 cm onto the sphere rests at phi = contact radius."""
    field = sphere_field()
    sc = scenes.Scene(garments=[centred_sheet(0.004, 0.004, SPHERE_R + 0.05)], body=field)
    s = make()
    s.build(sc)
    for _ in range(frames):
        s.step(1 / 60, 1, 2)
    import pyoracle
    phi, _ = pyoracle.field_query(field, s.read_positions().astype(np.float64))
    return phi


def slope_slide(make, friction):
    """test_body_collision.cpp:127-175: displacement uphill of a 2 cm patch on a 20 degree slope after 120 frames."""
    a = math.radians(20.0)
    up = np.array([-math.sin(a), math.cos(a), 0.0])
    uphill = np.array([math.cos(a), math.sin(a), 0.0])

    def dist(x, y, z):
        return x * up[0] + y * up[1] + z * up[2]
    field = scenes.sdf_field(dist, lo=(-0.12, -0.12, -0.12), hi=(0.12, 0.12, 0.12))
    fab = dict(scenes.test_jersey(), friction=friction)
    mesh = scenes.grid_mesh(3, 3, 0.01)
    r = mesh.rest - 0.01
    pos = np.array([ri[0] * uphill + ri[1] * np.array([0, 0, 1.0]) + 0.004 * up for ri in r])
    g = scenes.Garment(id="patch", mesh=mesh, initial_positions=pos, piece_fabric=[fab])
    sc = scenes.Scene(garments=[g], body=field)
    s = make()
    s.build(sc)
    start = s.read_positions().astype(np.float64).mean(axis=0)
    for _ in range(120):
        s.step(1 / 60, 1, 2)
    return float((s.read_positions().astype(np.float64).mean(axis=0) - start) @ uphill)


def sewn_pair(gap=0.1, w=0.05, h=0.01):
    """Two w x w squares in the XZ plane, the second shifted by `gap` beyond the first's right edge, sewn
    right edge <-> left edge (test_seams.cpp:13-30)."""
    n = int(round(w / h)) + 1
    a = scenes.grid_mesh(n, n, h)
    b = scenes.grid_mesh(n, n, h)
    mesh = scenes.join_meshes([a, b], [0.0, w + gap])
    na = a.n
    mesh.seam_pairs = np.asarray([(j * n + (n - 1), na + j * n, 0, j) for j in range(n)], np.uint32)
    pos = np.stack([mesh.rest[:, 0], np.zeros(mesh.n), mesh.rest[:, 1]], axis=1)
    fab = scenes.test_jersey()
    return scenes.Garment(id="pair", mesh=mesh, initial_positions=pos, piece_fabric=[fab, fab])


def seams_close(make):
    """test_seams.cpp:42-54: seam gap 47-53 mm after 12 frames of Assembling, < 1 mm after 30."""
    sc = scenes.Scene(garments=[sewn_pair()], settings=scenes.settings(gravity=(0.0, 0.0, 0.0)))
    s = make()
    s.build(sc)
    start = s.stats()["max_seam_gap"]
    s.set_phase("pair", 0)
    for _ in range(12):
        s.step(1 / 60, 10, 2)
    mid = s.stats()["max_seam_gap"]
    for _ in range(18):
        s.step(1 / 60, 10, 2)
    return start, mid, s.stats()["max_seam_gap"]


def snapshot_replay(make):
    """test_body_collision.cpp:108-123: snapshot/restore replays bit-identically in contact."""
    sc = scenes.Scene(garments=[centred_sheet(0.2, 0.02, SPHERE_R + 0.02)], body=sphere_field())
    s = make()
    s.build(sc)
    for _ in range(60):
        s.step(1 / 60, 1, 2)
    snap = s.snapshot()
    for _ in range(20):
        s.step(1 / 60, 1, 2)
    first = s.read_positions()
    s.restore(snap)
    for _ in range(20):
        s.step(1 / 60, 1, 2)
    return first, s.read_positions()


def test_aniso() -> dict:
    """testAniso (libs/core/src/test_fabrics.cpp:8): testStiff with a softer weft bend."""
    return scenes.fabric(20000, 20000, 5000, 0.02, 0.02, 100e-6, 40e-6, 0.450, 0.0008, 0.3, 1.0)


def _cdt_garment(gid, pattern, h, fabs, offsets=None):
    """A pattern meshed as the reference meshes it (buildSimMesh: boundary samples, hex lattice, constrained
    Delaunay triangulation) and laid flat in the XZ plane, piece k shifted by offsets[k] along x
    (flatGarment, sim_test_utils.hpp:15-25)."""
    from drape_amd import garment
    mesh = garment.build_sim_mesh(pattern, h)
    shift = np.zeros(mesh.n) if offsets is None else np.asarray(offsets)[mesh.vertex_piece]
    pos = np.stack([mesh.rest[:, 0] + shift, np.zeros(mesh.n), mesh.rest[:, 1]], axis=1)
    return scenes.Garment(id=gid, mesh=mesh, initial_positions=pos, piece_fabric=list(fabs))


def cantilever_ratio(make, fab, grain, B, h=0.002, width=0.025):
    """Peirce cantilever (test_cantilever.cpp:17-62): a strip overhanging l = 2 (B / W)^(1/3) beyond a 1 cm clamp
    (W = weight g), meshed at h, sags under gravity until still; the bending length c = overhang (cos(theta / 2) /
    (8 tan theta))^(1/3) from the tip's droop angle theta gives the flexural rigidity G = W c^3. Returns (frames,
    G / B, theta in degrees)."""
    from drape_amd import garment
    W = fab["weight"] * G
    length = 2.0 * (B / W) ** (1.0 / 3.0) + 0.01
    pat = garment.Pattern(pieces=[garment.rectangle_piece("strip", length, width, grain)], seams=[])
    g = _cdt_garment("strip", pat, h, [fab])
    m = g.mesh
    pinned = m.rest[:, 0] <= 0.01
    sc = scenes.Scene(garments=[g])
    sc.pins = [(int(v), tuple(g.initial_positions[v])) for v in np.flatnonzero(pinned)]
    # Clamp line: mean rest x of the pinned vertices that share a triangle with a free one.
    frontier = np.zeros(m.n, bool)
    for t in m.triangles:
        if pinned[t].any() and not pinned[t].all():
            frontier[t[pinned[t]]] = True
    x_clamp = float(m.rest[frontier, 0].mean())
    tip = np.flatnonzero(m.rest[:, 0] >= length - 1e-9)
    s = make()
    s.build(sc)
    frames = run_until_still(s, 1 / 60, 1, 2, 2e-4, 60, 10.0)
    x = s.read_positions().astype(np.float64)
    tip_pos = x[tip].mean(axis=0)
    theta = math.atan2(-tip_pos[1], tip_pos[0] - x_clamp)
    overhang = float(m.rest[tip, 0].mean()) - x_clamp
    c = overhang * (math.cos(theta / 2) / (8 * math.tan(theta))) ** (1.0 / 3.0)
    return frames, W * c ** 3 / B, math.degrees(theta)


def _tip_drop(make, garment_, tip_x):
    """Settled tip drop of a cantilever pinned where its initial x <= 1 cm (test_seam_hinge.cpp:17-36)."""
    pos = garment_.initial_positions
    sc = scenes.Scene(garments=[garment_])
    sc.pins = [(int(v), tuple(pos[v])) for v in np.flatnonzero(pos[:, 0] <= 0.01)]
    tip = np.flatnonzero(pos[:, 0] >= tip_x - 1e-9)
    s = make()
    s.build(sc)
    frames = run_until_still(s, 1 / 60, 1, 2, 2e-4, 60, 10.0)
    return frames, float(-s.read_positions().astype(np.float64)[tip, 1].mean())


def seam_hinge_drops(make, h=0.004):
    """SeamDoesNotHinge (test_seam_hinge.cpp:38-59): a 6 x 2 cm testStiff cantilever in one piece, and the same
    strip as two 3 cm pieces side by side, piece a's right edge sewn to piece b's left edge (zero-length seam
    springs, and seam hinges across consecutive seam pairs). Returns (single drop, sewn drop, frames of each)."""
    from drape_amd import garment
    fab = scenes.test_stiff()
    one = garment.Pattern(pieces=[garment.rectangle_piece("one", 0.06, 0.02)], seams=[])
    pa, pb = garment.rectangle_piece("a", 0.03, 0.02), garment.rectangle_piece("b", 0.03, 0.02)
    two = garment.Pattern(pieces=[pa, pb], seams=[garment.Seam("s", garment.segment_edge(pa, "right", "right", False),
                                                               garment.segment_edge(pb, "left", "left", True))])
    f1, d1 = _tip_drop(make, _cdt_garment("one", one, h, [fab]), 0.06)
    f2, d2 = _tip_drop(make, _cdt_garment("two", two, h, [fab, fab], offsets=[0.0, 0.03]), 0.06)
    return d1, d2, f1, f2


def hanging_settles(make):
    """HangingSheetSettlesSymmetric (test_drape_runner.cpp:108-125): the 0.5 m testJersey sheet hanging from its
    top corners drapes to rest within 15 simulated seconds; returns (settled, centre-of-mass x, bottom-left y,
    bottom-right y)."""
    from drape_amd import tee
    sc = scenes.hanging_sheet()
    n = int(round(0.5 / 0.01)) + 1
    s = make()
    s.build(sc)
    r = tee.drape_to_rest(s, ["sheet"], tee.DrapeOptions(max_sim_seconds=15.0))
    x = s.read_positions().astype(np.float64)
    return r.settled, float(x[:, 0].mean()), float(x[0, 1]), float(x[n - 1, 1])


def two_triangle_hinge(L, a, hA, b, hB, grain_angle, bend_warp, bend_weft):
    """One hinge: edge (0,0) -> (L,0), wings (a, hA) and (b, -hB), grain at grain_angle, as a one-garment scene."""
    rest = np.array([[0.0, 0.0], [L, 0.0], [a, hA], [b, -hB]])
    tris = np.array([[0, 1, 2], [1, 0, 3]], np.uint32)
    area = np.full(4, 0.25 * L * (hA + hB) / 2)
    grain = np.array([[math.cos(grain_angle), math.sin(grain_angle)]])
    mesh = scenes.Mesh(rest=rest, vertex_piece=np.zeros(4, np.uint32), triangles=tris,
                       triangle_piece=np.zeros(2, np.uint32), vertex_area=area, boundary=np.ones(4, np.uint8),
                       piece_grain=grain, seam_pairs=np.zeros((0, 4), np.uint32))
    fab = dict(scenes.test_jersey(), bend_warp=bend_warp, bend_weft=bend_weft)
    pos = np.stack([rest[:, 0], np.zeros(4), rest[:, 1]], axis=1)
    g = scenes.Garment(id="hinge", mesh=mesh, initial_positions=pos, piece_fabric=[fab])
    return scenes.Scene(garments=[g])


def hinge_k_closed_form(L, hA, hB, grain_angle, bend_warp, bend_weft):
    """Hinge stiffness of two_triangle_hinge from the geometry: the directional rigidity of the edge (the grain
    at angle g to the edge: p sin^2 g + q cos^2 g with p = max((3 B_warp - B_weft) / 2, f), q = max(B_weft, f),
    f = B_min / 20; both sides of one piece, half each) times |e|^2 over the two faces' area L (hA + hB) / 2."""
    floor = 0.05 * min(bend_warp, bend_weft)
    p = max((3 * bend_warp - bend_weft) / 2, floor)
    q = max(bend_weft, floor)
    rigidity = p * math.sin(grain_angle) ** 2 + q * math.cos(grain_angle) ** 2
    return rigidity * L * L / (0.5 * L * (hA + hB))
