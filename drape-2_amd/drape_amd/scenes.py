"""Scene builders for the benchmark configurations and the tests.

Each builder returns a `Scene`: garments on regular grids (rest positions in pattern space, counter-clockwise
triangles, vertexArea = 1/3 of incident rest areas, one grainline per piece — the SimMesh fields of
libs/core/include/drape/core/sim_mesh.hpp:11-27), a signed-distance body sampled on a regular grid in the
CollisionField layout (collision_field.hpp:11-20: origin, cell size, dims, band, values x-fastest, clamped to
[-band, band]), pins, forces and SimSettings (solver.hpp:19-30).

Material mapping from the configurations' continuum constants onto the reference's decoupled model
(energies.cpp:40-62): stretchWarp = stretchWeft = E t, stretchBias = E t / (2 (1 + nu)) (the membrane shear
modulus times thickness), strain limits off, bendWarp = bendWeft = k (the configurations' bending constant read
as the bending rigidity B in N m; the hinge stiffness is then K = B |e|^2 / (A1 + A2), sim_model.cpp:20-27),
weight = areal density, damping 0.3 (the reference's test fabrics' value; the configurations give none).
Poisson coupling has no counterpart in the decoupled model and is not represented.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

DEFAULT_SETTINGS = {
    "gravity": (0.0, -9.81, 0.0),
    "collision_offset": 0.003,
    "assembly_closing_speed": 0.25,
    "assembly_extra_damping": 10.0,
    "damping_scale": 5.0,
    "strain_limit_stiffness_scale": 100.0,
    "seam_stiffness": 1e4,
    "friction_slip_speed": 3e-4,
    "max_cg_iterations": 50,
    "cg_tolerance": 1e-2,
}


def settings(**over) -> dict:
    s = dict(DEFAULT_SETTINGS)
    s.update(over)
    return s


def fabric(stretch_warp, stretch_weft, stretch_bias, strain_limit_warp, strain_limit_weft, bend_warp, bend_weft,
           weight, thickness, friction, damping) -> dict:
    return dict(stretch_warp=stretch_warp, stretch_weft=stretch_weft, stretch_bias=stretch_bias,
                strain_limit_warp=strain_limit_warp, strain_limit_weft=strain_limit_weft, bend_warp=bend_warp,
                bend_weft=bend_weft, weight=weight, thickness=thickness, friction=friction, damping=damping)


# The reference's test fabrics (libs/core/src/test_fabrics.cpp:5-8).
def test_jersey() -> dict:
    return fabric(400, 200, 100, 0.5, 0.8, 2e-6, 1.5e-6, 0.160, 0.0006, 0.3, 0.3)


def test_stiff() -> dict:
    return fabric(20000, 20000, 5000, 0.02, 0.02, 100e-6, 100e-6, 0.450, 0.0008, 0.3, 1.0)


def membrane_fabric(young=5e4, poisson=0.3, thickness=5e-4, density=0.2, bend=1e-5, friction=0.3,
                    damping=0.3) -> dict:
    """Continuum constants of the benchmark configurations mapped onto FabricPhysical (module docstring)."""
    et = young * thickness
    return fabric(et, et, et / (2.0 * (1.0 + poisson)), 0.0, 0.0, bend, bend, density, thickness, friction, damping)


@dataclasses.dataclass
class Mesh:
    rest: np.ndarray            # (n, 2) float64
    vertex_piece: np.ndarray    # (n,) uint32
    triangles: np.ndarray       # (T, 3) uint32
    triangle_piece: np.ndarray  # (T,) uint32
    vertex_area: np.ndarray     # (n,) float64
    boundary: np.ndarray        # (n,) uint8
    piece_grain: np.ndarray     # (P, 2) float64
    seam_pairs: np.ndarray      # (K, 4) uint32: a, b, seam, index

    @property
    def n(self) -> int:
        return len(self.rest)


@dataclasses.dataclass
class Garment:
    id: str
    mesh: Mesh
    initial_positions: np.ndarray  # (n, 3) float64
    piece_fabric: list


@dataclasses.dataclass
class Field:
    origin: tuple
    cell_size: float
    dims: tuple  # (nx, ny, nz)
    band: float
    values: np.ndarray  # (nz, ny, nx) float32


@dataclasses.dataclass
class Scene:
    garments: list
    body: Field | None = None
    pins: list = dataclasses.field(default_factory=list)    # [(particle, (x, y, z))]
    forces: list = dataclasses.field(default_factory=list)  # [(particle, (fx, fy, fz))]
    settings: dict = dataclasses.field(default_factory=settings)
    name: str = ""

    @property
    def n(self) -> int:
        return sum(g.mesh.n for g in self.garments)


# ---------------------------------------------------------------------------------------------------------------
# Meshes.

def grid_mesh(nx: int, ny: int, spacing: float, grain=(1.0, 0.0)) -> Mesh:
    """nx x ny vertices (x fastest), every quad split along the same diagonal into two CCW triangles."""
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    rest = np.stack([ii.reshape(-1) * spacing, jj.reshape(-1) * spacing], axis=1).astype(np.float64)
    qi, qj = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="xy")
    qi = qi.reshape(-1)
    qj = qj.reshape(-1)
    v00 = qj * nx + qi
    v10 = v00 + 1
    v01 = v00 + nx
    v11 = v01 + 1
    # Quad (i, j) contributes (v00, v10, v11) then (v00, v11, v01).
    tris = np.empty((2 * len(v00), 3), dtype=np.uint32)
    tris[0::2] = np.stack([v00, v10, v11], axis=1)
    tris[1::2] = np.stack([v00, v11, v01], axis=1)
    area = np.zeros(nx * ny)
    a = rest[tris[:, 1]] - rest[tris[:, 0]]
    b = rest[tris[:, 2]] - rest[tris[:, 0]]
    tri_area = 0.5 * np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
    for k in range(3):
        np.add.at(area, tris[:, k], tri_area / 3.0)
    boundary = ((ii == 0) | (ii == nx - 1) | (jj == 0) | (jj == ny - 1)).reshape(-1).astype(np.uint8)
    g = np.asarray(grain, dtype=np.float64)
    return Mesh(rest=rest, vertex_piece=np.zeros(nx * ny, dtype=np.uint32), triangles=tris,
                triangle_piece=np.zeros(len(tris), dtype=np.uint32), vertex_area=area, boundary=boundary,
                piece_grain=(g / np.linalg.norm(g)).reshape(1, 2), seam_pairs=np.zeros((0, 4), dtype=np.uint32))


def join_meshes(meshes: list, offsets_x: list) -> Mesh:
    """Concatenates single-piece meshes into one multi-piece mesh (pieces in order); rest x shifted per piece."""
    rest, vp, tris, tp, area, bnd, grains = [], [], [], [], [], [], []
    base = 0
    for p, (m, dx) in enumerate(zip(meshes, offsets_x)):
        r = m.rest.copy()
        r[:, 0] += dx
        rest.append(r)
        vp.append(np.full(m.n, p, dtype=np.uint32))
        tris.append(m.triangles.astype(np.int64) + base)
        tp.append(np.full(len(m.triangles), p, dtype=np.uint32))
        area.append(m.vertex_area)
        bnd.append(m.boundary)
        grains.append(m.piece_grain[0])
        base += m.n
    return Mesh(rest=np.concatenate(rest), vertex_piece=np.concatenate(vp),
                triangles=np.concatenate(tris).astype(np.uint32), triangle_piece=np.concatenate(tp),
                vertex_area=np.concatenate(area), boundary=np.concatenate(bnd), piece_grain=np.stack(grains),
                seam_pairs=np.zeros((0, 4), dtype=np.uint32))


def flat_garment(gid: str, mesh: Mesh, fab: dict, origin=(0.0, 0.0, 0.0)) -> Garment:
    """Lays a mesh in the XZ plane: world = origin + (rest.x, 0, rest.y) (as flatGarment, sim_test_utils.hpp:15-25)."""
    o = np.asarray(origin, dtype=np.float64)
    pos = np.stack([mesh.rest[:, 0], np.zeros(mesh.n), mesh.rest[:, 1]], axis=1) + o
    return Garment(id=gid, mesh=mesh, initial_positions=pos, piece_fabric=[dict(fab)] * len(mesh.piece_grain))


# ---------------------------------------------------------------------------------------------------------------
# Signed-distance bodies on a grid.

def _grid_points(lo, hi, cell):
    lo = np.asarray(lo, dtype=np.float64)
    dims = tuple(int(math.ceil((hi[a] - lo[a]) / cell)) + 1 for a in range(3))
    return lo, dims


def sdf_field(distance, lo, hi, cell=0.004, band=0.02) -> Field:
    """Samples distance(x, y, z) (vectorised, negative inside) on a grid covering [lo, hi], clamped to the band."""
    origin, dims = _grid_points(lo, hi, cell)
    nx, ny, nz = dims
    xs = origin[0] + cell * np.arange(nx)
    ys = origin[1] + cell * np.arange(ny)
    values = np.empty((nz, ny, nx), dtype=np.float32)
    X, Y = np.meshgrid(xs, ys, indexing="xy")  # (ny, nx)
    for k in range(nz):
        z = origin[2] + cell * k
        d = distance(X, Y, np.full_like(X, z))
        values[k] = np.clip(d, -band, band).astype(np.float32)
    return Field(origin=tuple(origin), cell_size=cell, dims=dims, band=band, values=values)


def sphere_distance(center, radius):
    cx, cy, cz = center

    def d(x, y, z):
        return np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) - radius
    return d


def plane_distance(height):
    def d(x, y, z):
        return y - height
    return d


def capsule_distance(a, b, radius):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ab = b - a
    L2 = float(ab @ ab)

    def d(x, y, z):
        px, py, pz = x - a[0], y - a[1], z - a[2]
        t = np.clip((px * ab[0] + py * ab[1] + pz * ab[2]) / L2, 0.0, 1.0)
        qx, qy, qz = px - t * ab[0], py - t * ab[1], pz - t * ab[2]
        return np.sqrt(qx * qx + qy * qy + qz * qz) - radius
    return d


def union(*ds):
    def d(x, y, z):
        out = ds[0](x, y, z)
        for f in ds[1:]:
            out = np.minimum(out, f(x, y, z))
        return out
    return d


# ---------------------------------------------------------------------------------------------------------------
# Benchmark configurations.

def c1_sheet(n=32, spacing=0.01, seed=1, jitter=1e-4, steps_tolerance=1e-6, max_cg=2000) -> Scene:
    """32x32 sheet (spacing 1 cm) in the XZ plane at y = 0.5 m, centred on the origin, pinned at its two z-min
    corners; sphere r = 0.2 m at the origin; mu = 0.3; Newton <= 5, PCG to 1e-6 (maxCg raised to `max_cg`).
    Initial positions jittered per component by U(-jitter, jitter) (the 2D rest shape stays exact)."""
    mesh = grid_mesh(n, n, spacing)
    half = 0.5 * (n - 1) * spacing
    g = flat_garment("sheet", mesh, membrane_fabric(friction=0.3), origin=(-half, 0.5, -half))
    rng = np.random.default_rng(seed)
    g.initial_positions = g.initial_positions + rng.uniform(-jitter, jitter, size=g.initial_positions.shape)
    pins = [(0, tuple(g.initial_positions[0])), (n - 1, tuple(g.initial_positions[n - 1]))]
    body = sdf_field(sphere_distance((0, 0, 0), 0.2), lo=(-0.26, -0.26, -0.26), hi=(0.26, 0.56, 0.26))
    return Scene(garments=[g], body=body, pins=pins,
                 settings=settings(max_cg_iterations=max_cg, cg_tolerance=steps_tolerance), name="c1-sheet-32")


def c2_sheet(n=256, spacing=0.004, seed=2, jitter=5e-4, height=0.4, radius=0.25, ground=-0.3,
             tolerance=1e-6, max_cg=2000, half_box=None) -> Scene:
    """n x n sheet (spacing 4 mm) centred above the origin at y = height, no pins, dropped onto a sphere
    (radius at the origin) plus a ground plane; jitter U(-jitter, jitter) per component."""
    mesh = grid_mesh(n, n, spacing)
    half = 0.5 * (n - 1) * spacing
    g = flat_garment("sheet", mesh, membrane_fabric(friction=0.3), origin=(-half, height, -half))
    rng = np.random.default_rng(seed)
    g.initial_positions = g.initial_positions + rng.uniform(-jitter, jitter, size=g.initial_positions.shape)
    hb = half_box if half_box is not None else half + 0.04
    body = sdf_field(union(sphere_distance((0, 0, 0), radius), plane_distance(ground)),
                     lo=(-hb, ground - 0.04, -hb), hi=(hb, height + 0.05, hb))
    return Scene(garments=[g], body=body, settings=settings(max_cg_iterations=max_cg, cg_tolerance=tolerance),
                 name=f"c2-sheet-{n}")


def mannequin_distance():
    """Capsule-union body: torso (r 0.15, segment 0.55 m, vertical, centred on the origin) and two arms (r 0.05,
    0.5 m) leaving the shoulders (+-0.15, 0.275, 0) at 45 degrees below horizontal."""
    torso = capsule_distance((0, -0.275, 0), (0, 0.275, 0), 0.15)
    c = math.cos(math.pi / 4)
    arms = [capsule_distance((s * 0.15, 0.275, 0), (s * (0.15 + 0.5 * c), 0.275 - 0.5 * c, 0), 0.05)
            for s in (-1.0, 1.0)]
    return union(torso, *arms)


def c3_tube(nx=200, ny=300, spacing=0.003, seed=3, jitter=0.0, bottom=0.44, tolerance=1e-6, max_cg=2000) -> Scene:
    """Front and back panels (nx x ny vertices each, one garment, two pieces) bent into a tube of circumference
    2 (nx - 1) spacing around the vertical axis, hanging from y = bottom upwards, above the capsule mannequin;
    side edges sewn (seam springs k = 1e3 N/m, rest length 0): front column nx-1 <-> back column 0 (seam 0),
    front column 0 <-> back column nx-1 (seam 1). mu = 0.4."""
    front = grid_mesh(nx, ny, spacing)
    back = grid_mesh(nx, ny, spacing)
    width = (nx - 1) * spacing
    mesh = join_meshes([front, back], [0.0, width + 0.05])
    R = 2 * width / (2 * math.pi)
    n1 = nx * ny
    pos = np.zeros((2 * n1, 3))
    jj = np.repeat(np.arange(ny), nx)
    ii = np.tile(np.arange(nx), ny)
    for piece, phase0 in ((0, -math.pi / 2), (1, math.pi / 2)):
        phi = phase0 + math.pi * ii / (nx - 1)
        sl = slice(piece * n1, (piece + 1) * n1)
        pos[sl, 0] = R * np.sin(phi)
        pos[sl, 1] = bottom + jj * spacing
        pos[sl, 2] = R * np.cos(phi)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    pairs = []
    for j in range(ny):
        pairs.append((j * nx + (nx - 1), n1 + j * nx + 0, 0, j))
    for j in range(ny):
        pairs.append((j * nx + 0, n1 + j * nx + (nx - 1), 1, j))
    mesh.seam_pairs = np.asarray(pairs, dtype=np.uint32)
    fab = membrane_fabric(friction=0.4)
    g = Garment(id="tube", mesh=mesh, initial_positions=pos, piece_fabric=[dict(fab), dict(fab)])
    top = bottom + (ny - 1) * spacing
    body = sdf_field(mannequin_distance(), lo=(-0.62, -0.5, -0.26), hi=(0.62, top + 0.05, 0.26))
    return Scene(garments=[g], body=body,
                 settings=settings(seam_stiffness=1e3, max_cg_iterations=max_cg, cg_tolerance=tolerance),
                 name="c3-tube")


def c4_batch(count=64, seed=4, lo=2000, hi=40000, spacing=0.004, tolerance=1e-6, max_cg=2000,
             cell=0.008) -> list:
    """`count` independent sheets: particle counts log-uniform in [lo, hi], aspect U(0.5, 2), each dropped 0.1 m
    onto its own sphere or capsule (radius U(0.15, 0.3)) plus ground, each its own segment."""
    rng = np.random.default_rng(seed)
    scenes = []
    for k in range(count):
        target = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        aspect = rng.uniform(0.5, 2.0)
        nx = max(4, int(round(math.sqrt(target / aspect))))
        ny = max(4, int(round(target / nx)))
        radius = rng.uniform(0.15, 0.3)
        capsule = bool(rng.integers(0, 2))
        mesh = grid_mesh(nx, ny, spacing)
        hx, hz = 0.5 * (nx - 1) * spacing, 0.5 * (ny - 1) * spacing
        height = radius + 0.1
        g = flat_garment(f"sheet{k}", mesh, membrane_fabric(friction=0.3), origin=(-hx, height, -hz))
        g.initial_positions = g.initial_positions + rng.uniform(-5e-4, 5e-4, size=g.initial_positions.shape)
        ground = -radius - 0.05
        if capsule:
            body_d = capsule_distance((-0.2, 0, 0), (0.2, 0, 0), radius)
        else:
            body_d = sphere_distance((0, 0, 0), radius)
        ext = max(hx, hz) + 0.04
        body = sdf_field(union(body_d, plane_distance(ground)), lo=(-ext, ground - 0.04, -ext),
                         hi=(ext, height + 0.05, ext), cell=cell)
        scenes.append(Scene(garments=[g], body=body,
                            settings=settings(max_cg_iterations=max_cg, cg_tolerance=tolerance),
                            name=f"c4-{k}-{nx}x{ny}"))
    return scenes


def hanging_sheet(size=0.5, spacing=0.01) -> Scene:
    """A testJersey sheet hanging in the XY plane from its two top corners (hangingSheet, sim_test_utils.hpp:52-80)."""
    n = int(round(size / spacing)) + 1
    mesh = grid_mesh(n, n, spacing)
    pos = np.stack([mesh.rest[:, 0], 0.5 + mesh.rest[:, 1], np.zeros(mesh.n)], axis=1)
    g = Garment(id="sheet", mesh=mesh, initial_positions=pos, piece_fabric=[test_jersey()])
    tl, tr = (n - 1) * n, n * n - 1
    return Scene(garments=[g], pins=[(tl, tuple(pos[tl])), (tr, tuple(pos[tr]))], name="hanging-sheet")
