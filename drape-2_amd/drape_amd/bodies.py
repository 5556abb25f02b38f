"""Closed triangle meshes of simple bodies (the inputs of the device SDF bake): icosphere, capsule, and the
manifold check the bake requires. Triangles are counter-clockwise seen from outside."""
from __future__ import annotations

import math

import numpy as np


def icosphere(center=(0.0, 0.0, 0.0), radius=0.1, subdivisions=3):
    """Icosahedron subdivided `subdivisions` times (each triangle into four, midpoints shared), vertices projected
    onto the sphere: 20 * 4^s triangles. Returns (vertices [n, 3] float64, triangles [t, 3] uint32)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    x = np.asarray(center, float) + radius * np.array(verts)
    return x, np.array(f, np.uint32)


def capsule(a=(0.0, 0.0, 0.0), b=(0.0, 0.5, 0.0), radius=0.1, segments=24, rings=8):
    """Capsule around segment ab: a cylinder of `segments` sides and two hemispherical caps of `rings` rings each,
    closed at the poles. Returns (vertices, triangles)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    axis = b - a
    length = np.linalg.norm(axis)
    w = axis / length
    u = np.cross(w, [1.0, 0.0, 0.0] if abs(w[0]) < 0.9 else [0.0, 1.0, 0.0])
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    verts = [a - radius * w]  # bottom pole
    rows = []
    # rings from the bottom pole to the equator of cap a, then from the equator of cap b to its top pole
    lat = [(-math.pi / 2 + math.pi / 2 * k / rings, a) for k in range(1, rings + 1)]
    lat += [(math.pi / 2 * k / rings, b) for k in range(0, rings)]
    for phi, centre in lat:
        row = []
        for s in range(segments):
            th = 2 * math.pi * s / segments
            d = math.cos(phi) * (math.cos(th) * u + math.sin(th) * v) + math.sin(phi) * w
            verts.append(centre + radius * d)
            row.append(len(verts) - 1)
        rows.append(row)
    verts.append(b + radius * w)  # top pole
    top = len(verts) - 1
    tris = []
    for s in range(segments):
        tris.append((0, rows[0][(s + 1) % segments], rows[0][s]))
    for r in range(len(rows) - 1):
        for s in range(segments):
            p, q = rows[r][s], rows[r][(s + 1) % segments]
            p2, q2 = rows[r + 1][s], rows[r + 1][(s + 1) % segments]
            tris += [(p, q, q2), (p, q2, p2)]
    for s in range(segments):
        tris.append((top, rows[-1][s], rows[-1][(s + 1) % segments]))
    return np.array(verts), np.array(tris, np.uint32)


def is_closed_manifold(triangles) -> bool:
    """Every undirected edge used by exactly two triangles, in opposite directions (triangle_mesh.hpp:16-17)."""
    t = np.asarray(triangles, np.int64)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    directed = set(map(tuple, e))
    return len(directed) == len(e) and all((b, a) in directed for a, b in directed)


def signed_volume(vertices, triangles) -> float:
    x = np.asarray(vertices, float)[np.asarray(triangles, np.int64)]
    return float(np.sum(np.einsum("ij,ij->i", x[:, 0], np.cross(x[:, 1], x[:, 2]))) / 6.0)
