"""Constrained Delaunay triangulation of a simple polygon's boundary samples and interior points.

The reference meshes each pattern piece with the CDT library (artem-ogre/CDT 1.4.5, fetched at build time and
absent here; libs/garment/src/sim_mesh_builder.cpp:262-276: insertVertices, insertEdges on the closed boundary
loop, eraseOuterTriangles). Its published algorithm is restated here:

1. the Delaunay triangulation of all points (Qhull, through scipy.spatial.Delaunay);
2. each boundary edge missing from it is recovered by flipping the edges that cross it (Sloan, "A fast algorithm
   for generating constrained Delaunay triangulations", Computers & Structures 47(3), 1993): an edge whose two
   triangles form a strictly convex quadrilateral is flipped; one that does not waits for a later pass;
3. the Delaunay property is restored by Lawson flips of the non-constrained edges created in step 2 (an edge is
   flipped while the opposite vertex of one triangle lies inside the other's circumcircle);
4. triangles reachable from the convex hull without crossing a boundary edge are erased.

For points in general position the constrained Delaunay triangulation is unique, so this returns the same
triangle set as the library; triangle order and the vertex order within a triangle differ (triangles are
returned counter-clockwise, sorted).
"""
from __future__ import annotations

import numpy as np


def _incircle(pa, pb, pc, pd) -> float:
    """> 0 when pd lies inside the circumcircle of the counter-clockwise triangle (pa, pb, pc)."""
    adx, ady = pa[0] - pd[0], pa[1] - pd[1]
    bdx, bdy = pb[0] - pd[0], pb[1] - pd[1]
    cdx, cdy = pc[0] - pd[0], pc[1] - pd[1]
    return ((adx * adx + ady * ady) * (bdx * cdy - cdx * bdy) - (bdx * bdx + bdy * bdy) * (adx * cdy - cdx * ady)
            + (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady))


def _orient(pa, pb, pc) -> float:
    return (pb[0] - pa[0]) * (pc[1] - pa[1]) - (pb[1] - pa[1]) * (pc[0] - pa[0])


def _crosses(pa, pb, pc, pd) -> bool:
    """Segments ab and cd cross at a point interior to both."""
    o1, o2 = _orient(pa, pb, pc), _orient(pa, pb, pd)
    o3, o4 = _orient(pc, pd, pa), _orient(pc, pd, pb)
    return ((o1 > 0 and o2 < 0) or (o1 < 0 and o2 > 0)) and ((o3 > 0 and o4 < 0) or (o3 < 0 and o4 > 0))


class _Tri:
    """Triangles as counter-clockwise vertex triples with an edge -> triangles map."""

    def __init__(self, pts, simplices):
        self.p = pts
        self.t = []
        self.edges: dict = {}
        for a, b, c in simplices:
            if _orient(pts[a], pts[b], pts[c]) < 0:
                b, c = c, b
            self._add((a, b, c))

    @staticmethod
    def key(a, b):
        return (a, b) if a < b else (b, a)

    def _add(self, tri) -> int:
        i = len(self.t)
        self.t.append(tri)
        a, b, c = tri
        for e in ((a, b), (b, c), (c, a)):
            self.edges.setdefault(self.key(*e), []).append(i)
        return i

    def _replace(self, i, tri) -> None:
        a, b, c = self.t[i]
        for e in ((a, b), (b, c), (c, a)):
            lst = self.edges[self.key(*e)]
            lst.remove(i)
            if not lst:
                del self.edges[self.key(*e)]
        self.t[i] = tri
        a, b, c = tri
        for e in ((a, b), (b, c), (c, a)):
            self.edges.setdefault(self.key(*e), []).append(i)

    def opposite(self, i, a, b) -> int:
        for v in self.t[i]:
            if v != a and v != b:
                return v
        raise RuntimeError("degenerate triangle")

    def quad(self, e):
        """The two triangles on edge e and their vertices opposite it, or None on the hull."""
        lst = self.edges.get(e)
        if lst is None or len(lst) != 2:
            return None
        t0, t1 = lst
        return t0, t1, self.opposite(t0, *e), self.opposite(t1, *e)

    def flip(self, e):
        """Replaces the triangles (a, b, c) and (b, a, d) on edge e = {a, b} by (c, d, b) and (d, c, a); returns
        the new edge {c, d}, or None if the quadrilateral is not strictly convex."""
        q = self.quad(e)
        if q is None:
            return None
        t0, t1, c, d = q
        a, b = e
        # Orient so that (a, b, c) is counter-clockwise.
        if _orient(self.p[a], self.p[b], self.p[c]) < 0:
            a, b = b, a
        p = self.p
        if not (_orient(p[c], p[d], p[b]) > 0 and _orient(p[d], p[c], p[a]) > 0):
            return None
        self._replace(t0, (c, d, b))
        self._replace(t1, (d, c, a))
        return self.key(c, d)


def constrained_delaunay(points: np.ndarray, boundary_count: int) -> np.ndarray:
    """Triangles (T, 3) int64, counter-clockwise, of the constrained Delaunay triangulation of `points` whose first
    `boundary_count` entries form the closed boundary loop, restricted to the inside of that loop."""
    from scipy.spatial import Delaunay

    pts = np.ascontiguousarray(points, dtype=np.float64)
    nb = int(boundary_count)
    dt = Delaunay(pts)
    if len(dt.coplanar):
        raise ValueError("triangulation dropped coincident or coplanar points")
    P = [tuple(p) for p in pts.tolist()]
    T = _Tri(P, dt.simplices.tolist())
    constraints = {T.key(i, (i + 1) % nb) for i in range(nb)}

    created = []
    for i in range(nb):
        a, b = i, (i + 1) % nb
        if T.key(a, b) in T.edges:
            continue
        pa, pb = P[a], P[b]
        # Edges that cross the segment ab: walk the triangles around a towards b, then along the segment.
        crossing = [e for e in T.edges if a not in e and b not in e and _crosses(pa, pb, P[e[0]], P[e[1]])]
        queue = list(crossing)
        guard = 0
        while queue:
            guard += 1
            if guard > 100000:
                raise RuntimeError("boundary edge recovery did not converge")
            e = queue.pop(0)
            if e not in T.edges:
                continue
            ne = T.flip(e)
            if ne is None:
                queue.append(e)
                continue
            if ne[0] not in (a, b) and ne[1] not in (a, b) and _crosses(pa, pb, P[ne[0]], P[ne[1]]):
                queue.append(ne)
            else:
                created.append(ne)
        if T.key(a, b) not in T.edges:
            raise RuntimeError("boundary edge could not be recovered")

    # Lawson flips from the edges created above.
    work = [e for e in created if e not in constraints]
    guard = 0
    while work:
        guard += 1
        if guard > 1000000:
            raise RuntimeError("Delaunay restoration did not converge")
        e = work.pop()
        if e in constraints:
            continue
        q = T.quad(e)
        if q is None:
            continue
        t0, t1, c, d = q
        a, b, cc = T.t[t0]
        scale = max(abs(P[a][0] - P[d][0]), abs(P[a][1] - P[d][1]), 1e-300) ** 4
        if _incircle(P[a], P[b], P[cc], P[d]) <= 1e-12 * scale:
            continue
        ea, eb = e
        ne = T.flip(e)
        if ne is None:
            continue
        for x, y in ((ea, c), (c, eb), (eb, d), (d, ea)):
            k = T.key(x, y)
            if k not in constraints:
                work.append(k)

    # Erase the triangles outside the boundary loop: flood from hull edges that are not boundary edges.
    n = len(T.t)
    outside = np.zeros(n, dtype=bool)
    stack = []
    for e, lst in T.edges.items():
        if len(lst) == 1 and e not in constraints and not outside[lst[0]]:
            outside[lst[0]] = True
            stack.append(lst[0])
    while stack:
        i = stack.pop()
        a, b, c = T.t[i]
        for x, y in ((a, b), (b, c), (c, a)):
            k = T.key(x, y)
            if k in constraints:
                continue
            for j in T.edges[k]:
                if not outside[j]:
                    outside[j] = True
                    stack.append(j)
    out = np.asarray([T.t[i] for i in range(n) if not outside[i]], dtype=np.int64).reshape(-1, 3)
    # Canonical order: rotate each triangle to start at its smallest vertex, then sort.
    r = np.argmin(out, axis=1)
    out = np.stack([out[np.arange(len(out)), (r + k) % 3] for k in range(3)], axis=1)
    return out[np.lexsort((out[:, 2], out[:, 1], out[:, 0]))]
