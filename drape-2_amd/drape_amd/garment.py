"""Sewing patterns, the test tee, the simulation mesh builder and initial placement.

Restates the reference's garment pipeline on the host (it feeds the solver; none of it runs on the GPU):

* cubic Bezier outlines with arc-length tables (libs/core/src/bezier.cpp:7-78, outline.cpp:8-76),
* pattern pieces, seams and edge references (libs/core/include/drape/core/pattern.hpp:14-80, pattern.cpp:16-26),
* spec tables and size grading (libs/garment/src/spec_table.cpp:8-25),
* the men's and women's test tee (libs/garment/src/test_tee.cpp:15-221),
* pattern validation (libs/garment/src/validate.cpp:24-87, libs/core/src/polygon.cpp:7-89),
* the simulation mesh: boundary sampling shared across seams, a grain-aligned hexagonal lattice inside each
  piece, a constrained Delaunay triangulation (`cdt.constrained_delaunay`, in place of the CDT 1.4.5 library the
  reference links), three Jacobi smoothing passes and seam pairs (libs/garment/src/sim_mesh_builder.cpp:14-323),
* placement of each piece on its body-region cylinder (libs/garment/src/placement.cpp:23-46).

The geometry (the tee's control points, spec tables, sampling rules and tolerances) has to match the reference's
for the meshes to match: the same particle counts, the same seam pairs.
"""
from __future__ import annotations

import bisect
import dataclasses
import math

import numpy as np

from . import cdt
from .scenes import Mesh

# ---------------------------------------------------------------------------------------------------------------
# Cubic Bezier segments and arc length.

# Five-point Gauss-Legendre rule on [-1, 1] (nodes 0, +-sqrt(5 -+ 2 sqrt(10/7)) / 3; weights 128/225,
# (322 +- 13 sqrt(70)) / 900).
_GL_X = (-math.sqrt(5.0 + 2.0 * math.sqrt(10.0 / 7.0)) / 3.0, -math.sqrt(5.0 - 2.0 * math.sqrt(10.0 / 7.0)) / 3.0,
         0.0, math.sqrt(5.0 - 2.0 * math.sqrt(10.0 / 7.0)) / 3.0, math.sqrt(5.0 + 2.0 * math.sqrt(10.0 / 7.0)) / 3.0)
_GL_W = ((322.0 - 13.0 * math.sqrt(70.0)) / 900.0, (322.0 + 13.0 * math.sqrt(70.0)) / 900.0, 128.0 / 225.0,
         (322.0 + 13.0 * math.sqrt(70.0)) / 900.0, (322.0 - 13.0 * math.sqrt(70.0)) / 900.0)


def _v(x, y) -> tuple:
    return (float(x), float(y))


def _norm(x: float, y: float) -> float:
    return math.sqrt(x * x + y * y)  # as Eigen's norm() of a 2-vector: sqrt of the squared norm


@dataclasses.dataclass(frozen=True)
class Bezier:
    """Cubic Bezier p0..p3 (bezier.hpp)."""
    p0: tuple
    p1: tuple
    p2: tuple
    p3: tuple

    def point(self, t: float) -> tuple:
        u = 1.0 - t
        a, b, c, d = u * u * u, 3 * u * u * t, 3 * u * t * t, t * t * t
        return (a * self.p0[0] + b * self.p1[0] + c * self.p2[0] + d * self.p3[0],
                a * self.p0[1] + b * self.p1[1] + c * self.p2[1] + d * self.p3[1])

    def derivative(self, t: float) -> tuple:
        u = 1.0 - t
        a, b, c = 3 * u * u, 6 * u * t, 3 * t * t
        return (a * (self.p1[0] - self.p0[0]) + b * (self.p2[0] - self.p1[0]) + c * (self.p3[0] - self.p2[0]),
                a * (self.p1[1] - self.p0[1]) + b * (self.p2[1] - self.p1[1]) + c * (self.p3[1] - self.p2[1]))

    def reversed(self) -> "Bezier":
        return Bezier(self.p3, self.p2, self.p1, self.p0)

    def mirrored_x(self) -> "Bezier":
        return Bezier(*[(-p[0], p[1]) for p in (self.p0, self.p1, self.p2, self.p3)])

    @staticmethod
    def line(a, b) -> "Bezier":
        a, b = _v(*a), _v(*b)
        d = (b[0] - a[0], b[1] - a[1])
        return Bezier(a, (a[0] + d[0] / 3.0, a[1] + d[1] / 3.0), (a[0] + 2.0 * d[0] / 3.0, a[1] + 2.0 * d[1] / 3.0), b)


class ArcTable:
    """Cumulative arc length at `samples` + 1 evenly spaced parameters, Gauss-Legendre per interval; the parameter
    of an arc length interpolates linearly between samples (bezier.cpp:40-74)."""

    def __init__(self, c: Bezier, samples: int = 256):
        samples = max(samples, 1)
        cum = [0.0]
        for i in range(samples):
            a, b = i / samples, (i + 1) / samples
            half, mid = 0.5 * (b - a), 0.5 * (a + b)
            s = 0.0
            for x, w in zip(_GL_X, _GL_W):
                d = c.derivative(mid + half * x)
                s += w * _norm(d[0], d[1])
            cum.append(cum[-1] + s * half)
        self.cum = cum

    @property
    def length(self) -> float:
        return self.cum[-1]

    def t_at(self, s: float) -> float:
        total = self.length
        if total <= 0:
            return 0.0
        s = min(max(s, 0.0), total)
        k = bisect.bisect_right(self.cum, s)
        if k == len(self.cum):
            return 1.0
        i = k - 1
        s0, s1 = self.cum[i], self.cum[i + 1]
        frac = (s - s0) / (s1 - s0) if s1 > s0 else 0.0
        return (i + frac) / (len(self.cum) - 1)


def bezier_length(c: Bezier) -> float:
    return ArcTable(c).length


class Outline:
    """A closed chain of named Bezier segments with arc-length lookup (outline.hpp, outline.cpp)."""

    def __init__(self):
        self.segments: list = []
        self.names: list = []
        self.tables: list = []
        self.starts: list = [0.0]

    def add(self, name: str, seg: Bezier) -> None:
        self.segments.append(seg)
        self.names.append(name)
        self.tables.append(ArcTable(seg))
        self.starts.append(self.starts[-1] + self.tables[-1].length)

    def __len__(self) -> int:
        return len(self.segments)

    @property
    def perimeter(self) -> float:
        return self.starts[-1]

    def segment_start(self, i: int) -> float:
        return self.starts[i]

    def segment_end(self, i: int) -> float:
        return self.starts[i + 1]

    def find(self, name: str):
        return self.names.index(name) if name in self.names else None

    def is_closed(self, tol: float = 1e-9) -> bool:
        if not self.segments:
            return False
        for i, s in enumerate(self.segments):
            n = self.segments[(i + 1) % len(self.segments)]
            if _norm(s.p3[0] - n.p0[0], s.p3[1] - n.p0[1]) > tol:
                return False
        return True

    def _locate(self, s: float):
        s = min(max(s, 0.0), self.perimeter)
        if s >= self.perimeter:
            return 0, 0.0  # wraps to the start point
        i = bisect.bisect_right(self.starts, s) - 1
        return i, self.tables[i].t_at(s - self.starts[i])

    def point_at(self, s: float) -> tuple:
        i, t = self._locate(s)
        return self.segments[i].point(t)

    def polyline(self, max_spacing: float) -> list:
        out = []
        for seg, tab in zip(self.segments, self.tables):
            n = max(1, math.ceil(tab.length / max_spacing))
            out += [seg.point(tab.t_at(tab.length * k / n)) for k in range(n)]
        return out

    def __eq__(self, other) -> bool:
        return isinstance(other, Outline) and self.segments == other.segments and self.names == other.names


# ---------------------------------------------------------------------------------------------------------------
# Patterns.

REGIONS = ("Torso", "Neck", "Head", "ArmLeft", "ArmRight", "LegLeft", "LegRight", "Waist")


@dataclasses.dataclass
class Placement:
    """Wrap of a piece onto a body-region cylinder (pattern.hpp:14-22)."""
    region: str = "Torso"
    center_angle: float = 0.0     # radians, angle of pattern x = 0
    axial_at_origin: float = 0.0  # metres along the axis for pattern y = 0
    axial_sign: int = 1           # +1: pattern +y runs along +axis
    offset: float = 0.06          # radial distance outside the region radius
    wrap_radius: float | None = None  # absolute wrap radius, overrides radius + offset


@dataclasses.dataclass
class Piece:
    id: str
    outline: Outline
    grainline: tuple = (0.0, 1.0)
    fabric_slot: str = "body"
    placement: Placement = dataclasses.field(default_factory=Placement)


@dataclasses.dataclass
class EdgeRef:
    """Arc-length range [start, end] on a piece outline; `reversed` traverses it end to start."""
    piece: str
    start: float
    end: float
    reversed: bool = False


@dataclasses.dataclass
class Seam:
    id: str
    a: EdgeRef
    b: EdgeRef
    ease: float = 0.0


@dataclasses.dataclass
class Pattern:
    pieces: list
    seams: list

    def find_piece(self, pid: str):
        for p in self.pieces:
            if p.id == pid:
                return p
        return None


def rectangle_piece(pid: str, width: float, height: float, grain=(0.0, 1.0), fabric_slot: str = "body") -> Piece:
    """Segments bottom, right, top, left, counter-clockwise from the origin (shapes.cpp:5-17)."""
    o = Outline()
    o.add("bottom", Bezier.line((0, 0), (width, 0)))
    o.add("right", Bezier.line((width, 0), (width, height)))
    o.add("top", Bezier.line((width, height), (0, height)))
    o.add("left", Bezier.line((0, height), (0, 0)))
    return Piece(pid, o, _v(*grain), fabric_slot)


def segment_edge(piece: Piece, first: str, last: str, reversed_: bool) -> EdgeRef:
    """The edge from the start of segment `first` to the end of segment `last` (pattern.cpp:16-26)."""
    i, k = piece.outline.find(first), piece.outline.find(last)
    if i is None or k is None:
        raise ValueError(f"piece '{piece.id}' has no segment '{first if i is None else last}'")
    if k < i:
        raise ValueError(f"piece '{piece.id}': segment '{last}' comes before '{first}'")
    return EdgeRef(piece.id, piece.outline.segment_start(i), piece.outline.segment_end(k), reversed_)


# ---------------------------------------------------------------------------------------------------------------
# Spec tables (spec_table.hpp:9, spec_table.cpp:8-25).

SIZES = ("XS", "S", "M", "L", "XL", "XXL")


@dataclasses.dataclass
class SpecTable:
    base_size: str
    poms: list  # [(key, base, grade per size step)]


def resolve_spec(table: SpecTable, size: str, overrides: dict | None = None) -> dict:
    if size not in SIZES:
        raise ValueError(f"unknown size '{size}'")
    steps = SIZES.index(size) - SIZES.index(table.base_size)
    out = {k: base + grade * steps for k, base, grade in table.poms}
    for k, v in ((overrides or {}).get(size) or {}).items():
        out[k] = v
    return out


# ---------------------------------------------------------------------------------------------------------------
# The test tee (test_tee.cpp). Spec keys in the order the generator reads them.

_TEE_KEYS = ("chest_width", "body_length", "shoulder_width", "sleeve_length", "sleeve_opening", "bicep_width",
             "neck_width", "front_neck_drop", "back_neck_drop", "armhole_depth", "shoulder_drop", "neck_rib_height")

# Men's and women's basic crew-neck tee, base size M: (base, grade per size) per key (test_tee.cpp:131-175).
MENS_TEE = SpecTable("M", list(zip(_TEE_KEYS, (0.530, 0.720, 0.460, 0.210, 0.170, 0.190, 0.180, 0.090, 0.020,
                                               0.250, 0.045, 0.020),
                                   (0.025, 0.020, 0.020, 0.010, 0.008, 0.008, 0.005, 0.003, 0.0, 0.010, 0.0, 0.0))))
WOMENS_TEE = SpecTable("M", list(zip(_TEE_KEYS, (0.485, 0.650, 0.350, 0.180, 0.150, 0.165, 0.160, 0.085, 0.020,
                                                 0.215, 0.035, 0.020),
                                     (0.020, 0.015, 0.010, 0.005, 0.006, 0.007, 0.005, 0.003, 0.0, 0.008, 0.0,
                                      0.0))))


class GeneratorError(ValueError):
    pass


def _read_tee_spec(spec: dict, table: SpecTable, gid: str) -> dict:
    for key, _, _ in table.poms:
        if key not in spec:
            raise GeneratorError(f"{gid}: missing '{key}'")
        if not spec[key] > 0:
            raise GeneratorError(f"{gid}: '{key}' must be > 0")
    s = {k: float(spec[k]) for k in _TEE_KEYS}
    checks = ((s["armhole_depth"] > s["shoulder_drop"] + 0.02, "'armhole_depth' must exceed 'shoulder_drop' by at least 0.02 m"),
              (s["neck_width"] < s["shoulder_width"] - 0.02, "'neck_width' must be at least 0.02 m less than 'shoulder_width'"),
              (s["armhole_depth"] < s["body_length"] - 0.02, "'armhole_depth' must be at least 0.02 m less than 'body_length'"),
              (s["front_neck_drop"] < s["body_length"] - 0.02, "'front_neck_drop' must be at least 0.02 m less than 'body_length'"),
              (s["back_neck_drop"] < s["body_length"] - 0.02, "'back_neck_drop' must be at least 0.02 m less than 'body_length'"))
    for ok, msg in checks:
        if not ok:
            raise GeneratorError(f"{gid}: {msg}")
    return s


def _body_piece(pid: str, s: dict, neck_drop: float, center_angle: float) -> Piece:
    """Front or back: hem, right side, armhole, shoulder, neck, and their mirror images (test_tee.cpp:40-66)."""
    C, L, S, NW = s["chest_width"], s["body_length"], s["shoulder_width"], s["neck_width"]
    AD, SD = s["armhole_depth"], s["shoulder_drop"]
    hps = (NW / 2, L)
    sp = (S / 2, L - SD)
    ab = (C / 2, L - AD)
    cf = (0.0, L - neck_drop)
    armhole = Bezier(sp, (S / 2, L - SD - 0.5 * (AD - SD)), (S / 2 + 0.5 * (C / 2 - S / 2), L - AD), ab)
    neck = Bezier(cf, (0.55 * NW / 2, L - neck_drop), (NW / 2, L - 0.45 * neck_drop), hps)
    o = Outline()
    o.add("hem", Bezier.line((-C / 2, 0), (C / 2, 0)))
    o.add("side_r", Bezier.line((C / 2, 0), ab))
    o.add("armhole_r", armhole.reversed())
    o.add("shoulder_r", Bezier.line(sp, hps))
    o.add("neck_r", neck.reversed())
    o.add("neck_l", neck.mirrored_x())
    o.add("shoulder_l", Bezier.line((-hps[0], hps[1]), (-sp[0], sp[1])))
    o.add("armhole_l", armhole.mirrored_x())
    o.add("side_l", Bezier.line((-ab[0], ab[1]), (-C / 2, 0)))
    return Piece(pid, o, (0.0, 1.0), "body", Placement("Torso", center_angle, -L, 1, 0.06, None))


def _cap_right(s: dict, ch: float) -> Bezier:
    W, SL = s["bicep_width"], s["sleeve_length"]
    return Bezier((W, SL - ch), (0.75 * W, SL - 0.45 * ch), (0.40 * W, SL), (0.0, SL))


def _solve_cap_height(s: dict, armhole_length: float, gid: str) -> float:
    """Bisection on the cap height whose right cap curve is as long as the body's armhole (test_tee.cpp:72-91)."""
    def f(ch):
        return bezier_length(_cap_right(s, ch)) - armhole_length
    lo, hi = 0.02, s["sleeve_length"] - 0.02
    if hi <= lo or f(hi) < 0:
        raise GeneratorError(f"{gid}: no sleeve cap fits: 'sleeve_length' is too short for the armhole set by "
                             "'chest_width', 'shoulder_width', 'armhole_depth' and 'shoulder_drop'")
    if f(lo) > 0:
        raise GeneratorError(f"{gid}: no sleeve cap fits: 'bicep_width' is too large for the armhole set by "
                             "'chest_width', 'shoulder_width', 'armhole_depth' and 'shoulder_drop'")
    i = 0
    while i < 100 and hi - lo > 1e-10:
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
        i += 1
    return 0.5 * (lo + hi)


def _sleeve_piece(pid: str, s: dict, ch: float, region: str) -> Piece:
    SO, W, SL = s["sleeve_opening"], s["bicep_width"], s["sleeve_length"]
    cap = _cap_right(s, ch)
    u = (W, SL - ch)
    o = Outline()
    o.add("hem", Bezier.line((-SO, 0), (SO, 0)))
    o.add("underarm_r", Bezier.line((SO, 0), u))
    o.add("cap_r", cap)
    o.add("cap_l", cap.mirrored_x().reversed())
    o.add("underarm_l", Bezier.line((-u[0], u[1]), (-SO, 0)))
    return Piece(pid, o, (0.0, 1.0), "body", Placement(region, 0.0, SL, -1, 0.06, None))


def _seg_len(p: Piece, name: str) -> float:
    i = p.outline.find(name)
    return p.outline.segment_end(i) - p.outline.segment_start(i)


def generate_tee(spec: dict, table: SpecTable = MENS_TEE, gid: str = "test_tee") -> Pattern:
    """Front, back, two sleeves and the neck rib with their 13 seams (test_tee.cpp:180-218)."""
    s = _read_tee_spec(spec, table, gid)
    front = _body_piece("front", s, s["front_neck_drop"], 0.0)
    back = _body_piece("back", s, s["back_neck_drop"], math.pi)
    ch = _solve_cap_height(s, _seg_len(front, "armhole_r"), gid)
    sleeve_l = _sleeve_piece("sleeve_l", s, ch, "ArmLeft")
    sleeve_r = _sleeve_piece("sleeve_r", s, ch, "ArmRight")
    lf = _seg_len(front, "neck_r") + _seg_len(front, "neck_l")
    lb = _seg_len(back, "neck_r") + _seg_len(back, "neck_l")
    lr = lf + lb
    rib = rectangle_piece("neck_rib", lr, s["neck_rib_height"], (0.0, 1.0), "rib")
    rib.placement = Placement("Neck", -math.pi / 2, 0.0, 1, 0.0, lr / (2 * math.pi))
    e = segment_edge
    seams = [
        Seam("shoulder_wl", e(front, "shoulder_r", "shoulder_r", False), e(back, "shoulder_l", "shoulder_l", True)),
        Seam("shoulder_wr", e(front, "shoulder_l", "shoulder_l", False), e(back, "shoulder_r", "shoulder_r", True)),
        Seam("side_wl", e(front, "side_r", "side_r", False), e(back, "side_l", "side_l", True)),
        Seam("side_wr", e(front, "side_l", "side_l", False), e(back, "side_r", "side_r", True)),
        Seam("armhole_front_wl", e(front, "armhole_r", "armhole_r", False), e(sleeve_l, "cap_l", "cap_l", True)),
        Seam("armhole_back_wl", e(back, "armhole_l", "armhole_l", False), e(sleeve_l, "cap_r", "cap_r", True)),
        Seam("armhole_front_wr", e(front, "armhole_l", "armhole_l", False), e(sleeve_r, "cap_r", "cap_r", True)),
        Seam("armhole_back_wr", e(back, "armhole_r", "armhole_r", False), e(sleeve_r, "cap_l", "cap_l", True)),
        Seam("underarm_l", e(sleeve_l, "underarm_r", "underarm_r", False), e(sleeve_l, "underarm_l", "underarm_l", True)),
        Seam("underarm_r", e(sleeve_r, "underarm_r", "underarm_r", False), e(sleeve_r, "underarm_l", "underarm_l", True)),
        Seam("neck_front", EdgeRef("neck_rib", 0.0, lf, False), e(front, "neck_r", "neck_l", True)),
        Seam("neck_back", EdgeRef("neck_rib", lf, lr, False), e(back, "neck_r", "neck_l", True)),
        Seam("rib_close", e(rib, "right", "right", False), e(rib, "left", "left", True)),
    ]
    pattern = Pattern([front, back, sleeve_l, sleeve_r, rib], seams)
    issues = validate_pattern(pattern)
    if issues:
        raise GeneratorError(f"{gid}: these measurements give an invalid pattern ({issues[0][0]}: {issues[0][1]})")
    return pattern


def mens_tee(size: str = "M") -> Pattern:
    return generate_tee(resolve_spec(MENS_TEE, size), MENS_TEE, "test_tee")


def womens_tee(size: str = "M") -> Pattern:
    return generate_tee(resolve_spec(WOMENS_TEE, size), WOMENS_TEE, "test_tee_womens")


# ---------------------------------------------------------------------------------------------------------------
# Validation (validate.cpp, polygon.cpp).

def signed_area(poly) -> float:
    a = 0.0
    n = len(poly)
    for i in range(n):
        p, q = poly[i], poly[(i + 1) % n]
        a += p[0] * q[1] - q[0] * p[1]
    return 0.5 * a


def _orient(a, b, c) -> float:
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def _on_segment(a, b, p) -> bool:
    return min(a[0], b[0]) <= p[0] <= max(a[0], b[0]) and min(a[1], b[1]) <= p[1] <= max(a[1], b[1])


def _sign(v: float) -> int:
    return (v > 0) - (v < 0)


def segments_intersect(a, b, c, d) -> bool:
    o1, o2 = _sign(_orient(a, b, c)), _sign(_orient(a, b, d))
    o3, o4 = _sign(_orient(c, d, a)), _sign(_orient(c, d, b))
    if o1 != o2 and o3 != o4:
        return True
    return ((o1 == 0 and _on_segment(a, b, c)) or (o2 == 0 and _on_segment(a, b, d)) or
            (o3 == 0 and _on_segment(c, d, a)) or (o4 == 0 and _on_segment(c, d, b)))


def is_simple_polygon(poly) -> bool:
    n = len(poly)
    if n < 4:
        return True
    minx = [min(poly[i][0], poly[(i + 1) % n][0]) for i in range(n)]
    maxx = [max(poly[i][0], poly[(i + 1) % n][0]) for i in range(n)]
    order = sorted(range(n), key=lambda i: minx[i])
    for oi, i in enumerate(order):
        for oj in range(oi + 1, n):
            j = order[oj]
            if minx[j] > maxx[i]:
                break
            d = abs(i - j)
            if d == 1 or d == n - 1:
                continue
            if segments_intersect(poly[i], poly[(i + 1) % n], poly[j], poly[(j + 1) % n]):
                return False
    return True


def validate_pattern(pattern: Pattern) -> list:
    """[(rule, message)] for closed, simple, counter-clockwise outlines, non-zero grainlines, fabric slots, seam
    ranges, opposite traversal, equal lengths (2 mm + ease) and non-overlapping seam ranges."""
    eps, tol = 1e-9, 0.002
    issues = []
    for p in pattern.pieces:
        who = f"piece '{p.id}'"
        if not p.outline.is_closed():
            issues.append(("outline-closed", f"{who}: outline is not closed"))
        else:
            poly = p.outline.polyline(0.002)
            if not is_simple_polygon(poly):
                issues.append(("outline-simple", f"{who}: outline crosses itself"))
            elif signed_area(poly) <= 0:
                issues.append(("outline-ccw", f"{who}: outline must run counter-clockwise"))
        if _norm(*p.grainline) < eps:
            issues.append(("grainline", f"{who}: grainline is zero"))
        if not p.fabric_slot:
            issues.append(("fabric-slot", f"{who}: fabric slot is empty"))
    intervals: dict = {}
    for s in pattern.seams:
        who = f"seam '{s.id}'"
        ok = True
        for e in (s.a, s.b):
            piece = pattern.find_piece(e.piece)
            if piece is None:
                issues.append(("seam-piece", f"{who}: unknown piece '{e.piece}'"))
                ok = False
                continue
            if not (e.start >= 0 and e.start < e.end and e.end <= piece.outline.perimeter + eps):
                issues.append(("seam-range", f"{who}: edge on piece '{e.piece}' is outside [0, perimeter] or empty"))
                ok = False
                continue
            intervals.setdefault(e.piece, []).append((e.start, e.end, s.id))
        if s.a.reversed == s.b.reversed:
            issues.append(("seam-twisted", f"{who}: edges must be joined with opposite traversal (one reversed)"))
        if ok:
            la, lb = s.a.end - s.a.start, s.b.end - s.b.start
            if abs(la - lb) > tol + abs(s.ease):
                issues.append(("seam-length", f"{who}: edge lengths differ by {abs(la - lb) * 1000:.3f} mm"))
    for pid, lst in intervals.items():
        lst.sort(key=lambda x: x[0])
        for i in range(1, len(lst)):
            if lst[i][0] < lst[i - 1][1] - eps:
                issues.append(("seam-overlap", f"piece '{pid}': seams '{lst[i - 1][2]}' and '{lst[i][2]}' overlap"))
    return issues


# ---------------------------------------------------------------------------------------------------------------
# Simulation mesh (sim_mesh_builder.cpp).

QUALITY_SPACING = {"draft": 0.020, "standard": 0.010, "fine": 0.005}


def particle_distance(quality: str) -> float:
    return QUALITY_SPACING[quality.lower()]


class MeshError(ValueError):
    pass


_EPS = 1e-9


def _steps_ceil(length: float, h: float) -> int:
    return max(1, math.ceil(length / h - 1e-9))


def _steps_round(length: float, h: float) -> int:
    return max(1, int(math.floor(length / h + 0.5)))  # round half away from zero (lengths are positive)


class _Loop:
    """Sorted arc positions along one piece outline, deduplicated within 1e-9."""

    def __init__(self, perimeter: float):
        self.perimeter = perimeter
        self.positions: list = []

    def wrap(self, s: float) -> float:
        return 0.0 if s >= self.perimeter - _EPS else s

    def add(self, s: float) -> None:
        self.positions.append(self.wrap(s))

    def finish(self) -> None:
        out = []
        for s in sorted(self.positions):
            if not out or s - out[-1] > _EPS:
                out.append(s)
        self.positions = out

    def index_of(self, s: float) -> int:
        s = self.wrap(s)
        k = bisect.bisect_left(self.positions, s - _EPS)
        if k == len(self.positions) or abs(self.positions[k] - s) > _EPS:
            raise MeshError("internal: boundary sample not found")
        return k


def _corner_positions(o: Outline) -> list:
    """Arc positions of segment joints whose tangent turns by more than 10 degrees."""
    out = []
    n = len(o)
    for i in range(n):
        prev, nxt = o.segments[(i + n - 1) % n], o.segments[i]
        tin, tout = prev.derivative(1.0), nxt.derivative(0.0)
        if _norm(*tin) < 1e-14:
            tin = (prev.p3[0] - prev.p0[0], prev.p3[1] - prev.p0[1])
        if _norm(*tout) < 1e-14:
            tout = (nxt.p3[0] - nxt.p0[0], nxt.p3[1] - nxt.p0[1])
        ni, no = _norm(*tin), _norm(*tout)
        c = min(max((tin[0] / ni) * (tout[0] / no) + (tin[1] / ni) * (tout[1] / no), -1.0), 1.0)
        if math.acos(c) * 180.0 / math.pi > 10.0:
            out.append(o.segment_start(i))
    return out


def _distance_point_segment(p, a, b) -> float:
    abx, aby = b[0] - a[0], b[1] - a[1]
    l2 = abx * abx + aby * aby
    if l2 <= 0:
        return _norm(p[0] - a[0], p[1] - a[1])
    t = min(max(((p[0] - a[0]) * abx + (p[1] - a[1]) * aby) / l2, 0.0), 1.0)
    return _norm(p[0] - (a[0] + t * abx), p[1] - (a[1] + t * aby))


class _EdgeGrid:
    """Boundary edges bucketed into square cells (each edge in every cell its box, grown by one cell, touches);
    `nearest` looks in the query point's own cell only."""

    def __init__(self, poly: list, cell: float):
        self.poly, self.cell = poly, cell
        self.cells: dict = {}
        n = len(poly)
        for i in range(n):
            a, b = poly[i], poly[(i + 1) % n]
            x0, x1 = self._c(min(a[0], b[0]) - cell), self._c(max(a[0], b[0]) + cell)
            y0, y1 = self._c(min(a[1], b[1]) - cell), self._c(max(a[1], b[1]) + cell)
            for x in range(x0, x1 + 1):
                for y in range(y0, y1 + 1):
                    self.cells.setdefault((x, y), []).append(i)

    def _c(self, v: float) -> int:
        return math.floor(v / self.cell)

    def nearest(self, p) -> float:
        lst = self.cells.get((self._c(p[0]), self._c(p[1])))
        if lst is None:
            return math.inf
        n = len(self.poly)
        return min(_distance_point_segment(p, self.poly[i], self.poly[(i + 1) % n]) for i in lst)


def _lattice_points(poly: list, grain: tuple, h: float) -> list:
    """Grain-aligned hexagonal lattice points inside the polygon and at least h / 2 from its boundary: rows every
    sqrt(3)/2 h across the grain, odd rows shifted by h / 2, inside by the crossing parity along each row."""
    gn = _norm(*grain)
    u = (grain[0] / gn, grain[1] / gn)
    v = (-u[1], u[0])
    local = [(p[0] * u[0] + p[1] * u[1], p[0] * v[0] + p[1] * v[1]) for p in poly]
    umin = min(q[0] for q in local)
    umax = max(q[0] for q in local)
    vmin = min(q[1] for q in local)
    vmax = max(q[1] for q in local)
    row = math.sqrt(3.0) / 2.0 * h
    grid = _EdgeGrid(poly, h)
    out = []
    n = len(local)
    for j in range(math.floor(vmin / row) - 1, math.ceil(vmax / row) + 2):
        vy = j * row
        crossings = []
        for k in range(n):
            a, b = local[k], local[(k + 1) % n]
            if (a[1] > vy) != (b[1] > vy):
                crossings.append(a[0] + (vy - a[1]) * (b[0] - a[0]) / (b[1] - a[1]))
        crossings.sort()
        shift = 0.5 if j % 2 == 1 else 0.0
        for i in range(math.floor(umin / h) - 1, math.ceil(umax / h) + 2):
            ux = (i + shift) * h
            if bisect.bisect_left(crossings, ux) % 2 == 0:
                continue  # outside
            p = (ux * u[0] + vy * v[0], ux * u[1] + vy * v[1])
            if grid.nearest(p) < 0.5 * h:
                continue
            out.append(p)
    return out


def _tri_area(a, b, c) -> float:
    return 0.5 * ((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]))


def _smooth(pts: list, tris: np.ndarray, boundary_count: int, iterations: int) -> None:
    """Jacobi Laplacian smoothing of the interior vertices (mean of the previous pass's neighbours, sorted); a move
    is undone when it leaves an incident triangle with area <= 1e-14."""
    n = len(pts)
    nbrs = [set() for _ in range(n)]
    incident = [[] for _ in range(n)]
    for t, (a, b, c) in enumerate(tris.tolist()):
        for x, y, z in ((a, b, c), (b, c, a), (c, a, b)):
            incident[x].append(t)
            nbrs[x].add(y)
            nbrs[x].add(z)
    nb = [sorted(s) for s in nbrs]
    tl = tris.tolist()
    for _ in range(iterations):
        prev = list(pts)
        for i in range(boundary_count, n):
            if not nb[i]:
                continue
            mx = my = 0.0
            for k in nb[i]:
                mx += prev[k][0]
                my += prev[k][1]
            cnt = float(len(nb[i]))
            old = pts[i]
            pts[i] = (mx / cnt, my / cnt)
            for t in incident[i]:
                a, b, c = tl[t]
                if _tri_area(pts[a], pts[b], pts[c]) <= 1e-14:
                    pts[i] = old
                    break


def build_sim_mesh(pattern: Pattern, h: float) -> Mesh:
    """Pieces in pattern order: boundary samples first (seam sides sampled as a whole with the seam's shared step
    count, free spans between corners and seam ends at about h), then the lattice; triangles of the constrained
    Delaunay triangulation of the boundary polygon, counter-clockwise; seam pairs A's k-th sample with B's k-th."""
    issues = validate_pattern(pattern)
    if issues:
        raise MeshError("pattern is invalid:" + "".join(f" [{r}] {m};" for r, m in issues))
    pidx = {p.id: i for i, p in enumerate(pattern.pieces)}
    sides = [[] for _ in pattern.pieces]
    seam_steps = []
    for si, s in enumerate(pattern.seams):
        steps = _steps_ceil(max(s.a.end - s.a.start, s.b.end - s.b.start), h)
        seam_steps.append(steps)
        for e in (s.a, s.b):
            sides[pidx[e.piece]].append((si, e.start, e.end, e.reversed, steps))

    rest, vpiece, bnd, tris, tpiece, grains = [], [], [], [], [], []
    loops, offsets = [], []
    for pi, piece in enumerate(pattern.pieces):
        o = piece.outline
        loop = _Loop(o.perimeter)
        breaks = [0.0] + _corner_positions(o)
        for _, st, en, _, _ in sides[pi]:
            breaks += [st, en]
        breaks = sorted(loop.wrap(b) for b in breaks)
        uniq = []
        for b in breaks:
            if not uniq or b - uniq[-1] > _EPS:
                uniq.append(b)
        breaks = uniq
        for _, st, en, _, steps in sides[pi]:
            for k in range(steps + 1):
                loop.add(st + (en - st) * k / steps)
        for bi, s0 in enumerate(breaks):
            s1 = breaks[bi + 1] if bi + 1 < len(breaks) else loop.perimeter
            mid = 0.5 * (s0 + s1)
            if any(mid > st and mid < en for _, st, en, _, _ in sides[pi]):
                continue
            n = _steps_round(s1 - s0, h)
            for k in range(n):
                loop.add(s0 + (s1 - s0) * k / n)
        loop.finish()
        loops.append(loop)

        pts = [o.point_at(s) for s in loop.positions]
        nb = len(pts)
        gn = _norm(*piece.grainline)
        grain = (piece.grainline[0] / gn, piece.grainline[1] / gn)
        pts += _lattice_points(pts[:nb], grain, h)
        arr = np.asarray(pts, dtype=np.float64)
        if len(np.unique(arr, axis=0)) != len(arr):
            raise MeshError(f"piece '{piece.id}': duplicate mesh vertices")
        t = cdt.constrained_delaunay(arr, nb)
        _smooth(pts, t, nb, 3)

        off = len(rest)
        offsets.append(off)
        rest += pts
        vpiece += [pi] * len(pts)
        bnd += [1] * nb + [0] * (len(pts) - nb)
        for a, b, c in t.tolist():
            if not _tri_area(pts[a], pts[b], pts[c]) > 1e-12:
                raise MeshError(f"piece '{piece.id}': degenerate triangle")
        tris.append(t.astype(np.int64) + off)
        tpiece += [pi] * len(t)
        grains.append(grain)

    rest_a = np.asarray(rest, dtype=np.float64)
    tri_a = np.concatenate(tris).astype(np.uint32)
    area = np.zeros(len(rest_a))
    for a, b, c in tri_a.tolist():
        ar = _tri_area(rest[a], rest[b], rest[c]) / 3.0
        area[a] += ar
        area[b] += ar
        area[c] += ar

    pairs = []
    for si, s in enumerate(pattern.seams):
        n = seam_steps[si]

        def vertex_at(e: EdgeRef, k: int) -> int:
            pi = pidx[e.piece]
            kk = n - k if e.reversed else k
            return offsets[pi] + loops[pi].index_of(e.start + (e.end - e.start) * kk / n)
        for k in range(n + 1):
            pairs.append((vertex_at(s.a, k), vertex_at(s.b, k), si, k))
    return Mesh(rest=rest_a, vertex_piece=np.asarray(vpiece, dtype=np.uint32), triangles=tri_a,
                triangle_piece=np.asarray(tpiece, dtype=np.uint32), vertex_area=area,
                boundary=np.asarray(bnd, dtype=np.uint8), piece_grain=np.asarray(grains, dtype=np.float64),
                seam_pairs=np.asarray(pairs, dtype=np.uint32).reshape(-1, 4))


# ---------------------------------------------------------------------------------------------------------------
# Placement (placement.cpp:23-46).

@dataclasses.dataclass
class RegionCylinder:
    origin: tuple
    axis: tuple
    e0: tuple
    radius: float

    def e1(self) -> tuple:
        a, e = self.axis, self.e0
        return (a[1] * e[2] - a[2] * e[1], a[2] * e[0] - a[0] * e[2], a[0] * e[1] - a[1] * e[0])


def place_pattern(pattern: Pattern, mesh: Mesh, regions: dict) -> np.ndarray:
    """Each vertex wrapped onto its piece's region cylinder: angle center + sign x / R, axial position
    axial0 + sign y, radius R = wrap radius or region radius + offset."""
    cyl = []
    for p in pattern.pieces:
        if p.placement.region not in regions:
            raise ValueError(f"piece '{p.id}' needs body region {p.placement.region}, which this body does not "
                             "provide")
        cyl.append(regions[p.placement.region])
    out = np.zeros((mesh.n, 3))
    for v in range(mesh.n):
        pi = int(mesh.vertex_piece[v])
        pl, c = pattern.pieces[pi].placement, cyl[pi]
        R = pl.wrap_radius if pl.wrap_radius is not None else c.radius + pl.offset
        x, y = mesh.rest[v]
        theta = pl.center_angle + pl.axial_sign * x / R
        axial = pl.axial_at_origin + pl.axial_sign * y
        e1 = c.e1()
        ct, st = math.cos(theta), math.sin(theta)
        out[v] = [c.origin[k] + axial * c.axis[k] + R * (ct * c.e0[k] + st * e1[k]) for k in range(3)]
    return out
