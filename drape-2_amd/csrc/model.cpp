// model.cpp — scene descriptors -> host model (terms, particles, segments).
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <tuple>

#include "layout.hpp"

namespace dh {

Fabric fabricFromDesc(const drape_fabric& f) {
  Fabric o;
  o.stretchWarp = f.stretch_warp;
  o.stretchWeft = f.stretch_weft;
  o.stretchBias = f.stretch_bias;
  o.strainLimitWarp = f.strain_limit_warp;
  o.strainLimitWeft = f.strain_limit_weft;
  o.bendWarp = f.bend_warp;
  o.bendWeft = f.bend_weft;
  o.weight = f.weight;
  o.thickness = f.thickness;
  o.friction = f.friction;
  o.damping = f.damping;
  return o;
}

void bendingCoefficients(double bendWarp, double bendWeft, double& p, double& q) {
  const double floor = 0.05 * std::min(bendWarp, bendWeft);
  p = std::max((3.0 * bendWarp - bendWeft) / 2.0, floor);
  q = std::max(bendWeft, floor);
}

double edgeBendStiffness(double p, double q, const double* edge, const double* grain) {
  const double el = std::hypot(edge[0], edge[1]), gl = std::hypot(grain[0], grain[1]);
  const double dx = edge[0] / el, dy = edge[1] / el, ux = grain[0] / gl, uy = grain[1] / gl;
  const double c = dx * ux + dy * uy;
  const double s = dx * uy - dy * ux;
  return p * s * s + q * c * c;
}

double Model::hingeStiffness(const HingeTerm& h) const {
  double B = 0;
  for (const HingeSide& side : h.side) {
    const Fabric& f = fabrics[side.fabric];
    double p, q;
    bendingCoefficients(f.bendWarp, f.bendWeft, p, q);
    B += 0.5 * edgeBendStiffness(p, q, side.edge, side.grain);
  }
  return B * h.geom;
}

size_t Model::garmentIndex(const std::string& id) const {
  for (size_t i = 0; i < garments.size(); ++i)
    if (garments[i].id == id) return i;
  throw std::out_of_range("unknown garment '" + id + "'");
}

double Field::sample(const Vec3& p) const {
  if (values.empty()) return band;
  const double l[3] = {(p.x - origin[0]) / cell, (p.y - origin[1]) / cell, (p.z - origin[2]) / cell};
  int i[3];
  double t[3];
  for (int a = 0; a < 3; ++a) {
    if (!(l[a] >= 0.0 && l[a] <= dims[a] - 1)) return band;
    i[a] = std::min(static_cast<int>(std::floor(l[a])), dims[a] - 2);
    t[a] = l[a] - i[a];
  }
  double r = 0;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const double w = (dx ? t[0] : 1 - t[0]) * (dy ? t[1] : 1 - t[1]) * (dz ? t[2] : 1 - t[2]);
        r += w * values[(static_cast<size_t>(i[2] + dz) * dims[1] + (i[1] + dy)) * dims[0] + (i[0] + dx)];
      }
  return r;
}

namespace {

Settings settingsFromDesc(const drape_settings& s) {
  Settings o;
  for (int a = 0; a < 3; ++a) o.gravity[a] = s.gravity[a];
  o.collisionOffset = s.collision_offset;
  o.closingSpeed = s.assembly_closing_speed;
  o.extraDamping = s.assembly_extra_damping;
  o.dampingScale = s.damping_scale;
  o.limitScale = s.strain_limit_stiffness_scale;
  o.seamStiffness = s.seam_stiffness;
  o.frictionSlip = s.friction_slip_speed;
  o.maxCg = s.max_cg_iterations;
  o.cgTolerance = s.cg_tolerance;
  return o;
}

// Mesh edges, sorted by (min, max) vertex key and, within a key, by (triangle, corner) — the iteration order of the
// reference's std::map<edge, triangles> (sim_model.cpp:125-129).
struct EdgeTable {
  std::vector<std::pair<uint64_t, uint32_t>> e;  // key, triangle
  explicit EdgeTable(const drape_mesh_desc& m) {
    e.reserve(3 * size_t(m.n_triangles));
    for (uint32_t t = 0; t < m.n_triangles; ++t)
      for (int k = 0; k < 3; ++k) {
        const uint32_t a = m.triangles[3 * t + k], b = m.triangles[3 * t + (k + 1) % 3];
        e.emplace_back(key(a, b), t);
      }
    std::stable_sort(e.begin(), e.end(), [](const auto& p, const auto& q) { return p.first < q.first; });
  }
  static uint64_t key(uint32_t a, uint32_t b) {
    const uint32_t lo = std::min(a, b), hi = std::max(a, b);
    return (uint64_t(lo) << 32) | hi;
  }
  // The only triangle of a boundary edge, or -1 (sim_model.cpp:148-151).
  int64_t soleTriangle(uint32_t a, uint32_t b) const {
    const uint64_t k = key(a, b);
    auto it = std::lower_bound(e.begin(), e.end(), k, [](const auto& p, uint64_t v) { return p.first < v; });
    if (it == e.end() || it->first != k) return -1;
    if (it + 1 != e.end() && (it + 1)->first == k) return -1;
    return it->second;
  }
};

}  // namespace

Model Model::build(const drape_scene_desc* scenes, uint32_t nScenes) {
  if (!scenes && nScenes) throw std::invalid_argument("null scene list");
  if (nScenes == 0) throw std::invalid_argument("scene has no particles");
  Model m;
  uint32_t rowCursor = 0;
  for (uint32_t si = 0; si < nScenes; ++si) {
    const drape_scene_desc& sc = scenes[si];
    Segment seg;
    seg.particleBegin = static_cast<uint32_t>(m.x.size());
    seg.garmentBegin = static_cast<uint32_t>(m.garments.size());
    seg.settings = settingsFromDesc(sc.settings);
    if (sc.body && sc.body->values) {
      const drape_field_desc& f = *sc.body;
      for (int a = 0; a < 3; ++a) {
        seg.body.origin[a] = f.origin[a];
        seg.body.dims[a] = f.dims[a];
      }
      seg.body.cell = f.cell_size;
      seg.body.band = f.band;
      if (f.dims[0] < 2 || f.dims[1] < 2 || f.dims[2] < 2) throw std::invalid_argument("body field needs >= 2 samples per axis");
      const size_t count = size_t(f.dims[0]) * f.dims[1] * f.dims[2];
      seg.body.values.assign(f.values, f.values + count);
    }
    const uint32_t segIndex = static_cast<uint32_t>(m.segments.size());
    for (uint32_t gix = 0; gix < sc.n_garments; ++gix) {
      const drape_garment_desc& sg = sc.garments[gix];
      const drape_mesh_desc& mesh = sg.mesh;
      const std::string id = sg.id ? sg.id : "";
      if (!sg.initial_positions)
        throw std::invalid_argument("garment '" + id + "': one initial position per mesh vertex required");
      if (sg.n_piece_fabrics != mesh.n_pieces)
        throw std::invalid_argument("garment '" + id + "': one fabric per pattern piece required");
      // Every index the mesh carries must address its own arrays before anything is derived from it (an index past
      // them would read past the host arrays and write past the device term tables).
      for (uint32_t vtx = 0; vtx < mesh.n_vertices; ++vtx)
        if (mesh.vertex_piece[vtx] >= mesh.n_pieces)
          throw std::out_of_range("garment '" + id + "': vertex piece out of range");
      for (uint32_t t = 0; t < mesh.n_triangles; ++t) {
        if (mesh.triangle_piece[t] >= mesh.n_pieces)
          throw std::out_of_range("garment '" + id + "': triangle piece out of range");
        for (int k = 0; k < 3; ++k)
          if (mesh.triangles[3 * size_t(t) + k] >= mesh.n_vertices)
            throw std::out_of_range("garment '" + id + "': triangle corner out of range");
      }
      Garment g;
      g.id = id;
      g.offset = static_cast<uint32_t>(m.x.size());
      g.count = mesh.n_vertices;
      g.segment = segIndex;
      g.fabricOffset = static_cast<uint32_t>(m.fabrics.size());
      g.pieceCount = sg.n_piece_fabrics;
      const uint32_t gi = static_cast<uint32_t>(m.garments.size());
      const uint32_t off = g.offset;
      for (uint32_t k = 0; k < sg.n_piece_fabrics; ++k) m.fabrics.push_back(fabricFromDesc(sg.piece_fabric[k]));
      for (uint32_t vtx = 0; vtx < mesh.n_vertices; ++vtx) {
        const double* p = sg.initial_positions + 3 * size_t(vtx);
        m.x.push_back({static_cast<float>(p[0]), static_cast<float>(p[1]), static_cast<float>(p[2])});
        m.vertexArea.push_back(mesh.vertex_area[vtx]);
        m.rest.push_back(mesh.rest[2 * size_t(vtx)]);
        m.rest.push_back(mesh.rest[2 * size_t(vtx) + 1]);
        m.pFabric.push_back(g.fabricOffset + mesh.vertex_piece[vtx]);
        m.pGarment.push_back(gi);
        m.pSegment.push_back(segIndex);
      }
      const double* R = mesh.rest;
      auto corner = [&](uint32_t t, int k) { return mesh.triangles[3 * size_t(t) + k]; };
      auto restArea = [&](uint32_t t) {
        const uint32_t a = corner(t, 0), b = corner(t, 1), c = corner(t, 2);
        const double e1x = R[2 * b] - R[2 * a], e1y = R[2 * b + 1] - R[2 * a + 1];
        const double e2x = R[2 * c] - R[2 * a], e2y = R[2 * c + 1] - R[2 * a + 1];
        return 0.5 * std::abs(e1x * e2y - e1y * e2x);
      };
      auto grainOf = [&](uint32_t piece, double* out) {
        out[0] = mesh.piece_grain[2 * piece];
        out[1] = mesh.piece_grain[2 * piece + 1];
      };
      // Stretch: rest triangle in the piece's (grain, grain-perpendicular) frame, dmInv = Dm^-1 (sim_model.cpp:108-123).
      for (uint32_t t = 0; t < mesh.n_triangles; ++t) {
        const uint32_t piece = mesh.triangle_piece[t];
        double u[2];
        grainOf(piece, u);
        double q[3][2];
        for (int k = 0; k < 3; ++k) {
          const double px = R[2 * corner(t, k)], py = R[2 * corner(t, k) + 1];
          q[k][0] = px * u[0] + py * u[1];
          q[k][1] = -px * u[1] + py * u[0];
        }
        const double a = q[1][0] - q[0][0], b = q[2][0] - q[0][0], c = q[1][1] - q[0][1], d = q[2][1] - q[0][1];
        const double det = a * d - c * b;
        const double inv = 1.0 / det;
        StretchTerm st;
        for (int k = 0; k < 3; ++k) st.v[k] = off + corner(t, k);
        st.dmInv[0] = d * inv;
        st.dmInv[1] = -c * inv;
        st.dmInv[2] = -b * inv;
        st.dmInv[3] = a * inv;
        st.area = 0.5 * std::abs(det);
        st.fabric = g.fabricOffset + piece;
        m.stretch.push_back(st);
      }
      const EdgeTable edges(mesh);
      auto opposite = [&](uint32_t t, uint32_t e0, uint32_t e1) {
        for (int k = 0; k < 3; ++k)
          if (corner(t, k) != e0 && corner(t, k) != e1) return corner(t, k);
        return e0;
      };
      // Hinges: interior edges shared by exactly two triangles of one piece (sim_model.cpp:124-146).
      for (size_t i = 0; i < edges.e.size();) {
        size_t j = i;
        while (j < edges.e.size() && edges.e[j].first == edges.e[i].first) ++j;
        if (j - i == 2) {
          const uint32_t t0 = edges.e[i].second, t1 = edges.e[i + 1].second;
          if (mesh.triangle_piece[t0] == mesh.triangle_piece[t1]) {
            const uint32_t lo = uint32_t(edges.e[i].first >> 32), hi = uint32_t(edges.e[i].first & 0xffffffffu);
            const uint32_t piece = mesh.triangle_piece[t0];
            HingeTerm h;
            h.v[0] = off + opposite(t0, lo, hi);
            h.v[1] = off + opposite(t1, lo, hi);
            h.v[2] = off + lo;
            h.v[3] = off + hi;
            HingeSide side;
            side.fabric = g.fabricOffset + piece;
            side.edge[0] = R[2 * hi] - R[2 * lo];
            side.edge[1] = R[2 * hi + 1] - R[2 * lo + 1];
            grainOf(piece, side.grain);
            h.side[0] = h.side[1] = side;
            h.geom = (side.edge[0] * side.edge[0] + side.edge[1] * side.edge[1]) / (restArea(t0) + restArea(t1));
            m.hinges.push_back(h);
          }
        }
        i = j;
      }
      // Seam springs; a seam hinge follows each pair whose successor on the same seam spans boundary edges on
      // both sides (sim_model.cpp:147-175).
      auto pair = [&](uint32_t k, int c) { return mesh.seam_pairs[4 * size_t(k) + c]; };
      for (uint32_t k = 0; k < mesh.n_seam_pairs; ++k) {
        const uint32_t a = pair(k, 0), b = pair(k, 1);
        if (a >= mesh.n_vertices || b >= mesh.n_vertices) throw std::out_of_range("seam pair vertex out of range");
        SeamTerm s;
        s.a = off + a;
        s.b = off + b;
        s.garment = gi;
        const double* pa = sg.initial_positions + 3 * size_t(a);
        const double* pb = sg.initial_positions + 3 * size_t(b);
        s.startGap = static_cast<float>(std::sqrt((pa[0] - pb[0]) * (pa[0] - pb[0]) + (pa[1] - pb[1]) * (pa[1] - pb[1]) +
                                                  (pa[2] - pb[2]) * (pa[2] - pb[2])));
        m.seams.push_back(s);
        if (k + 1 >= mesh.n_seam_pairs || pair(k + 1, 2) != pair(k, 2)) continue;
        const uint32_t na = pair(k + 1, 0), nb = pair(k + 1, 1);
        const int64_t tA = edges.soleTriangle(a, na), tB = edges.soleTriangle(b, nb);
        if (tA < 0 || tB < 0) continue;
        const uint32_t ta = uint32_t(tA), tb = uint32_t(tB);
        HingeTerm h;
        h.v[0] = off + opposite(ta, a, na);
        h.v[1] = off + opposite(tb, b, nb);
        h.v[2] = off + a;
        h.v[3] = off + na;
        const uint32_t pieceA = mesh.triangle_piece[ta], pieceB = mesh.triangle_piece[tb];
        h.side[0].fabric = g.fabricOffset + pieceA;
        h.side[0].edge[0] = R[2 * na] - R[2 * a];
        h.side[0].edge[1] = R[2 * na + 1] - R[2 * a + 1];
        grainOf(pieceA, h.side[0].grain);
        h.side[1].fabric = g.fabricOffset + pieceB;
        h.side[1].edge[0] = R[2 * nb] - R[2 * b];
        h.side[1].edge[1] = R[2 * nb + 1] - R[2 * b + 1];
        grainOf(pieceB, h.side[1].grain);
        h.geom = (h.side[0].edge[0] * h.side[0].edge[0] + h.side[0].edge[1] * h.side[0].edge[1]) /
                 (restArea(ta) + restArea(tb));
        m.hinges.push_back(h);
      }
      m.garments.push_back(g);
    }
    seg.particleCount = static_cast<uint32_t>(m.x.size()) - seg.particleBegin;
    if (seg.particleCount == 0) throw std::invalid_argument("scene has no particles");
    seg.garmentCount = static_cast<uint32_t>(m.garments.size()) - seg.garmentBegin;
    seg.rowBegin = rowCursor;
    seg.rowCount = std::max<uint32_t>(dk::kTile, (seg.particleCount + dk::kTile - 1) / dk::kTile * dk::kTile);
    rowCursor += seg.rowCount;
    // Pins and forces index the particles of this scene (global indices when there is one scene).
    const size_t n = m.x.size();
    m.v.resize(n);
    m.fExt.resize(n);
    m.pinned.resize(n, 0);
    m.mass.resize(n, 0.0);
    for (uint32_t k = 0; k < sc.n_pins; ++k) {
      const uint32_t pidx = sc.pin_particles[k];
      if (pidx >= seg.particleCount) throw std::out_of_range("pin particle out of range");
      const uint32_t gidx = seg.particleBegin + pidx;
      const double* p = sc.pin_positions + 3 * size_t(k);
      const Vec3f pos{static_cast<float>(p[0]), static_cast<float>(p[1]), static_cast<float>(p[2])};
      m.pinned[gidx] = 1;
      m.pins.push_back({gidx, pos});
      m.x[gidx] = pos;
    }
    for (uint32_t k = 0; k < sc.n_forces; ++k) {
      const uint32_t pidx = sc.force_particles[k];
      if (pidx >= seg.particleCount) throw std::out_of_range("force particle out of range");
      Vec3& f = m.fExt[seg.particleBegin + pidx];
      f.x += sc.force_vectors[3 * size_t(k)];
      f.y += sc.force_vectors[3 * size_t(k) + 1];
      f.z += sc.force_vectors[3 * size_t(k) + 2];
    }
    m.segments.push_back(std::move(seg));
  }
  for (uint32_t i = 0; i < m.x.size(); ++i) m.updateMass(i);
  for (auto& h : m.hinges) h.K = m.hingeStiffness(h);
  m.contactForce.assign(m.x.size(), 0.0);
  return m;
}

}  // namespace dh
