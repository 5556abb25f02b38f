// oracle.cpp — FP64 CPU oracle for the implicit-Euler Newton–PCG cloth step.
//
// TEST INFRASTRUCTURE ONLY: the checker and the timed CPU baseline. Nothing in the product links or loads it.
//
// This is this project's own double-precision restatement of the reference's CPU path, written from the
// reference files as a behavioural specification (no reference source is copied; Eigen is replaced by the small
// vector/matrix helpers below). Each function names the reference lines it restates:
//   scene model and structure ........ libs/sim/src/sim_model.cpp:12-274
//   stretch strain ................... libs/sim/src/stretch.hpp:15-49
//   stretch / hinge / seam / friction  libs/sim/src/energies.cpp:10-110
//   dihedral angle and bending ....... libs/sim/src/bending.hpp:29-51, bending.cpp:7-18
//   body field ....................... libs/core/src/collision_field.cpp:8-41
//   post-solve push-out .............. libs/sim/src/collision.cpp:5-34
//   solver (evaluate, PCG, Newton) ... libs/sim/src/cpu_solver.cpp:34-355
//   GPU gather lists ................. libs/sim/src/gpu/gpu_solver_core.cpp:184-224
//
// Bit-level parity with the reference is unpinned wherever the reference calls Eigen (LDLT inverse, 2x2 inverse,
// reduction order); the pins are the closed forms, finite-difference checks and physics known-answer tests in
// tests/test_oracle_*.py.
//
// Threading: term evaluation, SpMV and per-particle loops may run on several threads; every accumulation keeps
// the serial order, so results are bit-identical for any thread count.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------------------------------------------
// Small FP64 vector / matrix helpers (Eigen replacement).

struct V3 {
  double x = 0, y = 0, z = 0;
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double sqnorm(V3 a) { return dot(a, a); }
inline double norm(V3 a) { return std::sqrt(sqnorm(a)); }
inline bool isZero(V3 a) { return a.x == 0.0 && a.y == 0.0 && a.z == 0.0; }

struct V3f {
  float x = 0, y = 0, z = 0;
};
inline V3 toD(V3f a) { return {a.x, a.y, a.z}; }
inline V3f toF(V3 a) { return {static_cast<float>(a.x), static_cast<float>(a.y), static_cast<float>(a.z)}; }
inline V3f operator+(V3f a, V3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3f operator-(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3f operator*(float s, V3f a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3f operator*(V3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dotf(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float normf(V3f a) { return std::sqrt(dotf(a, a)); }
inline bool allFinite(V3f a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

struct M3 {  // row-major
  double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double& operator()(int r, int c) { return a[3 * r + c]; }
  double operator()(int r, int c) const { return a[3 * r + c]; }
};
inline M3 identity() {
  M3 m;
  m.a[0] = m.a[4] = m.a[8] = 1.0;
  return m;
}
inline M3 outer(V3 u, V3 v) {
  M3 m;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m(r, c) = u[r] * v[c];
  return m;
}
inline M3 operator*(double s, const M3& m) {
  M3 o;
  for (int k = 0; k < 9; ++k) o.a[k] = s * m.a[k];
  return o;
}
inline M3 operator+(const M3& x, const M3& y) {
  M3 o;
  for (int k = 0; k < 9; ++k) o.a[k] = x.a[k] + y.a[k];
  return o;
}
inline M3 operator-(const M3& x, const M3& y) {
  M3 o;
  for (int k = 0; k < 9; ++k) o.a[k] = x.a[k] - y.a[k];
  return o;
}
inline M3& operator+=(M3& x, const M3& y) {
  for (int k = 0; k < 9; ++k) x.a[k] += y.a[k];
  return x;
}
inline M3& operator-=(M3& x, const M3& y) {
  for (int k = 0; k < 9; ++k) x.a[k] -= y.a[k];
  return x;
}
inline V3 operator*(const M3& m, V3 v) {
  return {m(0, 0) * v.x + m(0, 1) * v.y + m(0, 2) * v.z, m(1, 0) * v.x + m(1, 1) * v.y + m(1, 2) * v.z,
          m(2, 0) * v.x + m(2, 1) * v.y + m(2, 2) * v.z};
}

// 2x2 matrix stored as (0,0), (1,0), (0,1), (1,1) — column-major, as the GPU layout packs dmInv.
struct M2 {
  double d[4] = {1, 0, 0, 1};
  double operator()(int r, int c) const { return d[r + 2 * c]; }
};

// ---------------------------------------------------------------------------------------------------------------
// Energy terms.

template <int N>
struct TermEval {
  double energy = 0;
  V3 grad[N];
  M3 hess[N][N];
};

// Anisotropic strain of one triangle (stretch.hpp:15-49). which: 0 warp |f_u|-1, 1 weft |f_v|-1, 2 bias n_u.n_v.
inline void evalStretch(const V3* x, const M2& dmInv, int which, double& C, V3* grad) {
  const V3 e1 = x[1] - x[0];
  const V3 e2 = x[2] - x[0];
  const V3 fu = e1 * dmInv(0, 0) + e2 * dmInv(1, 0);
  const V3 fv = e1 * dmInv(0, 1) + e2 * dmInv(1, 1);
  const double lu = norm(fu), lv = norm(fv);
  const double tiny = 1e-12;
  V3 dfu, dfv;
  C = 0;
  if (which == 0) {
    if (lu > tiny) {
      C = lu - 1.0;
      dfu = fu / lu;
    }
  } else if (which == 1) {
    if (lv > tiny) {
      C = lv - 1.0;
      dfv = fv / lv;
    }
  } else if (lu > tiny && lv > tiny) {
    const V3 nu = fu / lu, nv = fv / lv;
    const double c = dot(nu, nv);
    C = c;
    dfu = (nv - c * nu) / lu;
    dfv = (nu - c * nv) / lv;
  }
  grad[1] = dfu * dmInv(0, 0) + dfv * dmInv(0, 1);
  grad[2] = dfu * dmInv(1, 0) + dfv * dmInv(1, 1);
  grad[0] = -(grad[1] + grad[2]);
}

// Coefficient of vertex `local` in f (column) = sum_k d_k x_k (energies.cpp:13-18).
inline double vertexCoefficient(const M2& dmInv, int column, int local) {
  if (local == 1) return dmInv(0, column);
  if (local == 2) return dmInv(1, column);
  return -(dmInv(0, column) + dmInv(1, column));
}

// 1/2 K C^2 with Gauss-Newton Hessian K g g^T (energies.cpp:20-27).
template <int N>
inline void addQuadratic(TermEval<N>& out, double K, double C, const V3* g) {
  out.energy += 0.5 * K * C * C;
  for (int i = 0; i < N; ++i) {
    out.grad[i] += K * C * g[i];
    for (int j = 0; j < N; ++j) out.hess[i][j] += outer(K * g[i], g[j]);
  }
}

// Tension stiffness K C d_i d_j (I - n n^T) / |f| for C > 0 (energies.cpp:29-39).
inline void addTension(TermEval<3>& out, double K, double C, const M2& dmInv, int column, V3 f) {
  if (C <= 0) return;
  const double len = norm(f);
  if (len < 1e-12) return;
  const V3 n = f / len;
  const M3 P = (K * C / len) * (identity() - outer(n, n));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out.hess[i][j] += (vertexCoefficient(dmInv, column, i) * vertexCoefficient(dmInv, column, j)) * P;
}

struct Fabric {
  double stretchWarp = 0, stretchWeft = 0, stretchBias = 0, strainLimitWarp = 0, strainLimitWeft = 0;
  double bendWarp = 0, bendWeft = 0, weight = 0, thickness = 0, friction = 0, damping = 0;
};
inline Fabric fromDesc(const drape_fabric& f) {
  return {f.stretch_warp, f.stretch_weft, f.stretch_bias, f.strain_limit_warp, f.strain_limit_weft, f.bend_warp,
          f.bend_weft,    f.weight,       f.thickness,    f.friction,          f.damping};
}

struct StretchTerm {
  std::array<uint32_t, 3> v{};
  M2 dmInv;
  double area = 0;
  uint32_t fabric = 0;
};

// Warp, weft, bias and strain-limit energies (energies.cpp:40-62).
inline void evalStretchTerm(const StretchTerm& t, const V3* x, const Fabric& f, double limitScale, TermEval<3>& out) {
  const double k[3] = {f.stretchWarp, f.stretchWeft, f.stretchBias};
  const double limit[2] = {f.strainLimitWarp, f.strainLimitWeft};
  const V3 e1 = x[1] - x[0], e2 = x[2] - x[0];
  const V3 fcol[2] = {e1 * t.dmInv(0, 0) + e2 * t.dmInv(1, 0), e1 * t.dmInv(0, 1) + e2 * t.dmInv(1, 1)};
  V3 g[3];
  double C = 0;
  for (int which = 0; which < 3; ++which) {
    if (k[which] <= 0) continue;
    evalStretch(x, t.dmInv, which, C, g);
    const double K = k[which] * t.area;
    addQuadratic<3>(out, K, C, g);
    if (which == 2) continue;
    addTension(out, K, C, t.dmInv, which, fcol[which]);
    if (limit[which] > 0 && C > limit[which]) {
      const double KL = limitScale * K;
      const double excess = C - limit[which];
      addQuadratic<3>(out, KL, excess, g);
      addTension(out, KL, excess, t.dmInv, which, fcol[which]);
    }
  }
}

// Dihedral angle and its gradient, Bridson–Marino–Fedkiw 2003 with the sign flipped (bending.hpp:29-51).
inline void evalDihedral(const V3* x, double& theta, V3* grad) {
  const V3 n1 = cross(x[0] - x[2], x[0] - x[3]);
  const V3 n2 = cross(x[1] - x[3], x[1] - x[2]);
  const V3 E = x[3] - x[2];
  const double n1sq = sqnorm(n1), n2sq = sqnorm(n2), elen = norm(E);
  if (n1sq < 1e-24 || n2sq < 1e-24 || elen < 1e-12) {
    theta = 0;
    for (int k = 0; k < 4; ++k) grad[k] = V3{};
    return;
  }
  const V3 e = E / elen;
  const V3 m1 = n1 / std::sqrt(n1sq), m2 = n2 / std::sqrt(n2sq);
  theta = std::atan2(dot(cross(m1, m2), e), dot(m1, m2));
  const V3 a = n1 / n1sq, b = n2 / n2sq;
  grad[0] = -elen * a;
  grad[1] = -elen * b;
  grad[2] = -((dot(x[0] - x[3], E) / elen) * a + (dot(x[1] - x[3], E) / elen) * b);
  grad[3] = (dot(x[0] - x[2], E) / elen) * a + (dot(x[1] - x[2], E) / elen) * b;
}

struct HingeTerm {
  std::array<uint32_t, 4> v{};  // opposite A, opposite B, edge start, edge end
  double K = 0;
  double restAngle = 0;
};

// 1/2 K (theta - theta0)^2, Gauss-Newton Hessian (energies.cpp:64-69).
inline void evalHingeTerm(const HingeTerm& t, const V3* x, TermEval<4>& out) {
  double theta = 0;
  V3 g[4];
  evalDihedral(x, theta, g);
  addQuadratic<4>(out, t.K, theta - t.restAngle, g);
}

// One-sided seam spring (energies.cpp:71-91).
inline void evalSeamTerm(const V3* x, double target, double K, TermEval<2>& out) {
  const V3 diff = x[0] - x[1];
  const double d = norm(diff);
  if (d <= target) return;
  M3 block;
  if (d < 1e-12) {
    block = K * identity();
  } else {
    const V3 n = diff / d;
    const double ext = d - target;
    out.energy += 0.5 * K * ext * ext;
    out.grad[0] += K * ext * n;
    out.grad[1] -= K * ext * n;
    const M3 nn = outer(n, n);
    block = K * (nn + (ext / d) * (identity() - nn));
  }
  out.hess[0][0] += block;
  out.hess[1][1] += block;
  out.hess[0][1] -= block;
  out.hess[1][0] -= block;
}

// IPC smoothed Coulomb friction (energies.cpp:93-110).
inline void evalFrictionTerm(V3 x, V3 x0, V3 n, double muLambda, double eps, TermEval<1>& out) {
  if (muLambda <= 0 || eps <= 0) return;
  const M3 P = identity() - outer(n, n);
  const V3 u = P * (x - x0);
  const double s = norm(u);
  if (s < eps) {
    out.energy += muLambda * (-s * s * s / (3 * eps * eps) + s * s / eps + eps / 3);
    const double f1OverS = 2 / eps - s / (eps * eps);
    out.grad[0] += muLambda * f1OverS * u;
    out.hess[0][0] += (muLambda * f1OverS) * P;
    if (s > 0) out.hess[0][0] -= (muLambda / (eps * eps * s)) * outer(u, u);
  } else {
    const V3 dir = u / s;
    out.energy += muLambda * s;
    out.grad[0] += muLambda * dir;
    out.hess[0][0] += (muLambda / s) * (P - outer(dir, dir));
  }
}

// p = (3 Bwarp - Bweft)/2, q = Bweft, each >= 0.05 min (bending.cpp:7-10).
inline void bendingCoefficients(double bendWarp, double bendWeft, double& p, double& q) {
  const double floor = 0.05 * std::min(bendWarp, bendWeft);
  p = std::max((3.0 * bendWarp - bendWeft) / 2.0, floor);
  q = std::max(bendWeft, floor);
}

// B(phi) = p sin^2 + q cos^2 of the angle between rest edge and grain (bending.cpp:12-18).
inline double edgeBendStiffness(double p, double q, const double* edge, const double* grain) {
  const double dn = std::sqrt(edge[0] * edge[0] + edge[1] * edge[1]);
  const double un = std::sqrt(grain[0] * grain[0] + grain[1] * grain[1]);
  const double d0 = edge[0] / dn, d1 = edge[1] / dn, u0 = grain[0] / un, u1 = grain[1] / un;
  const double cosPhi = d0 * u0 + d1 * u1;
  const double sinPhi = d0 * u1 - d1 * u0;
  return p * sinPhi * sinPhi + q * cosPhi * cosPhi;
}

// ---------------------------------------------------------------------------------------------------------------
// Body field (collision_field.cpp:8-41) and push-out helpers (collision.cpp:5-34).

struct Field {
  V3 origin;
  double cellSize = 0.004;
  int dims[3] = {0, 0, 0};
  double band = 0.02;
  std::vector<float> values;

  double sample(V3 p) const {
    if (values.empty()) return band;
    const V3 local = (p - origin) / cellSize;
    int i[3];
    double t[3];
    for (int a = 0; a < 3; ++a) {
      const double maxCoord = dims[a] - 1;
      if (!(local[a] >= 0.0 && local[a] <= maxCoord)) return band;
      i[a] = std::min(static_cast<int>(std::floor(local[a])), dims[a] - 2);
      t[a] = local[a] - i[a];
    }
    double result = 0;
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const double w = (dx ? t[0] : 1 - t[0]) * (dy ? t[1] : 1 - t[1]) * (dz ? t[2] : 1 - t[2]);
          const size_t idx = (static_cast<size_t>(i[2] + dz) * dims[1] + (i[1] + dy)) * dims[0] + (i[0] + dx);
          result += w * static_cast<double>(values[idx]);
        }
    return result;
  }

  V3 gradient(V3 p) const {
    V3 g;
    for (int a = 0; a < 3; ++a) {
      V3 d;
      d[a] = cellSize;
      g[a] = (sample(p + d) - sample(p - d)) / (2 * cellSize);
    }
    const double n = norm(g);
    return n < 1e-9 ? V3{} : g / n;
  }
};

inline Field fieldFromDesc(const drape_field_desc* f) {
  Field out;
  if (!f) return out;
  out.origin = {f->origin[0], f->origin[1], f->origin[2]};
  out.cellSize = f->cell_size;
  for (int a = 0; a < 3; ++a) out.dims[a] = f->dims[a];
  out.band = f->band;
  const size_t count = static_cast<size_t>(f->dims[0]) * f->dims[1] * f->dims[2];
  if (count && f->values) out.values.assign(f->values, f->values + count);
  return out;
}

inline bool isZeroF(V3f a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

// collideBody (collision.cpp:5-22).
inline void collideBody(const Field& body, float radius, float friction, V3f xPrev, V3f& x) {
  const V3 p = toD(x);
  const double phi = body.sample(p);
  if (phi >= radius) return;
  const V3f n = toF(body.gradient(p));
  if (isZeroF(n)) return;
  const float depth = static_cast<float>(radius - phi);
  x = x + depth * n;
  const V3f delta = x - xPrev;
  const V3f tangential = delta - dotf(delta, n) * n;
  const float t = normf(tangential);
  if (t <= friction * depth) {
    x = x - tangential;
  } else if (t > 0) {
    x = x - tangential * (friction * depth / t);
  }
}

// landOnBody (collision.cpp:24-34).
inline V3f landOnBody(const Field& body, float radius, V3f from, V3f to) {
  auto clear = [&](float t) { return body.sample(toD(from + t * (to - from))) >= radius; };
  if (!clear(0)) return from;
  float lo = 0, hi = 1;
  for (int i = 0; i < 40; ++i) {
    const float mid = 0.5f * (lo + hi);
    (clear(mid) ? lo : hi) = mid;
  }
  return from + lo * (to - from);
}

// Inverse of a symmetric 3x3 block; false when it is not positive definite. The reference uses Eigen's LDLT
// (cpu_solver.cpp:176-177); for the positive definite blocks this solver produces, both give the unique inverse.
inline bool invertSpd3(const M3& A, M3& inv) {
  const double a00 = A(0, 0), a10 = A(1, 0), a20 = A(2, 0), a11 = A(1, 1), a21 = A(2, 1), a22 = A(2, 2);
  const double d0 = a00;
  if (!(d0 > 0)) return false;
  const double l10 = a10 / d0, l20 = a20 / d0;
  const double d1 = a11 - l10 * l10 * d0;
  if (!(d1 > 0)) return false;
  const double l21 = (a21 - l20 * l10 * d0) / d1;
  const double d2 = a22 - l20 * l20 * d0 - l21 * l21 * d1;
  if (!(d2 > 0)) return false;
  for (int k = 0; k < 3; ++k) {
    const double e0 = k == 0, e1 = k == 1, e2 = k == 2;
    const double y0 = e0, y1 = e1 - l10 * y0, y2 = e2 - l20 * y0 - l21 * y1;
    const double z0 = y0 / d0, z1 = y1 / d1, z2 = y2 / d2;
    const double x2 = z2, x1 = z1 - l21 * x2, x0 = z0 - l10 * x1 - l20 * x2;
    inv(0, k) = x0;
    inv(1, k) = x1;
    inv(2, k) = x2;
  }
  return true;
}

// ---------------------------------------------------------------------------------------------------------------
// Scene model (sim_model.cpp).

constexpr double kContactReach = 0.01;   // sim_model.hpp:14
constexpr double kContactSnap = 0.0005;  // sim_model.hpp:15
enum Phase { Assembling = 0, Settling = 1 };

struct Settings {
  V3 gravity{0, -9.81, 0};
  double collisionOffset = 0.003, assemblyClosingSpeed = 0.25, assemblyExtraDamping = 10.0, dampingScale = 5.0;
  double strainLimitStiffnessScale = 100.0, seamStiffness = 1e4, frictionSlipSpeed = 3e-4;
  int maxCgIterations = 50;
  double cgTolerance = 1e-2;
};

struct Garment {
  std::string id;
  uint32_t offset = 0, count = 0;
  int phase = Settling;
  double assemblyTime = 0;
  uint32_t fabricOffset = 0, pieceCount = 0;
};
struct HingeSide {
  uint32_t fabric = 0;
  double edgeDir[2] = {1, 0}, grain[2] = {0, 1};
};
struct SeamSpring {
  uint32_t a = 0, b = 0, garment = 0;
  float startGap = 0;
};

struct Model {
  Settings settings;
  Field body;
  bool hasBody = false;
  std::vector<Garment> garments;
  std::vector<Fabric> fabrics;
  std::vector<V3f> x, v;
  std::vector<double> mass, vertexArea;
  std::vector<V3> fExt;
  std::vector<uint32_t> pFabric, pGarment;
  std::vector<uint8_t> pinned;
  std::vector<std::pair<uint32_t, V3f>> pins;
  std::vector<StretchTerm> stretch;
  std::vector<HingeTerm> hinges;
  std::vector<std::array<HingeSide, 2>> hingeSides;
  std::vector<double> hingeGeom;
  std::vector<SeamSpring> seams;
  std::vector<uint32_t> rowStart, colIndex, diagSlot;
  std::vector<std::array<uint32_t, 9>> stretchSlots;
  std::vector<std::array<uint32_t, 16>> hingeSlots;
  std::vector<std::array<uint32_t, 4>> seamSlots;
  std::vector<double> contactForce;

  size_t garmentIndex(const std::string& id) const {
    for (size_t i = 0; i < garments.size(); ++i)
      if (garments[i].id == id) return i;
    throw std::out_of_range("unknown garment '" + id + "'");
  }
  double contactRadius(size_t i) const { return settings.collisionOffset + 0.5 * fabrics[pFabric[i]].thickness; }
  void updateMass(uint32_t i) { mass[i] = std::max(vertexArea[i] * fabrics[pFabric[i]].weight, 1e-12); }

  // sim_model.cpp:20-27
  double hingeStiffness(size_t hgi) const {
    double B = 0;
    for (const auto& side : hingeSides[hgi]) {
      const Fabric& f = fabrics[side.fabric];
      double p, q;
      bendingCoefficients(f.bendWarp, f.bendWeft, p, q);
      B += 0.5 * edgeBendStiffness(p, q, side.edgeDir, side.grain);
    }
    return B * hingeGeom[hgi];
  }
  // sim_model.cpp:29-33
  double seamTarget(const SeamSpring& sp) const {
    const Garment& g = garments[sp.garment];
    if (g.phase != Assembling) return 0.0;
    return std::max(0.0, static_cast<double>(sp.startGap) - settings.assemblyClosingSpeed * g.assemblyTime);
  }
  // sim_model.cpp:35-39
  uint32_t slot(uint32_t row, uint32_t col) const {
    const auto begin = colIndex.begin() + rowStart[row];
    const auto end = colIndex.begin() + rowStart[row + 1];
    return static_cast<uint32_t>(std::lower_bound(begin, end, col) - colIndex.begin());
  }

  // sim_model.cpp:41-74
  void buildStructure() {
    const size_t n = x.size();
    std::vector<std::vector<uint32_t>> cols(n);
    auto couple = [&](const uint32_t* vs, int m) {
      for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) cols[vs[a]].push_back(vs[b]);
    };
    for (size_t i = 0; i < n; ++i) cols[i].push_back(static_cast<uint32_t>(i));
    for (const auto& t : stretch) couple(t.v.data(), 3);
    for (const auto& h : hinges) couple(h.v.data(), 4);
    for (const auto& s : seams) {
      const uint32_t ab[2] = {s.a, s.b};
      couple(ab, 2);
    }
    rowStart.assign(n + 1, 0);
    colIndex.clear();
    for (size_t i = 0; i < n; ++i) {
      std::sort(cols[i].begin(), cols[i].end());
      cols[i].erase(std::unique(cols[i].begin(), cols[i].end()), cols[i].end());
      colIndex.insert(colIndex.end(), cols[i].begin(), cols[i].end());
      rowStart[i + 1] = static_cast<uint32_t>(colIndex.size());
    }
    diagSlot.resize(n);
    for (uint32_t i = 0; i < n; ++i) diagSlot[i] = slot(i, i);
    auto slotsFor = [&](const uint32_t* vs, int m, uint32_t* out) {
      for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) out[a * m + b] = slot(vs[a], vs[b]);
    };
    stretchSlots.resize(stretch.size());
    for (size_t k = 0; k < stretch.size(); ++k) slotsFor(stretch[k].v.data(), 3, stretchSlots[k].data());
    hingeSlots.resize(hinges.size());
    for (size_t k = 0; k < hinges.size(); ++k) slotsFor(hinges[k].v.data(), 4, hingeSlots[k].data());
    seamSlots.resize(seams.size());
    for (size_t k = 0; k < seams.size(); ++k) {
      const uint32_t ab[2] = {seams[k].a, seams[k].b};
      slotsFor(ab, 2, seamSlots[k].data());
    }
    contactForce.assign(n, 0.0);
  }

  // sim_model.cpp:76-197
  static Model build(const drape_scene_desc& scene) {
    Model next;
    const drape_settings& s = scene.settings;
    next.settings.gravity = {s.gravity[0], s.gravity[1], s.gravity[2]};
    next.settings.collisionOffset = s.collision_offset;
    next.settings.assemblyClosingSpeed = s.assembly_closing_speed;
    next.settings.assemblyExtraDamping = s.assembly_extra_damping;
    next.settings.dampingScale = s.damping_scale;
    next.settings.strainLimitStiffnessScale = s.strain_limit_stiffness_scale;
    next.settings.seamStiffness = s.seam_stiffness;
    next.settings.frictionSlipSpeed = s.friction_slip_speed;
    next.settings.maxCgIterations = s.max_cg_iterations;
    next.settings.cgTolerance = s.cg_tolerance;
    if (scene.body) {
      next.body = fieldFromDesc(scene.body);
      next.hasBody = true;
    }
    for (uint32_t gix = 0; gix < scene.n_garments; ++gix) {
      const drape_garment_desc& sg = scene.garments[gix];
      const drape_mesh_desc& mesh = sg.mesh;
      const std::string id = sg.id ? sg.id : "";
      if (!sg.initial_positions) {
        throw std::invalid_argument("garment '" + id + "': one initial position per mesh vertex required");
      }
      if (sg.n_piece_fabrics != mesh.n_pieces) {
        throw std::invalid_argument("garment '" + id + "': one fabric per pattern piece required");
      }
      Garment g;
      g.id = id;
      g.offset = static_cast<uint32_t>(next.x.size());
      g.count = mesh.n_vertices;
      g.fabricOffset = static_cast<uint32_t>(next.fabrics.size());
      g.pieceCount = sg.n_piece_fabrics;
      const auto gi = static_cast<uint32_t>(next.garments.size());
      const uint32_t off = g.offset;
      for (uint32_t k = 0; k < sg.n_piece_fabrics; ++k) next.fabrics.push_back(fromDesc(sg.piece_fabric[k]));
      auto init = [&](uint32_t vtx) {
        return V3{sg.initial_positions[3 * vtx], sg.initial_positions[3 * vtx + 1], sg.initial_positions[3 * vtx + 2]};
      };
      auto rest = [&](uint32_t vtx, int c) { return mesh.rest[2 * vtx + c]; };
      for (uint32_t vtx = 0; vtx < mesh.n_vertices; ++vtx) {
        next.x.push_back(toF(init(vtx)));
        next.vertexArea.push_back(mesh.vertex_area[vtx]);
        next.pFabric.push_back(g.fabricOffset + mesh.vertex_piece[vtx]);
        next.pGarment.push_back(gi);
      }
      auto tri = [&](uint32_t t, int k) { return mesh.triangles[3 * t + k]; };
      auto triArea = [&](uint32_t t) {
        const double ax = rest(tri(t, 1), 0) - rest(tri(t, 0), 0), ay = rest(tri(t, 1), 1) - rest(tri(t, 0), 1);
        const double bx = rest(tri(t, 2), 0) - rest(tri(t, 0), 0), by = rest(tri(t, 2), 1) - rest(tri(t, 0), 1);
        return 0.5 * std::abs(ax * by - ay * bx);
      };
      // Stretch terms in each piece's grain frame (sim_model.cpp:107-123).
      for (uint32_t t = 0; t < mesh.n_triangles; ++t) {
        const uint32_t piece = mesh.triangle_piece[t];
        const double ux = mesh.piece_grain[2 * piece], uy = mesh.piece_grain[2 * piece + 1];
        const double vx = -uy, vy = ux;
        double rr[3][2];
        for (int k = 0; k < 3; ++k) {
          const double px = rest(tri(t, k), 0), py = rest(tri(t, k), 1);
          rr[k][0] = px * ux + py * uy;
          rr[k][1] = px * vx + py * vy;
        }
        // dm = [[r1-r0].u, (r2-r0).u; (r1-r0).v, (r2-r0).v]; inverse by the closed 2x2 form.
        const double m00 = rr[1][0] - rr[0][0], m01 = rr[2][0] - rr[0][0];
        const double m10 = rr[1][1] - rr[0][1], m11 = rr[2][1] - rr[0][1];
        const double det = m00 * m11 - m10 * m01;
        const double invdet = 1.0 / det;
        StretchTerm st;
        for (int k = 0; k < 3; ++k) st.v[k] = off + tri(t, k);
        st.dmInv.d[0] = m11 * invdet;   // (0,0)
        st.dmInv.d[1] = -m10 * invdet;  // (1,0)
        st.dmInv.d[2] = -m01 * invdet;  // (0,1)
        st.dmInv.d[3] = m00 * invdet;   // (1,1)
        st.area = 0.5 * std::abs(det);
        st.fabric = g.fabricOffset + piece;
        next.stretch.push_back(st);
      }
      // Hinges on interior edges shared by two triangles of one piece, in (min, max) edge-key order
      // (sim_model.cpp:124-146).
      std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> edgeTris;
      for (uint32_t t = 0; t < mesh.n_triangles; ++t)
        for (int k = 0; k < 3; ++k) edgeTris[std::minmax(tri(t, k), tri(t, (k + 1) % 3))].push_back(t);
      auto opposite = [&](uint32_t t, uint32_t e0, uint32_t e1) {
        for (int k = 0; k < 3; ++k)
          if (tri(t, k) != e0 && tri(t, k) != e1) return tri(t, k);
        return e0;
      };
      for (const auto& [edge, tris] : edgeTris) {
        if (tris.size() != 2 || mesh.triangle_piece[tris[0]] != mesh.triangle_piece[tris[1]]) continue;
        const uint32_t piece = mesh.triangle_piece[tris[0]];
        HingeTerm hg;
        hg.v = {off + opposite(tris[0], edge.first, edge.second), off + opposite(tris[1], edge.first, edge.second),
                off + edge.first, off + edge.second};
        next.hinges.push_back(hg);
        HingeSide side;
        side.fabric = g.fabricOffset + piece;
        side.edgeDir[0] = rest(edge.second, 0) - rest(edge.first, 0);
        side.edgeDir[1] = rest(edge.second, 1) - rest(edge.first, 1);
        side.grain[0] = mesh.piece_grain[2 * piece];
        side.grain[1] = mesh.piece_grain[2 * piece + 1];
        next.hingeSides.push_back({side, side});
        const double dd = side.edgeDir[0] * side.edgeDir[0] + side.edgeDir[1] * side.edgeDir[1];
        next.hingeGeom.push_back(dd / (triArea(tris[0]) + triArea(tris[1])));
      }
      // Seam springs, each followed by a seam hinge across consecutive pairs (sim_model.cpp:147-175).
      auto triangleWithEdge = [&](uint32_t e0, uint32_t e1) -> int64_t {
        auto it = edgeTris.find(std::minmax(e0, e1));
        return (it == edgeTris.end() || it->second.size() != 1) ? -1 : it->second[0];
      };
      auto sp = [&](uint32_t k, int c) { return mesh.seam_pairs[4 * k + c]; };
      for (uint32_t k = 0; k < mesh.n_seam_pairs; ++k) {
        SeamSpring spring;
        spring.a = off + sp(k, 0);
        spring.b = off + sp(k, 1);
        spring.garment = gi;
        spring.startGap = static_cast<float>(norm(init(sp(k, 0)) - init(sp(k, 1))));
        next.seams.push_back(spring);
        if (k + 1 >= mesh.n_seam_pairs || sp(k + 1, 2) != sp(k, 2)) continue;
        const uint32_t nxa = sp(k + 1, 0), nxb = sp(k + 1, 1), spa = sp(k, 0), spb = sp(k, 1);
        const int64_t tA = triangleWithEdge(spa, nxa);
        const int64_t tB = triangleWithEdge(spb, nxb);
        if (tA < 0 || tB < 0) continue;
        const auto ta = static_cast<uint32_t>(tA), tb = static_cast<uint32_t>(tB);
        const uint32_t pieceA = mesh.triangle_piece[ta], pieceB = mesh.triangle_piece[tb];
        HingeTerm hg;
        hg.v = {off + opposite(ta, spa, nxa), off + opposite(tb, spb, nxb), off + spa, off + nxa};
        next.hinges.push_back(hg);
        HingeSide sa, sb;
        sa.fabric = g.fabricOffset + pieceA;
        sa.edgeDir[0] = rest(nxa, 0) - rest(spa, 0);
        sa.edgeDir[1] = rest(nxa, 1) - rest(spa, 1);
        sa.grain[0] = mesh.piece_grain[2 * pieceA];
        sa.grain[1] = mesh.piece_grain[2 * pieceA + 1];
        sb.fabric = g.fabricOffset + pieceB;
        sb.edgeDir[0] = rest(nxb, 0) - rest(spb, 0);
        sb.edgeDir[1] = rest(nxb, 1) - rest(spb, 1);
        sb.grain[0] = mesh.piece_grain[2 * pieceB];
        sb.grain[1] = mesh.piece_grain[2 * pieceB + 1];
        next.hingeSides.push_back({sa, sb});
        const double dd = sa.edgeDir[0] * sa.edgeDir[0] + sa.edgeDir[1] * sa.edgeDir[1];
        next.hingeGeom.push_back(dd / (triArea(ta) + triArea(tb)));
      }
      next.garments.push_back(g);
    }
    const size_t n = next.x.size();
    next.v.assign(n, V3f{});
    next.fExt.assign(n, V3{});
    next.pinned.assign(n, 0);
    next.mass.assign(n, 0.0);
    for (uint32_t k = 0; k < scene.n_pins; ++k) {
      const uint32_t pidx = scene.pin_particles[k];
      if (pidx >= n) throw std::out_of_range("pin particle out of range");
      const V3f pos = toF(V3{scene.pin_positions[3 * k], scene.pin_positions[3 * k + 1], scene.pin_positions[3 * k + 2]});
      next.pinned[pidx] = 1;
      next.pins.push_back({pidx, pos});
      next.x[pidx] = pos;
    }
    for (uint32_t k = 0; k < scene.n_forces; ++k) {
      const uint32_t pidx = scene.force_particles[k];
      if (pidx >= n) throw std::out_of_range("force particle out of range");
      next.fExt[pidx] += V3{scene.force_vectors[3 * k], scene.force_vectors[3 * k + 1], scene.force_vectors[3 * k + 2]};
    }
    for (uint32_t i = 0; i < n; ++i) next.updateMass(i);
    for (size_t h = 0; h < next.hinges.size(); ++h) next.hinges[h].K = next.hingeStiffness(h);
    next.buildStructure();
    return next;
  }

  // sim_model.cpp:199-208
  void updateFabric(const std::string& id, const std::vector<Fabric>& pieceFabric) {
    Garment& g = garments[garmentIndex(id)];
    if (pieceFabric.size() != g.pieceCount) throw std::invalid_argument("one fabric per pattern piece required");
    std::copy(pieceFabric.begin(), pieceFabric.end(), fabrics.begin() + g.fabricOffset);
    for (uint32_t i = g.offset; i < g.offset + g.count; ++i) updateMass(i);
    for (size_t h = 0; h < hinges.size(); ++h) {
      const uint32_t fi = hingeSides[h][0].fabric;
      if (fi >= g.fabricOffset && fi < g.fabricOffset + g.pieceCount) hinges[h].K = hingeStiffness(h);
    }
  }

  // sim_model.cpp:210-220
  void setPhase(const std::string& id, int phase) {
    const size_t gi = garmentIndex(id);
    Garment& g = garments[gi];
    g.phase = phase;
    if (phase == Assembling) {
      g.assemblyTime = 0;
      for (auto& sp : seams)
        if (sp.garment == gi) sp.startGap = normf(x[sp.a] - x[sp.b]);
    }
  }

  // sim_model.cpp:222-241
  drape_stats stats() const {
    drape_stats st{};
    double sumV2 = 0;
    size_t moving = 0;
    for (size_t i = 0; i < x.size(); ++i) {
      if (!allFinite(x[i]) || !allFinite(v[i])) st.has_non_finite = 1;
      if (pinned[i]) continue;
      const double v2 = sqnorm(toD(v[i]));
      st.kinetic_energy += 0.5 * mass[i] * v2;
      sumV2 += v2;
      ++moving;
    }
    st.rms_speed = moving ? std::sqrt(sumV2 / static_cast<double>(moving)) : 0.0;
    for (const auto& sp : seams) st.max_seam_gap = std::max(st.max_seam_gap, static_cast<double>(normf(x[sp.a] - x[sp.b])));
    if (hasBody)
      for (const auto& p : x) st.max_body_penetration = std::max(st.max_body_penetration, -body.sample(toD(p)));
    return st;
  }
};

// ---------------------------------------------------------------------------------------------------------------
// The solver (cpu_solver.cpp:34-355).

int gThreads = 1;

struct Solver {
  Model m;
  std::vector<V3> contactNormal;
  std::vector<double> contactOffset, frictionLambda;
  double substepH = 1.0 / 60.0;
  std::vector<uint8_t> contactNear, contactActive;
  std::vector<V3> xd, xStart, y, grad, step, r, z, p, ap, trial, apFull, resid;
  std::vector<M3> blocks, precond;
  drape_counters counters{};

  void allocate() {
    const size_t n = m.x.size();
    blocks.assign(m.colIndex.size(), M3{});
    precond.assign(n, identity());
    for (auto* vec : {&xd, &xStart, &y, &grad, &step, &r, &z, &p, &ap, &trial, &apFull, &resid}) vec->assign(n, V3{});
    contactNormal.assign(n, V3{});
    contactOffset.assign(n, 0.0);
    frictionLambda.assign(n, 0.0);
    contactNear.assign(n, 0);
    contactActive.assign(n, 0);
  }

  double contactGap(size_t i, V3 pos) const { return contactOffset[i] - dot(contactNormal[i], pos - xStart[i]); }
  void filter(std::vector<V3>& vec) const {
    for (size_t i = 0; i < vec.size(); ++i)
      if (contactActive[i]) vec[i] -= contactNormal[i] * dot(contactNormal[i], vec[i]);
  }

  // Evaluates terms of one kind in chunks: the evaluations run in parallel, the accumulation keeps term order.
  template <int N, class Verts, class Slots, class Eval>
  void addTerms(size_t count, const Verts& verts, const Slots& slots, const std::vector<V3>& pos, bool withDerivatives,
                double& energy, Eval&& eval) {
    constexpr size_t kChunk = 8192;
    std::vector<TermEval<N>> buf(std::min(count, kChunk));
    for (size_t base = 0; base < count; base += kChunk) {
      const size_t len = std::min(kChunk, count - base);
#pragma omp parallel for schedule(static) num_threads(gThreads) if (gThreads > 1)
      for (long long k = 0; k < static_cast<long long>(len); ++k) {
        const uint32_t* vs = verts(base + k);
        V3 xs[N];
        for (int a = 0; a < N; ++a) xs[a] = pos[vs[a]];
        TermEval<N>& te = buf[k];
        te = TermEval<N>{};
        eval(base + k, xs, te);
      }
      for (size_t k = 0; k < len; ++k) {
        const TermEval<N>& te = buf[k];
        energy += te.energy;
        if (!withDerivatives) continue;
        const uint32_t* vs = verts(base + k);
        const uint32_t* sl = slots(base + k);
        for (int a = 0; a < N; ++a) {
          grad[vs[a]] += te.grad[a];
          for (int b = 0; b < N; ++b) blocks[sl[a * N + b]] += te.hess[a][b];
        }
      }
    }
  }

  // cpu_solver.cpp:92-144
  double evaluate(const std::vector<V3>& pos, double invH2, bool withDerivatives) {
    const size_t n = pos.size();
    double energy = 0;
    if (withDerivatives) {
      std::fill(grad.begin(), grad.end(), V3{});
      std::fill(blocks.begin(), blocks.end(), M3{});
    }
    for (size_t i = 0; i < n; ++i) {
      if (m.pinned[i]) continue;
      const double k = m.mass[i] * invH2;
      const V3 d = pos[i] - y[i];
      energy += 0.5 * k * sqnorm(d);
      if (withDerivatives) {
        grad[i] += k * d;
        blocks[m.diagSlot[i]] += k * identity();
      }
    }
    const double limitScale = m.settings.strainLimitStiffnessScale;
    addTerms<3>(
        m.stretch.size(), [&](size_t k) { return m.stretch[k].v.data(); },
        [&](size_t k) { return m.stretchSlots[k].data(); }, pos, withDerivatives, energy,
        [&](size_t k, const V3* xs, TermEval<3>& te) {
          evalStretchTerm(m.stretch[k], xs, m.fabrics[m.stretch[k].fabric], limitScale, te);
        });
    addTerms<4>(
        m.hinges.size(), [&](size_t k) { return m.hinges[k].v.data(); }, [&](size_t k) { return m.hingeSlots[k].data(); },
        pos, withDerivatives, energy, [&](size_t k, const V3* xs, TermEval<4>& te) { evalHingeTerm(m.hinges[k], xs, te); });
    std::vector<std::array<uint32_t, 2>> seamAB(m.seams.size());
    for (size_t k = 0; k < m.seams.size(); ++k) seamAB[k] = {m.seams[k].a, m.seams[k].b};
    addTerms<2>(
        m.seams.size(), [&](size_t k) { return seamAB[k].data(); }, [&](size_t k) { return m.seamSlots[k].data(); }, pos,
        withDerivatives, energy, [&](size_t k, const V3* xs, TermEval<2>& te) {
          evalSeamTerm(xs, m.seamTarget(m.seams[k]), m.settings.seamStiffness, te);
        });
    for (size_t i = 0; i < n; ++i) {
      if (frictionLambda[i] <= 0) continue;
      TermEval<1> te;
      evalFrictionTerm(pos[i], xStart[i], contactNormal[i], m.fabrics[m.pFabric[i]].friction * frictionLambda[i],
                       m.settings.frictionSlipSpeed * substepH, te);
      energy += te.energy;
      if (withDerivatives) {
        grad[i] += te.grad[0];
        blocks[m.diagSlot[i]] += te.hess[0][0];
      }
    }
    if (withDerivatives)
      for (size_t i = 0; i < n; ++i)
        if (m.pinned[i]) grad[i] = V3{};
    return energy;
  }

  // cpu_solver.cpp:147-159 (rows are independent: parallel and bit-identical)
  void multiply(const std::vector<V3>& in, std::vector<V3>& out) const {
    const long long n = static_cast<long long>(in.size());
#pragma omp parallel for schedule(static) num_threads(gThreads) if (gThreads > 1)
    for (long long i = 0; i < n; ++i) {
      V3 acc;
      if (!m.pinned[i]) {
        for (uint32_t s = m.rowStart[i]; s < m.rowStart[i + 1]; ++s) {
          const uint32_t j = m.colIndex[s];
          if (!m.pinned[j]) acc += blocks[s] * in[j];
        }
      }
      out[i] = acc;
    }
  }

  static double vdot(const std::vector<V3>& a, const std::vector<V3>& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += dot(a[i], b[i]);
    return s;
  }

  // cpu_solver.cpp:169-213
  void solveLinear() {
    const size_t n = grad.size();
    for (size_t i = 0; i < n; ++i) {
      if (m.pinned[i]) {
        precond[i] = M3{};
        continue;
      }
      M3 inv;
      precond[i] = invertSpd3(blocks[m.diagSlot[i]], inv) ? inv : identity();
    }
    for (size_t i = 0; i < n; ++i) {
      step[i] = V3{};
      resid[i] = -grad[i];
      r[i] = resid[i];
    }
    filter(r);
    for (size_t i = 0; i < n; ++i) z[i] = precond[i] * r[i];
    filter(z);
    p = z;
    double rz = vdot(r, z);
    const double r0 = std::sqrt(vdot(r, r));
    for (int it = 0; r0 > 0 && it < m.settings.maxCgIterations; ++it) {
      multiply(p, apFull);
      ap = apFull;
      filter(ap);
      const double pAp = vdot(p, ap);
      if (!(pAp > 0)) break;
      ++counters.pcg_iterations;
      const double alpha = rz / pAp;
      for (size_t i = 0; i < n; ++i) {
        step[i] += alpha * p[i];
        r[i] -= alpha * ap[i];
        resid[i] -= alpha * apFull[i];
      }
      if (std::sqrt(vdot(r, r)) <= m.settings.cgTolerance * r0) break;
      for (size_t i = 0; i < n; ++i) z[i] = precond[i] * r[i];
      filter(z);
      const double rzNew = vdot(r, z);
      const double beta = rzNew / rz;
      rz = rzNew;
      for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    for (size_t i = 0; i < n; ++i)
      m.contactForce[i] = contactActive[i] ? std::max(0.0, -dot(resid[i], contactNormal[i])) : 0.0;
  }

  // cpu_solver.cpp:217-228
  double evaluateWithContact(double invH2) {
    for (size_t i = 0; i < xd.size(); ++i) {
      if (!contactNear[i]) continue;
      const double c = contactGap(i, xd[i]);
      if (c > -kContactSnap) contactActive[i] = 1;
      if (contactActive[i] && c > 0) xd[i] += c * contactNormal[i];
    }
    const double e = evaluate(xd, invH2, true);
    for (size_t i = 0; i < xd.size(); ++i)
      if (contactActive[i] && dot(grad[i], contactNormal[i]) < 0) contactActive[i] = 0;
    return e;
  }

  // Substep start: xStart, y (inertia target) and the warm start xd = xStart + h v (cpu_solver.cpp:234-245).
  void predict(double h) {
    const V3 gravity = m.settings.gravity;
    const size_t n = m.x.size();
    for (size_t i = 0; i < n; ++i) {
      xStart[i] = toD(m.x[i]);
      const V3 vel = toD(m.v[i]);
      if (m.pinned[i]) {
        y[i] = xStart[i];
        xd[i] = xStart[i];
        continue;
      }
      const V3 g = m.garments[m.pGarment[i]].phase == Settling ? gravity : V3{};
      y[i] = xStart[i] + h * vel + (h * h) * (g + m.fExt[i] / m.mass[i]);
      xd[i] = xStart[i] + h * vel;
    }
    substepH = h;
  }

  // cpu_solver.cpp:230-314
  void substep(double h, int newtonIterations) {
    const double invH2 = 1.0 / (h * h);
    const size_t n = m.x.size();
    predict(h);
    for (size_t i = 0; i < n; ++i) {
      contactNear[i] = 0;
      contactActive[i] = 0;
      frictionLambda[i] = 0;
      const double lastForce = m.contactForce[i];
      m.contactForce[i] = 0;
      if (!m.hasBody || m.pinned[i]) continue;
      const double radius = m.contactRadius(i);
      const double phi = m.body.sample(xStart[i]);
      if (phi >= radius + kContactReach) continue;
      const V3 normal = m.body.gradient(xStart[i]);
      if (isZero(normal)) continue;
      contactNear[i] = 1;
      contactNormal[i] = normal;
      contactOffset[i] = radius - phi;
      frictionLambda[i] = lastForce;
    }
    for (int it = 0; it < newtonIterations; ++it) {
      ++counters.newton_iterations;
      const double e0 = evaluateWithContact(invH2);
      solveLinear();
      const double slope = vdot(grad, step);
      if (!(slope < 0)) break;
      double alpha = 1.0;
      for (int ls = 0; ls < 20; ++ls) {
        ++counters.line_search_trials;
        for (size_t i = 0; i < n; ++i) trial[i] = xd[i] + alpha * step[i];
        if (evaluate(trial, invH2, false) <= e0 + 1e-4 * alpha * slope) break;
        alpha *= 0.5;
      }
      double maxMove = 0;
      for (size_t i = 0; i < n; ++i) {
        xd[i] += alpha * step[i];
        const V3 mv = alpha * step[i];
        maxMove = std::max(maxMove, std::max(std::abs(mv.x), std::max(std::abs(mv.y), std::abs(mv.z))));
      }
      for (size_t i = 0; i < n; ++i) frictionLambda[i] = contactActive[i] ? m.contactForce[i] : 0.0;
      if (maxMove < 1e-9) break;
    }
    for (size_t i = 0; i < n; ++i) m.x[i] = toF(xd[i]);
    if (m.hasBody) {
      for (size_t i = 0; i < n; ++i) {
        if (m.pinned[i]) continue;
        const double radius = m.contactRadius(i);
        if (m.body.sample(toD(m.x[i])) >= radius - kContactSnap) continue;
        if (isZero(m.body.gradient(toD(m.x[i])))) {
          m.x[i] = landOnBody(m.body, static_cast<float>(radius), toF(xStart[i]), m.x[i]);
          continue;
        }
        collideBody(m.body, static_cast<float>(radius), static_cast<float>(m.fabrics[m.pFabric[i]].friction),
                    toF(xStart[i]), m.x[i]);
      }
    }
    for (const auto& [i, pos] : m.pins) m.x[i] = pos;
    for (size_t i = 0; i < n; ++i) {
      m.v[i] = toF((toD(m.x[i]) - xStart[i]) / h);
      double c = m.settings.dampingScale * m.fabrics[m.pFabric[i]].damping;
      if (m.garments[m.pGarment[i]].phase == Assembling) c += m.settings.assemblyExtraDamping;
      m.v[i] = m.v[i] * static_cast<float>(std::max(0.0, 1.0 - c * h));
    }
    for (auto& g : m.garments)
      if (g.phase == Assembling) g.assemblyTime += h;
    ++counters.substeps;
  }

  // cpu_solver.cpp:333-338
  void stepFrame(double dt, int substeps, int iterations) {
    substeps = std::max(1, substeps);
    const double h = dt / substeps;
    iterations = std::max(1, iterations);
    for (int s = 0; s < substeps; ++s) substep(h, iterations);
  }
};

}  // namespace orc

// =================================================================================================================
// C entry points.

struct oracle_solver {
  orc::Solver s;
  bool built = false;
};

namespace {

thread_local std::string gError;

template <class F>
int guard(F&& f) {
  try {
    gError.clear();
    f();
    return DRAPE_OK;
  } catch (const std::invalid_argument& e) {
    gError = e.what();
    return DRAPE_ERR_INVALID_ARGUMENT;
  } catch (const std::out_of_range& e) {
    gError = e.what();
    return DRAPE_ERR_OUT_OF_RANGE;
  } catch (const std::logic_error& e) {
    gError = e.what();
    return DRAPE_ERR_LOGIC;
  } catch (const std::exception& e) {
    gError = e.what();
    return DRAPE_ERR_RUNTIME;
  }
}

void requireBuilt(const oracle_solver* s) {
  if (!s) throw std::invalid_argument("null solver");
  if (!s->built) throw std::logic_error("solver not built");
}

using orc::M3;
using orc::V3;

void writeBlock(const M3& b, double* out) { std::memcpy(out, b.a, sizeof(double) * 9); }

template <int N>
void writeTerm(const orc::TermEval<N>& te, double* energy, double* grad, double* hess) {
  if (energy) *energy = te.energy;
  if (grad)
    for (int a = 0; a < N; ++a) {
      grad[3 * a] = te.grad[a].x;
      grad[3 * a + 1] = te.grad[a].y;
      grad[3 * a + 2] = te.grad[a].z;
    }
  if (hess)
    for (int a = 0; a < N; ++a)
      for (int b = 0; b < N; ++b) writeBlock(te.hess[a][b], hess + 9 * (a * N + b));
}

template <int N>
void readPoints(const double* x, V3* out) {
  for (int a = 0; a < N; ++a) out[a] = {x[3 * a], x[3 * a + 1], x[3 * a + 2]};
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// bakeCollisionField (collision_bake.cpp:98-208), restated sequentially in FP64.

namespace bk {

struct P3 {
  double x, y, z;
};
inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 scale(double s, P3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dotp(P3 a, P3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline P3 crossp(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline P3 normalizedp(P3 a) {  // Eigen's normalized(): unchanged when the norm is zero
  const double z = dotp(a, a);
  if (!(z > 0.0)) return a;
  const double nn = std::sqrt(z);
  return {a.x / nn, a.y / nn, a.z / nn};
}

enum Feature { kFace = 0, kVa, kVb, kVc, kEab, kEbc, kEca };

// Ericson RTCD 5.1.5 (collision_bake.cpp:19-55).
P3 closestPointOnTriangle(P3 p, P3 a, P3 b, P3 c, int& f) {
  const P3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const double d1 = dotp(ab, ap), d2 = dotp(ac, ap);
  if (d1 <= 0 && d2 <= 0) {
    f = kVa;
    return a;
  }
  const P3 bp = sub(p, b);
  const double d3 = dotp(ab, bp), d4 = dotp(ac, bp);
  if (d3 >= 0 && d4 <= d3) {
    f = kVb;
    return b;
  }
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    f = kEab;
    return add(a, scale(d1 / (d1 - d3), ab));
  }
  const P3 cp = sub(p, c);
  const double d5 = dotp(ab, cp), d6 = dotp(ac, cp);
  if (d6 >= 0 && d5 <= d6) {
    f = kVc;
    return c;
  }
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    f = kEca;
    return add(a, scale(d2 / (d2 - d6), ac));
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    f = kEbc;
    return add(b, scale((d4 - d3) / ((d4 - d3) + (d5 - d6)), sub(c, b)));
  }
  const double denom = 1.0 / (va + vb + vc);
  f = kFace;
  return add(add(a, scale(vb * denom, ab)), scale(vc * denom, ac));
}

struct Mesh {
  std::vector<P3> v;
  std::vector<std::array<uint32_t, 3>> t;
};

bool closedManifold(const Mesh& m) {
  std::map<std::pair<uint32_t, uint32_t>, int> directed;
  for (const auto& tri : m.t)
    for (int k = 0; k < 3; ++k)
      if (++directed[{tri[k], tri[(k + 1) % 3]}] > 1) return false;
  for (const auto& [e, c] : directed)
    if (!directed.count({e.second, e.first})) return false;
  return !m.t.empty();
}

// Pseudonormals (collision_bake.cpp:57-95).
struct Normals {
  std::vector<P3> face, vertex;
  std::map<std::pair<uint32_t, uint32_t>, P3> edge;
  explicit Normals(const Mesh& m) : face(m.t.size()), vertex(m.v.size(), P3{0, 0, 0}) {
    for (size_t t = 0; t < m.t.size(); ++t) {
      const auto& tri = m.t[t];
      face[t] = normalizedp(crossp(sub(m.v[tri[1]], m.v[tri[0]]), sub(m.v[tri[2]], m.v[tri[0]])));
      for (int k = 0; k < 3; ++k) {
        const P3 p = m.v[tri[k]];
        const P3 e1 = normalizedp(sub(m.v[tri[(k + 1) % 3]], p));
        const P3 e2 = normalizedp(sub(m.v[tri[(k + 2) % 3]], p));
        vertex[tri[k]] = add(vertex[tri[k]], scale(std::acos(std::clamp(dotp(e1, e2), -1.0, 1.0)), face[t]));
        auto key = std::minmax(tri[k], tri[(k + 1) % 3]);
        auto it = edge.try_emplace(key, P3{0, 0, 0}).first;
        it->second = add(it->second, face[t]);
      }
    }
  }
  P3 of(const Mesh& m, size_t t, int f) const {
    const auto& tri = m.t[t];
    switch (f) {
      case kVa: return vertex[tri[0]];
      case kVb: return vertex[tri[1]];
      case kVc: return vertex[tri[2]];
      case kEab: return edge.at(std::minmax(tri[0], tri[1]));
      case kEbc: return edge.at(std::minmax(tri[1], tri[2]));
      case kEca: return edge.at(std::minmax(tri[2], tri[0]));
      default: return face[t];
    }
  }
};

std::vector<Mesh> meshes(const drape_body_mesh* c, uint32_t n) {
  if (!c || n == 0) throw std::invalid_argument("bake: no components");
  std::vector<Mesh> out(n);
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t k = 0; k < c[i].n_vertices; ++k)
      out[i].v.push_back({c[i].vertices[3 * k], c[i].vertices[3 * k + 1], c[i].vertices[3 * k + 2]});
    for (uint32_t k = 0; k < c[i].n_triangles; ++k) {
      std::array<uint32_t, 3> tri{c[i].triangles[3 * k], c[i].triangles[3 * k + 1], c[i].triangles[3 * k + 2]};
      for (uint32_t x : tri)
        if (x >= c[i].n_vertices) throw std::out_of_range("bake: triangle vertex out of range");
      out[i].t.push_back(tri);
    }
    if (!closedManifold(out[i]))
      throw std::invalid_argument("collision component " + std::to_string(i) + " is not a closed manifold");
  }
  return out;
}

struct Grid {
  P3 origin;
  int dims[3];
  double cell, band, padding;
};

Grid gridOf(const std::vector<Mesh>& ms, const drape_bake_options* o) {
  Grid g{};
  g.cell = o ? o->cell_size : 0.004;
  g.padding = o ? o->padding : 0.05;
  g.band = o ? o->band : 0.02;
  const double inf = std::numeric_limits<double>::infinity();
  double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
  for (const auto& m : ms)
    for (const auto& p : m.v) {
      const double c[3] = {p.x, p.y, p.z};
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], c[a]);
        hi[a] = std::max(hi[a], c[a]);
      }
    }
  double org[3];
  for (int a = 0; a < 3; ++a) {
    org[a] = lo[a] - g.padding;
    g.dims[a] = static_cast<int>(std::ceil(((hi[a] - lo[a]) + 2 * g.padding) / g.cell)) + 1;
  }
  g.origin = {org[0], org[1], org[2]};
  return g;
}

void bake(const std::vector<Mesh>& ms, const Grid& g, float* values) {
  const size_t nx = g.dims[0], ny = g.dims[1], nz = g.dims[2], count = nx * ny * nz;
  for (size_t i = 0; i < count; ++i) values[i] = static_cast<float>(g.band);
  auto index = [&](size_t x, size_t y, size_t z) { return (z * ny + y) * nx + x; };
  auto position = [&](size_t x, size_t y, size_t z) {
    return P3{g.origin.x + g.cell * double(x), g.origin.y + g.cell * double(y), g.origin.z + g.cell * double(z)};
  };
  const double o[3] = {g.origin.x, g.origin.y, g.origin.z};
  auto cellRange = [&](double minV, double maxV, int axis, size_t& i0, size_t& i1) {
    i0 = static_cast<size_t>(std::max(0.0, std::ceil((minV - o[axis]) / g.cell)));
    i1 = static_cast<size_t>(std::min<double>(g.dims[axis] - 1, std::floor((maxV - o[axis]) / g.cell)));
  };
  const double inf = std::numeric_limits<double>::infinity();
  std::vector<double> dist2(count);
  std::vector<int8_t> sign(count);
  std::vector<uint8_t> outside(count);
  for (const auto& m : ms) {
    std::fill(dist2.begin(), dist2.end(), inf);
    std::fill(sign.begin(), sign.end(), 0);
    const Normals nrm(m);
    for (size_t t = 0; t < m.t.size(); ++t) {
      const P3 a = m.v[m.t[t][0]], b = m.v[m.t[t][1]], c = m.v[m.t[t][2]];
      const P3 tmin{std::min(std::min(a.x, b.x), c.x), std::min(std::min(a.y, b.y), c.y), std::min(std::min(a.z, b.z), c.z)};
      const P3 tmax{std::max(std::max(a.x, b.x), c.x), std::max(std::max(a.y, b.y), c.y), std::max(std::max(a.z, b.z), c.z)};
      const double mn[3] = {tmin.x, tmin.y, tmin.z}, mx[3] = {tmax.x, tmax.y, tmax.z};
      size_t r[3][2];
      for (int axis = 0; axis < 3; ++axis) cellRange(mn[axis] - g.band, mx[axis] + g.band, axis, r[axis][0], r[axis][1]);
      for (size_t z = r[2][0]; z <= r[2][1]; ++z)
        for (size_t y = r[1][0]; y <= r[1][1]; ++y)
          for (size_t x = r[0][0]; x <= r[0][1]; ++x) {
            const size_t i = index(x, y, z);
            const P3 p = position(x, y, z);
            const P3 gap{std::max(std::max(tmin.x - p.x, p.x - tmax.x), 0.0), std::max(std::max(tmin.y - p.y, p.y - tmax.y), 0.0),
                         std::max(std::max(tmin.z - p.z, p.z - tmax.z), 0.0)};
            if (dotp(gap, gap) >= dist2[i]) continue;
            int f;
            const P3 q = closestPointOnTriangle(p, a, b, c, f);
            const P3 pq = sub(p, q);
            const double d2 = dotp(pq, pq);
            if (d2 < dist2[i]) {
              dist2[i] = d2;
              sign[i] = dotp(pq, nrm.of(m, t, f)) >= 0 ? 1 : -1;
            }
          }
    }
    std::fill(outside.begin(), outside.end(), 0);
    std::deque<size_t> queue;
    auto seed = [&](size_t x, size_t y, size_t z) {
      const size_t i = index(x, y, z);
      if (!outside[i] && dist2[i] > g.band * g.band) {
        outside[i] = 1;
        queue.push_back(i);
      }
    };
    for (size_t z = 0; z < nz; ++z)
      for (size_t y = 0; y < ny; ++y)
        for (size_t x = 0; x < nx; ++x)
          if (x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 || z == nz - 1) seed(x, y, z);
    while (!queue.empty()) {
      const size_t i = queue.front();
      queue.pop_front();
      const size_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
      if (x > 0) seed(x - 1, y, z);
      if (x + 1 < nx) seed(x + 1, y, z);
      if (y > 0) seed(x, y - 1, z);
      if (y + 1 < ny) seed(x, y + 1, z);
      if (z > 0) seed(x, y, z - 1);
      if (z + 1 < nz) seed(x, y, z + 1);
    }
    for (size_t i = 0; i < count; ++i) {
      double v;
      if (dist2[i] <= g.band * g.band)
        v = sign[i] * std::sqrt(dist2[i]);
      else
        v = outside[i] ? g.band : -g.band;
      values[i] = std::min(values[i], static_cast<float>(std::clamp(v, -g.band, g.band)));
    }
  }
}

}  // namespace bk

extern "C" {

const char* oracle_last_error(void) { return gError.c_str(); }

int oracle_set_threads(int32_t n) {
  return guard([&] {
    if (n < 1) throw std::invalid_argument("threads must be >= 1");
    orc::gThreads = n;
  });
}

int oracle_solver_create(oracle_solver** out) {
  return guard([&] {
    if (!out) throw std::invalid_argument("null output");
    *out = new oracle_solver();
  });
}

void oracle_solver_destroy(oracle_solver* s) { delete s; }

int oracle_solver_build(oracle_solver* s, const drape_scene_desc* scene) {
  return guard([&] {
    if (!s || !scene) throw std::invalid_argument("null argument");
    orc::Solver next;
    next.m = orc::Model::build(*scene);
    next.allocate();
    s->s = std::move(next);
    s->built = true;
  });
}

int oracle_solver_update_fabric(oracle_solver* s, const char* garment, uint32_t n, const drape_fabric* fabrics) {
  return guard([&] {
    requireBuilt(s);
    std::vector<orc::Fabric> f;
    for (uint32_t k = 0; k < n; ++k) f.push_back(orc::fromDesc(fabrics[k]));
    s->s.m.updateFabric(garment ? garment : "", f);
  });
}

int oracle_solver_set_phase(oracle_solver* s, const char* garment, int32_t phase) {
  return guard([&] {
    requireBuilt(s);
    s->s.m.setPhase(garment ? garment : "", phase);
  });
}

int oracle_solver_step(oracle_solver* s, double dt, int32_t substeps, int32_t iterations) {
  return guard([&] {
    requireBuilt(s);
    if (s->s.m.x.empty()) return;
    s->s.stepFrame(dt, substeps, iterations);
  });
}

int oracle_solver_particle_count(const oracle_solver* s, size_t* out) {
  return guard([&] {
    requireBuilt(s);
    *out = s->s.m.x.size();
  });
}

int oracle_solver_range(const oracle_solver* s, const char* garment, uint32_t* offset, uint32_t* count) {
  return guard([&] {
    requireBuilt(s);
    const auto& g = s->s.m.garments[s->s.m.garmentIndex(garment ? garment : "")];
    *offset = g.offset;
    *count = g.count;
  });
}

int oracle_solver_read_positions(const oracle_solver* s, float* out, size_t n) {
  return guard([&] {
    requireBuilt(s);
    if (n != s->s.m.x.size()) throw std::invalid_argument("readPositions: span size must equal particleCount()");
    for (size_t i = 0; i < n; ++i) {
      out[3 * i] = s->s.m.x[i].x;
      out[3 * i + 1] = s->s.m.x[i].y;
      out[3 * i + 2] = s->s.m.x[i].z;
    }
  });
}

int oracle_solver_read_velocities(const oracle_solver* s, float* out, size_t n) {
  return guard([&] {
    requireBuilt(s);
    if (n != s->s.m.v.size()) throw std::invalid_argument("span size must equal particleCount()");
    for (size_t i = 0; i < n; ++i) {
      out[3 * i] = s->s.m.v[i].x;
      out[3 * i + 1] = s->s.m.v[i].y;
      out[3 * i + 2] = s->s.m.v[i].z;
    }
  });
}

int oracle_solver_write_state(oracle_solver* s, const float* x, const float* v, size_t n) {
  return guard([&] {
    requireBuilt(s);
    if (n != s->s.m.x.size()) throw std::invalid_argument("span size must equal particleCount()");
    for (size_t i = 0; i < n; ++i) {
      if (x) s->s.m.x[i] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
      if (v) s->s.m.v[i] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    }
  });
}

int oracle_solver_stats(const oracle_solver* s, drape_stats* out) {
  return guard([&] {
    requireBuilt(s);
    *out = s->s.m.stats();
  });
}

// Snapshot blob: u64 n, u64 garments, u64 seams; positions, velocities (3 floats each per particle); phases
// (int32 per garment); assemblyTime (double per garment); seamStartGap (float per seam); contactForce (double per
// particle) — the fields of Snapshot (solver.hpp:76-82).
static size_t snapshotBytes(const orc::Model& m) {
  return 3 * sizeof(uint64_t) + m.x.size() * 24 + m.garments.size() * (4 + 8) + m.seams.size() * 4 + m.x.size() * 8;
}

int oracle_solver_snapshot_size(const oracle_solver* s, size_t* bytes) {
  return guard([&] {
    requireBuilt(s);
    *bytes = snapshotBytes(s->s.m);
  });
}

int oracle_solver_snapshot_save(const oracle_solver* s, void* out, size_t bytes) {
  return guard([&] {
    requireBuilt(s);
    const auto& m = s->s.m;
    if (bytes != snapshotBytes(m)) throw std::invalid_argument("snapshot buffer size mismatch");
    char* p = static_cast<char*>(out);
    const uint64_t head[3] = {m.x.size(), m.garments.size(), m.seams.size()};
    std::memcpy(p, head, sizeof(head));
    p += sizeof(head);
    std::memcpy(p, m.x.data(), m.x.size() * 12);
    p += m.x.size() * 12;
    std::memcpy(p, m.v.data(), m.v.size() * 12);
    p += m.v.size() * 12;
    for (const auto& g : m.garments) {
      const int32_t ph = g.phase;
      std::memcpy(p, &ph, 4);
      p += 4;
    }
    for (const auto& g : m.garments) {
      std::memcpy(p, &g.assemblyTime, 8);
      p += 8;
    }
    for (const auto& sp : m.seams) {
      std::memcpy(p, &sp.startGap, 4);
      p += 4;
    }
    std::memcpy(p, m.contactForce.data(), m.contactForce.size() * 8);
  });
}

int oracle_solver_snapshot_restore(oracle_solver* s, const void* in, size_t bytes) {
  return guard([&] {
    requireBuilt(s);
    auto& m = s->s.m;
    const char* p = static_cast<const char*>(in);
    if (bytes < 3 * sizeof(uint64_t)) throw std::invalid_argument("snapshot does not match this scene");
    uint64_t head[3];
    std::memcpy(head, p, sizeof(head));
    p += sizeof(head);
    const size_t expect = 3 * sizeof(uint64_t) + head[0] * 24 + head[1] * 12 + head[2] * 4 + head[0] * 8;
    if (head[0] != m.x.size() || head[1] != m.garments.size() || bytes != expect)
      throw std::invalid_argument("snapshot does not match this scene");
    std::memcpy(m.x.data(), p, m.x.size() * 12);
    p += m.x.size() * 12;
    std::memcpy(m.v.data(), p, m.v.size() * 12);
    p += m.v.size() * 12;
    for (auto& g : m.garments) {
      int32_t ph;
      std::memcpy(&ph, p, 4);
      g.phase = ph;
      p += 4;
    }
    for (auto& g : m.garments) {
      std::memcpy(&g.assemblyTime, p, 8);
      p += 8;
    }
    if (head[2] == m.seams.size()) {
      for (auto& sp : m.seams) {
        std::memcpy(&sp.startGap, p, 4);
        p += 4;
      }
    } else {
      p += head[2] * 4;
    }
    std::memcpy(m.contactForce.data(), p, m.contactForce.size() * 8);
  });
}

int oracle_solver_counters(const oracle_solver* s, drape_counters* out) {
  return guard([&] {
    requireBuilt(s);
    *out = s->s.counters;
  });
}

int oracle_solver_structure_sizes(const oracle_solver* s, drape_structure_sizes* out) {
  return guard([&] {
    requireBuilt(s);
    const auto& m = s->s.m;
    out->n_particles = static_cast<uint32_t>(m.x.size());
    out->n_stretch = static_cast<uint32_t>(m.stretch.size());
    out->n_hinges = static_cast<uint32_t>(m.hinges.size());
    out->n_seams = static_cast<uint32_t>(m.seams.size());
    out->n_blocks = static_cast<uint32_t>(m.colIndex.size());
  });
}

int oracle_solver_read_structure(const oracle_solver* s, uint32_t* row_start, uint32_t* col_index, uint32_t* diag_slot,
                                 uint32_t* stretch_slots, uint32_t* hinge_slots, uint32_t* seam_slots) {
  return guard([&] {
    requireBuilt(s);
    const auto& m = s->s.m;
    if (row_start) std::copy(m.rowStart.begin(), m.rowStart.end(), row_start);
    if (col_index) std::copy(m.colIndex.begin(), m.colIndex.end(), col_index);
    if (diag_slot) std::copy(m.diagSlot.begin(), m.diagSlot.end(), diag_slot);
    if (stretch_slots)
      for (size_t k = 0; k < m.stretchSlots.size(); ++k) std::copy(m.stretchSlots[k].begin(), m.stretchSlots[k].end(), stretch_slots + 9 * k);
    if (hinge_slots)
      for (size_t k = 0; k < m.hingeSlots.size(); ++k) std::copy(m.hingeSlots[k].begin(), m.hingeSlots[k].end(), hinge_slots + 16 * k);
    if (seam_slots)
      for (size_t k = 0; k < m.seamSlots.size(); ++k) std::copy(m.seamSlots[k].begin(), m.seamSlots[k].end(), seam_slots + 4 * k);
  });
}

int oracle_solver_read_terms(const oracle_solver* s, uint32_t* stretch_v, uint32_t* hinge_v, double* hinge_k,
                             uint32_t* seam_ab) {
  return guard([&] {
    requireBuilt(s);
    const auto& m = s->s.m;
    for (size_t k = 0; k < m.stretch.size(); ++k)
      if (stretch_v) std::copy(m.stretch[k].v.begin(), m.stretch[k].v.end(), stretch_v + 3 * k);
    for (size_t k = 0; k < m.hinges.size(); ++k) {
      if (hinge_v) std::copy(m.hinges[k].v.begin(), m.hinges[k].v.end(), hinge_v + 4 * k);
      if (hinge_k) hinge_k[k] = m.hinges[k].K;
    }
    for (size_t k = 0; k < m.seams.size(); ++k)
      if (seam_ab) {
        seam_ab[2 * k] = m.seams[k].a;
        seam_ab[2 * k + 1] = m.seams[k].b;
      }
  });
}

// gpu_solver_core.cpp:184-224: per particle [own i, stretch n+3t+a, hinge n+3S+4h+a, seam n+3S+4H+2k+a]; per BSR
// slot [diag i, stretch n+9t+3a+c, hinge n+9S+16h+4a+c, seam n+9S+16H+4k+2a+c], each in term order.
static void gatherLists(const orc::Model& m, std::vector<uint32_t>& gs, std::vector<uint32_t>& gi,
                        std::vector<uint32_t>& bs, std::vector<uint32_t>& bi) {
  const uint32_t n = static_cast<uint32_t>(m.x.size()), S = static_cast<uint32_t>(m.stretch.size());
  const uint32_t H = static_cast<uint32_t>(m.hinges.size()), K = static_cast<uint32_t>(m.seams.size());
  std::vector<std::vector<uint32_t>> g(n), b(m.colIndex.size());
  for (uint32_t i = 0; i < n; ++i) {
    g[i].push_back(i);
    b[m.diagSlot[i]].push_back(i);
  }
  for (uint32_t t = 0; t < S; ++t)
    for (uint32_t a = 0; a < 3; ++a) {
      g[m.stretch[t].v[a]].push_back(n + 3 * t + a);
      for (uint32_t c = 0; c < 3; ++c) b[m.stretchSlots[t][3 * a + c]].push_back(n + 9 * t + 3 * a + c);
    }
  for (uint32_t h = 0; h < H; ++h)
    for (uint32_t a = 0; a < 4; ++a) {
      g[m.hinges[h].v[a]].push_back(n + 3 * S + 4 * h + a);
      for (uint32_t c = 0; c < 4; ++c) b[m.hingeSlots[h][4 * a + c]].push_back(n + 9 * S + 16 * h + 4 * a + c);
    }
  for (uint32_t k = 0; k < K; ++k) {
    const uint32_t vs[2] = {m.seams[k].a, m.seams[k].b};
    for (uint32_t a = 0; a < 2; ++a) {
      g[vs[a]].push_back(n + 3 * S + 4 * H + 2 * k + a);
      for (uint32_t c = 0; c < 2; ++c) b[m.seamSlots[k][2 * a + c]].push_back(n + 9 * S + 16 * H + 4 * k + 2 * a + c);
    }
  }
  auto flatten = [](const std::vector<std::vector<uint32_t>>& lists, std::vector<uint32_t>& start,
                    std::vector<uint32_t>& idx) {
    start.assign(1, 0);
    for (const auto& l : lists) {
      idx.insert(idx.end(), l.begin(), l.end());
      start.push_back(static_cast<uint32_t>(idx.size()));
    }
  };
  flatten(g, gs, gi);
  flatten(b, bs, bi);
}

int oracle_solver_gather_sizes(const oracle_solver* s, uint32_t* n_grad, uint32_t* n_hess) {
  return guard([&] {
    requireBuilt(s);
    std::vector<uint32_t> gs, gi, bs, bi;
    gatherLists(s->s.m, gs, gi, bs, bi);
    *n_grad = static_cast<uint32_t>(gi.size());
    *n_hess = static_cast<uint32_t>(bi.size());
  });
}

int oracle_solver_read_gather(const oracle_solver* s, uint32_t* grad_start, uint32_t* grad_idx, uint32_t* block_start,
                              uint32_t* block_idx) {
  return guard([&] {
    requireBuilt(s);
    std::vector<uint32_t> gs, gi, bs, bi;
    gatherLists(s->s.m, gs, gi, bs, bi);
    if (grad_start) std::copy(gs.begin(), gs.end(), grad_start);
    if (grad_idx) std::copy(gi.begin(), gi.end(), grad_idx);
    if (block_start) std::copy(bs.begin(), bs.end(), block_start);
    if (block_idx) std::copy(bi.begin(), bi.end(), block_idx);
  });
}

int oracle_solver_debug_assemble(oracle_solver* s, double h, const double* u, double* energy, double* grad,
                                 double* blocks) {
  return guard([&] {
    requireBuilt(s);
    auto& so = s->s;
    const size_t n = so.m.x.size();
    so.predict(h);
    for (size_t i = 0; i < n; ++i) {
      so.contactNear[i] = so.contactActive[i] = 0;
      so.frictionLambda[i] = 0;
      so.xd[i] = so.xStart[i] + V3{u[3 * i], u[3 * i + 1], u[3 * i + 2]};
    }
    const double e = so.evaluate(so.xd, 1.0 / (h * h), true);
    if (energy) *energy = e;
    if (grad)
      for (size_t i = 0; i < n; ++i) {
        grad[3 * i] = so.grad[i].x;
        grad[3 * i + 1] = so.grad[i].y;
        grad[3 * i + 2] = so.grad[i].z;
      }
    if (blocks)
      for (size_t k = 0; k < so.blocks.size(); ++k) writeBlock(so.blocks[k], blocks + 9 * k);
  });
}

int oracle_solver_debug_spmv(oracle_solver* s, const double* in, double* out) {
  return guard([&] {
    requireBuilt(s);
    const size_t n = s->s.m.x.size();
    std::vector<V3> a(n), b(n);
    for (size_t i = 0; i < n; ++i) a[i] = {in[3 * i], in[3 * i + 1], in[3 * i + 2]};
    s->s.multiply(a, b);
    for (size_t i = 0; i < n; ++i) {
      out[3 * i] = b[i].x;
      out[3 * i + 1] = b[i].y;
      out[3 * i + 2] = b[i].z;
    }
  });
}

int oracle_eval_stretch(const double* x9, const double* dm_inv4, double area, const drape_fabric* f,
                        double limit_scale, double* energy, double* grad9, double* hess81) {
  return guard([&] {
    orc::StretchTerm t;
    for (int k = 0; k < 4; ++k) t.dmInv.d[k] = dm_inv4[k];
    t.area = area;
    V3 xs[3];
    readPoints<3>(x9, xs);
    orc::TermEval<3> te;
    orc::evalStretchTerm(t, xs, orc::fromDesc(*f), limit_scale, te);
    writeTerm(te, energy, grad9, hess81);
  });
}

int oracle_eval_hinge(const double* x12, double K, double rest_angle, double* energy, double* grad12,
                      double* hess144) {
  return guard([&] {
    orc::HingeTerm t;
    t.K = K;
    t.restAngle = rest_angle;
    V3 xs[4];
    readPoints<4>(x12, xs);
    orc::TermEval<4> te;
    orc::evalHingeTerm(t, xs, te);
    writeTerm(te, energy, grad12, hess144);
  });
}

int oracle_eval_seam(const double* x6, double target, double K, double* energy, double* grad6, double* hess36) {
  return guard([&] {
    V3 xs[2];
    readPoints<2>(x6, xs);
    orc::TermEval<2> te;
    orc::evalSeamTerm(xs, target, K, te);
    writeTerm(te, energy, grad6, hess36);
  });
}

int oracle_eval_friction(const double* x3, const double* x0, const double* n3, double mu_lambda, double eps,
                         double* energy, double* grad3, double* hess9) {
  return guard([&] {
    orc::TermEval<1> te;
    orc::evalFrictionTerm({x3[0], x3[1], x3[2]}, {x0[0], x0[1], x0[2]}, {n3[0], n3[1], n3[2]}, mu_lambda, eps, te);
    writeTerm(te, energy, grad3, hess9);
  });
}

int oracle_strain(const double* x9, const double* dm_inv4, int32_t which, double* C, double* grad9) {
  return guard([&] {
    orc::M2 d;
    for (int k = 0; k < 4; ++k) d.d[k] = dm_inv4[k];
    V3 xs[3], g[3];
    readPoints<3>(x9, xs);
    orc::evalStretch(xs, d, which, *C, g);
    for (int a = 0; a < 3; ++a) {
      grad9[3 * a] = g[a].x;
      grad9[3 * a + 1] = g[a].y;
      grad9[3 * a + 2] = g[a].z;
    }
  });
}

int oracle_dihedral(const double* x12, double* theta, double* grad12) {
  return guard([&] {
    V3 xs[4], g[4];
    readPoints<4>(x12, xs);
    orc::evalDihedral(xs, *theta, g);
    for (int a = 0; a < 4; ++a) {
      grad12[3 * a] = g[a].x;
      grad12[3 * a + 1] = g[a].y;
      grad12[3 * a + 2] = g[a].z;
    }
  });
}

int oracle_bending_coefficients(double bend_warp, double bend_weft, double* p, double* q) {
  return guard([&] { orc::bendingCoefficients(bend_warp, bend_weft, *p, *q); });
}

double oracle_edge_bend_stiffness(double p, double q, const double* edge2, const double* grain2) {
  return orc::edgeBendStiffness(p, q, edge2, grain2);
}

int oracle_field_query(const drape_field_desc* field, uint32_t n, const double* p, double* phi, double* normal) {
  return guard([&] {
    const orc::Field f = orc::fieldFromDesc(field);
    for (uint32_t i = 0; i < n; ++i) {
      const V3 q{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
      if (phi) phi[i] = f.sample(q);
      if (normal) {
        const V3 g = f.gradient(q);
        normal[3 * i] = g.x;
        normal[3 * i + 1] = g.y;
        normal[3 * i + 2] = g.z;
      }
    }
  });
}

int oracle_collide_body(const drape_field_desc* field, float radius, float friction, const float* x_prev, float* x) {
  return guard([&] {
    const orc::Field f = orc::fieldFromDesc(field);
    orc::V3f xv{x[0], x[1], x[2]};
    orc::collideBody(f, radius, friction, {x_prev[0], x_prev[1], x_prev[2]}, xv);
    x[0] = xv.x;
    x[1] = xv.y;
    x[2] = xv.z;
  });
}

int oracle_land_on_body(const drape_field_desc* field, float radius, const float* from, const float* to, float* out) {
  return guard([&] {
    const orc::Field f = orc::fieldFromDesc(field);
    const orc::V3f r = orc::landOnBody(f, radius, {from[0], from[1], from[2]}, {to[0], to[1], to[2]});
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
  });
}

int oracle_invert_spd3(const double* a9, double* inv9) {
  M3 A, inv;
  std::memcpy(A.a, a9, sizeof(double) * 9);
  const bool ok = orc::invertSpd3(A, inv);
  if (ok) std::memcpy(inv9, inv.a, sizeof(double) * 9);
  return ok ? 1 : 0;
}

int oracle_bsr_pcg_contact(uint32_t n, const uint32_t* row_start, const uint32_t* col_index, const double* blocks,
                           const uint8_t* pinned, const uint8_t* active, const double* normals, const double* rhs,
                           double tol, int32_t max_iter, double* x, int32_t* iterations, double* contact_force) {
  return guard([&] {
    orc::Solver so;
    so.m.x.assign(n, orc::V3f{});
    so.m.pinned.assign(n, 0);
    if (pinned) so.m.pinned.assign(pinned, pinned + n);
    so.m.rowStart.assign(row_start, row_start + n + 1);
    so.m.colIndex.assign(col_index, col_index + row_start[n]);
    so.m.diagSlot.resize(n);
    for (uint32_t i = 0; i < n; ++i) so.m.diagSlot[i] = so.m.slot(i, i);
    so.m.contactForce.assign(n, 0.0);
    so.m.settings.cgTolerance = tol;
    so.m.settings.maxCgIterations = max_iter;
    so.allocate();
    if (active) {
      if (!normals) throw std::invalid_argument("contact normals required with active contacts");
      for (uint32_t i = 0; i < n; ++i) {
        so.contactActive[i] = active[i] ? 1 : 0;
        so.contactNormal[i] = {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
      }
    }
    for (size_t k = 0; k < so.blocks.size(); ++k) std::memcpy(so.blocks[k].a, blocks + 9 * k, sizeof(double) * 9);
    for (uint32_t i = 0; i < n; ++i) so.grad[i] = {-rhs[3 * i], -rhs[3 * i + 1], -rhs[3 * i + 2]};
    so.solveLinear();
    for (uint32_t i = 0; i < n; ++i) {
      x[3 * i] = so.step[i].x;
      x[3 * i + 1] = so.step[i].y;
      x[3 * i + 2] = so.step[i].z;
    }
    if (iterations) *iterations = static_cast<int32_t>(so.counters.pcg_iterations);
    if (contact_force)
      for (uint32_t i = 0; i < n; ++i) contact_force[i] = so.m.contactForce[i];
  });
}

int oracle_bsr_pcg(uint32_t n, const uint32_t* row_start, const uint32_t* col_index, const double* blocks,
                   const uint8_t* pinned, const double* rhs, double tol, int32_t max_iter, double* x,
                   int32_t* iterations) {
  return oracle_bsr_pcg_contact(n, row_start, col_index, blocks, pinned, nullptr, nullptr, rhs, tol, max_iter, x,
                                iterations, nullptr);
}


int oracle_bake_field_grid(const drape_body_mesh* components, uint32_t n, const drape_bake_options* options,
                           drape_field_desc* out) {
  return guard([&] {
    const bk::Grid g = bk::gridOf(bk::meshes(components, n), options);
    *out = drape_field_desc{};
    out->origin[0] = g.origin.x;
    out->origin[1] = g.origin.y;
    out->origin[2] = g.origin.z;
    for (int a = 0; a < 3; ++a) out->dims[a] = g.dims[a];
    out->cell_size = g.cell;
    out->band = g.band;
  });
}

int oracle_bake_collision_field(const drape_body_mesh* components, uint32_t n, const drape_bake_options* options,
                                float* values) {
  return guard([&] {
    const auto ms = bk::meshes(components, n);
    bk::bake(ms, bk::gridOf(ms, options), values);
  });
}

// vertexNormals (libs/render/src/normals.cpp:5-24). FP32 as in the reference's render path; each product and
// difference rounded separately (this build uses no fused multiply-add on the host).
int oracle_vertex_normals(const float* positions, uint32_t n, const uint32_t* triangles, uint32_t n_triangles,
                          const uint32_t* weld_pairs, uint32_t n_welds, float* out) {
  return guard([&] {
    std::vector<float> acc(3 * size_t(n), 0.0f);
    auto P = [&](uint32_t v, int k) {
      if (v >= n) throw std::out_of_range("normals: vertex out of range");
      return positions[3 * size_t(v) + k];
    };
    for (uint32_t t = 0; t < n_triangles; ++t) {
      const uint32_t* tri = triangles + 3 * size_t(t);
      const float ux = P(tri[1], 0) - P(tri[0], 0), uy = P(tri[1], 1) - P(tri[0], 1), uz = P(tri[1], 2) - P(tri[0], 2);
      const float vx = P(tri[2], 0) - P(tri[0], 0), vy = P(tri[2], 1) - P(tri[0], 1), vz = P(tri[2], 2) - P(tri[0], 2);
      const float c[3] = {uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx};
      for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) acc[3 * size_t(tri[k]) + a] += c[a];
    }
    for (uint32_t w = 0; w < n_welds; ++w) {
      const uint32_t a = weld_pairs[2 * size_t(w)], b = weld_pairs[2 * size_t(w) + 1];
      if (a >= n || b >= n) throw std::out_of_range("normals: weld vertex out of range");
      for (int k = 0; k < 3; ++k) {
        const float s = acc[3 * size_t(a) + k] + acc[3 * size_t(b) + k];
        acc[3 * size_t(a) + k] = s;
        acc[3 * size_t(b) + k] = s;
      }
    }
    for (uint32_t v = 0; v < n; ++v) {
      const float* a = &acc[3 * size_t(v)];
      const float len = std::sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
      for (int k = 0; k < 3; ++k) out[3 * size_t(v) + k] = len > 0.0f ? a[k] / len : (k == 1 ? 1.0f : 0.0f);
    }
  });
}

}  // extern "C"
