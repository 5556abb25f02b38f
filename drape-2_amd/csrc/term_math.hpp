// term_math.hpp — FP32 device math of the cloth energy terms, the body field and the 3x3 SPD inverse.
//
// Each function states the reference expression it evaluates (energies.cpp, stretch.hpp, bending.hpp,
// collision_field.cpp, collision.cpp, kernels/term_math.h); the arithmetic is written for one lane per row: a row
// owner evaluates a term and keeps only its own row of the gradient and Hessian (block (a, b) for b = 0..N-1),
// so no Hessian is ever stored per term.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace dk {

struct M3 {
  float m[9];  // row-major
};

__device__ inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ inline F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ inline F3 operator*(float s, F3 a) { return F3{s * a.x, s * a.y, s * a.z}; }
__device__ inline F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ inline F3 operator/(F3 a, float s) { return F3{a.x / s, a.y / s, a.z / s}; }
__device__ inline F3& operator+=(F3& a, F3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
__device__ inline F3& operator-=(F3& a, F3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
__device__ inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline F3 cross(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline float len2(F3 a) { return dot(a, a); }
__device__ inline float len(F3 a) { return sqrtf(dot(a, a)); }
__device__ inline bool isZero(F3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ inline float comp(F3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

__device__ inline F3 ld3(const F3* p, uint32_t i) { return p[i]; }
__device__ inline void st3(F3* p, uint32_t i, F3 v) { p[i] = v; }

__device__ inline M3 zeroM() {
  M3 r;
#pragma unroll
  for (int k = 0; k < 9; ++k) r.m[k] = 0.0f;
  return r;
}
__device__ inline M3 identityScaled(float s) {
  M3 r = zeroM();
  r.m[0] = r.m[4] = r.m[8] = s;
  return r;
}
__device__ inline F3 mul(const M3& a, F3 v) {
  return F3{a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
// h += K (u v^T): the product u_r v_c is formed first, as scale(outer(u, v), K) does (kernels/term_math.h:20-27).
__device__ inline void addOuterScaled(M3& h, F3 u, F3 v, float K) {
  const float uu[3] = {u.x, u.y, u.z}, vv[3] = {v.x, v.y, v.z};
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) h.m[3 * r + c] += (uu[r] * vv[c]) * K;
}
__device__ inline void addScaled(M3& h, const M3& p, float s) {
#pragma unroll
  for (int k = 0; k < 9; ++k) h.m[k] += p.m[k] * s;
}
// s (I - n n^T)
__device__ inline M3 projector(F3 n, float s) {
  const float nn[3] = {n.x, n.y, n.z};
  M3 r;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) r.m[3 * a + b] = ((a == b ? 1.0f : 0.0f) - nn[a] * nn[b]) * s;
  return r;
}

// Removes the normal component at particles held on their contact plane (filter, cpu_solver.cpp:69-73).
__device__ inline F3 filterContact(uint32_t flags, float4 cn, F3 v) {
  if (!(flags & kContactActive)) return v;
  const F3 n = F3{cn.x, cn.y, cn.z};
  return v - n * dot(n, v);
}

// A row's block-Jacobi preconditioner is stored as the upper triangle of the inverse diagonal block (00, 01, 02,
// 11, 12, 22: the inverse of a symmetric block is symmetric) and applied with the contact filter after it,
// filter(M x) (cpu_solver.cpp:179-181).
__device__ inline F3 precondApply(const float* m, uint32_t flags, float4 cn, F3 x) {
  const F3 y = F3{m[0] * x.x + m[1] * x.y + m[2] * x.z, m[1] * x.x + m[3] * x.y + m[4] * x.z,
                  m[2] * x.x + m[4] * x.y + m[5] * x.z};
  return filterContact(flags, cn, y);
}

// ---- Stretch (stretch.hpp:15-49, energies.cpp:10-62) ----

// Coefficient of vertex `local` in f (column 0: f_u, 1: f_v) = sum_k d_k x_k.
__device__ inline float vertexCoefficient(const float d[4], int column, int local) {
  const float c0 = column == 0 ? d[0] : d[2];  // dmInv(0, column)
  const float c1 = column == 0 ? d[1] : d[3];  // dmInv(1, column)
  if (local == 1) return c0;
  if (local == 2) return c1;
  return -(c0 + c1);
}

// which: 0 warp (|f_u| - 1), 1 weft (|f_v| - 1), 2 bias (n_u . n_v); grad of C w.r.t. the three vertices.
__device__ inline void stretchStrain(const F3 x[3], const float d[4], int which, float& C, F3 g[3]) {
  const F3 e1 = x[1] - x[0], e2 = x[2] - x[0];
  const F3 fu = e1 * d[0] + e2 * d[1];
  const F3 fv = e1 * d[2] + e2 * d[3];
  const float lu = len(fu), lv = len(fv);
  const float tiny = 1e-12f;
  F3 dfu = f3(0, 0, 0), dfv = f3(0, 0, 0);
  C = 0.0f;
  if (which == 0) {
    if (lu > tiny) {
      C = lu - 1.0f;
      dfu = fu / lu;
    }
  } else if (which == 1) {
    if (lv > tiny) {
      C = lv - 1.0f;
      dfv = fv / lv;
    }
  } else if (lu > tiny && lv > tiny) {
    const F3 nu = fu / lu, nv = fv / lv;
    const float c = dot(nu, nv);
    C = c;
    dfu = (nv - c * nu) / lu;
    dfv = (nu - c * nv) / lv;
  }
  g[1] = dfu * d[0] + dfv * d[2];
  g[2] = dfu * d[1] + dfv * d[3];
  g[0] = -(g[1] + g[2]);
}

// Row `a` of one stretch term: energy, gradient of vertex a, blocks (a, 0..2). Energy only when !Deriv.
template <bool Deriv>
__device__ inline float stretchRow(const Stretch32& t, const F3 x[3], const Fabric32& f, float limitScale, int a,
                                   F3& ga, M3 H[3]) {
  const float k[3] = {f.stretchWarp, f.stretchWeft, f.stretchBias};
  const float limit[2] = {f.strainLimitWarp, f.strainLimitWeft};
  const F3 e1 = x[1] - x[0], e2 = x[2] - x[0];
  const F3 fcol[2] = {e1 * t.dmInv[0] + e2 * t.dmInv[1], e1 * t.dmInv[2] + e2 * t.dmInv[3]};
  float energy = 0.0f;
  for (int which = 0; which < 3; ++which) {
    if (k[which] <= 0.0f) continue;
    F3 g[3];
    float C;
    stretchStrain(x, t.dmInv, which, C, g);
    const float K = k[which] * t.area;
    energy += 0.5f * K * C * C;
    if (Deriv) {
      ga += (K * C) * g[a];
#pragma unroll
      for (int b = 0; b < 3; ++b) addOuterScaled(H[b], g[a], g[b], K);
    }
    if (which == 2) continue;
    // Tension stiffness K C d_i d_j (I - n n^T) / |f| for C > 0 (energies.cpp:29-39).
    const float lenf = len(fcol[which]);
    if (Deriv && C > 0.0f && !(lenf < 1e-12f)) {
      const M3 P = projector(fcol[which] / lenf, K * C / lenf);
      const float ca = vertexCoefficient(t.dmInv, which, a);
#pragma unroll
      for (int b = 0; b < 3; ++b) addScaled(H[b], P, ca * vertexCoefficient(t.dmInv, which, b));
    }
    if (limit[which] > 0.0f && C > limit[which]) {
      const float KL = limitScale * K;
      const float excess = C - limit[which];
      energy += 0.5f * KL * excess * excess;
      if (Deriv) {
        ga += (KL * excess) * g[a];
#pragma unroll
        for (int b = 0; b < 3; ++b) addOuterScaled(H[b], g[a], g[b], KL);
        if (!(lenf < 1e-12f)) {
          const M3 P = projector(fcol[which] / lenf, KL * excess / lenf);
          const float ca = vertexCoefficient(t.dmInv, which, a);
#pragma unroll
          for (int b = 0; b < 3; ++b) addScaled(H[b], P, ca * vertexCoefficient(t.dmInv, which, b));
        }
      }
    }
  }
  return energy;
}

// ---- Hinge (bending.hpp:29-51, energies.cpp:64-69) ----

// theta = atan2((m1 x m2) . e, m1 . m2) and its gradient (Bridson, Marino & Fedkiw 2003, sign flipped).
__device__ inline float dihedral(const F3 x[4], F3 g[4]) {
  const F3 n1 = cross(x[0] - x[2], x[0] - x[3]);
  const F3 n2 = cross(x[1] - x[3], x[1] - x[2]);
  const F3 E = x[3] - x[2];
  const float n1sq = len2(n1), n2sq = len2(n2), elen = len(E);
  if (n1sq < 1e-24f || n2sq < 1e-24f || elen < 1e-12f) {
    g[0] = g[1] = g[2] = g[3] = f3(0, 0, 0);
    return 0.0f;
  }
  const F3 e = E / elen;
  const F3 m1 = n1 / sqrtf(n1sq), m2 = n2 / sqrtf(n2sq);
  const float theta = atan2f(dot(cross(m1, m2), e), dot(m1, m2));
  const F3 a = n1 / n1sq, b = n2 / n2sq;
  g[0] = -elen * a;
  g[1] = -elen * b;
  g[2] = -((dot(x[0] - x[3], E) / elen) * a + (dot(x[1] - x[3], E) / elen) * b);
  g[3] = (dot(x[0] - x[2], E) / elen) * a + (dot(x[1] - x[2], E) / elen) * b;
  return theta;
}

template <bool Deriv>
__device__ inline float hingeRow(const Hinge32& t, const F3 x[4], int a, F3& ga, M3 H[4]) {
  F3 g[4];
  const float theta = dihedral(x, g);
  const float C = theta - t.restAngle;
  if (Deriv) {
    ga += (t.K * C) * g[a];
#pragma unroll
    for (int b = 0; b < 4; ++b) addOuterScaled(H[b], g[a], g[b], t.K);
  }
  return 0.5f * t.K * C * C;
}

// ---- Seam spring (energies.cpp:71-91) ----

template <bool Deriv>
__device__ inline float seamRow(const F3 x[2], float target, float K, int a, F3& ga, M3 H[2]) {
  const F3 diff = x[0] - x[1];
  const float d = len(diff);
  if (d <= target) return 0.0f;
  float energy = 0.0f;
  M3 block;
  if (d < 1e-12f) {
    block = identityScaled(K);
  } else {
    const F3 n = diff / d;
    const float ext = d - target;
    energy = 0.5f * K * ext * ext;
    if (Deriv) {
      const F3 gn = (K * ext) * n;
      ga += a == 0 ? gn : -gn;
      // K (n n^T + (ext / d) (I - n n^T))
      const float nn[3] = {n.x, n.y, n.z};
      const float s = ext / d;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float o = nn[r] * nn[c];
          block.m[3 * r + c] = (o + ((r == c ? 1.0f : 0.0f) - o) * s) * K;
        }
    }
  }
  if (Deriv) {
    addScaled(H[a], block, 1.0f);
    addScaled(H[1 - a], block, -1.0f);
  }
  return energy;
}

// ---- Friction (energies.cpp:93-110) ----

// Smoothed Coulomb friction with dx = x - x0 and contact normal n; adds to energy/grad/H.
template <bool Deriv>
__device__ inline float frictionTerm(F3 dx, F3 n, float muLambda, float eps, F3& g, M3& H) {
  if (muLambda <= 0.0f || eps <= 0.0f) return 0.0f;
  const M3 P = projector(n, 1.0f);
  const F3 u = mul(P, dx);
  const float s = len(u);
  if (s < eps) {
    const float e = muLambda * (-s * s * s / (3.0f * eps * eps) + s * s / eps + eps / 3.0f);
    if (Deriv) {
      const float f1OverS = 2.0f / eps - s / (eps * eps);
      g += (muLambda * f1OverS) * u;
      addScaled(H, P, muLambda * f1OverS);
      if (s > 0.0f) addOuterScaled(H, u, u, -(muLambda / (eps * eps * s)));
    }
    return e;
  }
  const F3 dir = u / s;
  if (Deriv) {
    g += muLambda * dir;
    M3 Q = P;
    addOuterScaled(Q, dir, dir, -1.0f);
    addScaled(H, Q, muLambda / s);
  }
  return muLambda * s;
}

// ---- Body field (collision_field.cpp:8-41; kernels/term_math.h:203-243) ----

__device__ inline float fieldSample(const float* values, const Field32& f, F3 p) {
  if (!f.present) return f.band;
  const float l[3] = {(p.x - f.ox) / f.cell, (p.y - f.oy) / f.cell, (p.z - f.oz) / f.cell};
  const uint32_t dims[3] = {f.dx, f.dy, f.dz};
  int i[3];
  float t[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float maxCoord = float(dims[a]) - 1.0f;
    if (!(l[a] >= 0.0f && l[a] <= maxCoord)) return f.band;
    const int fl = int(floorf(l[a]));
    const int top = int(dims[a]) - 2;
    i[a] = fl < top ? fl : top;
    t[a] = l[a] - float(i[a]);
  }
  const float* v = values + f.offset;
  float result = 0.0f;
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const float w = (dx ? t[0] : 1.0f - t[0]) * (dy ? t[1] : 1.0f - t[1]) * (dz ? t[2] : 1.0f - t[2]);
        const size_t idx = (size_t(uint32_t(i[2] + dz)) * f.dy + uint32_t(i[1] + dy)) * f.dx + uint32_t(i[0] + dx);
        result += w * v[idx];
      }
  return result;
}

// Central differences with step cell, normalised; zero where |g| < 1e-4 (the FP32 flatness threshold of the
// reference's GPU path, kernels/term_math.h:178-181: in float the trilinear weights of a clamped region leave
// ~1e-6 of noise, the true slope of a distance field is ~1).
__device__ inline F3 fieldGradient(const float* values, const Field32& f, F3 p) {
  const float h = f.cell;
  const F3 g = f3((fieldSample(values, f, p + f3(h, 0, 0)) - fieldSample(values, f, p - f3(h, 0, 0))) / (2.0f * h),
                  (fieldSample(values, f, p + f3(0, h, 0)) - fieldSample(values, f, p - f3(0, h, 0))) / (2.0f * h),
                  (fieldSample(values, f, p + f3(0, 0, h)) - fieldSample(values, f, p - f3(0, 0, h))) / (2.0f * h));
  const float n = len(g);
  return n < 1e-4f ? f3(0, 0, 0) : g / n;
}

// collideBody (collision.cpp:5-22): push out along the gradient, then positional Coulomb friction.
__device__ inline void collideBody(const float* values, const Field32& f, float radius, float friction, F3 xPrev,
                                   F3& x) {
  const float phi = fieldSample(values, f, x);
  if (phi >= radius) return;
  const F3 n = fieldGradient(values, f, x);
  if (isZero(n)) return;
  const float depth = radius - phi;
  x += depth * n;
  const F3 delta = x - xPrev;
  const F3 tangential = delta - dot(delta, n) * n;
  const float t = len(tangential);
  if (t <= friction * depth) {
    x -= tangential;
  } else if (t > 0.0f) {
    x -= tangential * (friction * depth / t);
  }
}

// landOnBody (collision.cpp:24-34): last point of from -> to at least radius from the body (40 bisections).
__device__ inline F3 landOnBody(const float* values, const Field32& f, float radius, F3 from, F3 to) {
  if (!(fieldSample(values, f, from) >= radius)) return from;
  float lo = 0.0f, hi = 1.0f;
  for (int i = 0; i < 40; ++i) {
    const float mid = 0.5f * (lo + hi);
    if (fieldSample(values, f, from + mid * (to - from)) >= radius)
      lo = mid;
    else
      hi = mid;
  }
  return from + lo * (to - from);
}

// ---- 3x3 symmetric positive definite inverse by LDL^T without pivoting (kernels/term_math.h:285-312); false
// (and no output) when a pivot is not positive. Reads the lower triangle of a row-major block.
__device__ inline bool invertSpd3(const float* A, float* inv) {
  const float a00 = A[0], a10 = A[3], a20 = A[6], a11 = A[4], a21 = A[7], a22 = A[8];
  const float d0 = a00;
  if (!(d0 > 0.0f)) return false;
  const float l10 = a10 / d0, l20 = a20 / d0;
  const float d1 = a11 - l10 * l10 * d0;
  if (!(d1 > 0.0f)) return false;
  const float l21 = (a21 - l20 * l10 * d0) / d1;
  const float d2 = a22 - l20 * l20 * d0 - l21 * l21 * d1;
  if (!(d2 > 0.0f)) return false;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float e0 = k == 0 ? 1.0f : 0.0f, e1 = k == 1 ? 1.0f : 0.0f, e2 = k == 2 ? 1.0f : 0.0f;
    const float y0 = e0, y1 = e1 - l10 * y0, y2 = e2 - l20 * y0 - l21 * y1;
    const float z0 = y0 / d0, z1 = y1 / d1, z2 = y2 / d2;
    const float x2 = z2, x1 = z1 - l21 * x2, x0 = z0 - l10 * x1 - l20 * x2;
    inv[0 + k] = x0;  // column k, row-major storage
    inv[3 + k] = x1;
    inv[6 + k] = x2;
  }
  return true;
}

}  // namespace dk
