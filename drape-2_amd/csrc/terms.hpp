// terms.hpp — FP32 device evaluation of the cloth energy terms for the MI355X Newton step.
//
// Physics follows the reference's FP64 definitions (stretch: libs/sim/src/stretch.hpp:15-49 and
// energies.cpp:10-62; hinge: bending.hpp:29-51 and energies.cpp:64-69; seam spring: energies.cpp:71-91; friction:
// energies.cpp:93-110; inertia: cpu_solver.cpp:99-108). The formulation here is this build's own and is shaped
// for the row-owner assembly: every term is reduced once to a few per-term quantities, from which any vertex
// row forms its gradient and its 3x3 blocks directly.
//
//  * Stretch is written through the deformation-gradient columns f_u = sum_k cu_k x_k and f_v = sum_k cv_k x_k
//    (cu, cv: the vertex coefficients given by Dm^-1). One evaluation gives dE/df_u, dE/df_v and the three 3x3
//    second-derivative blocks Huu, Hvv, Huv; vertex row a then has gradient cu_a dE/df_u + cv_a dE/df_v and
//    block (a, b) = cu_b R_a + cv_b Q_a with R_a = cu_a Huu + cv_a Huv^T, Q_a = cv_a Hvv + cu_a Huv.
//  * The dihedral angle is atan2 of the unnormalised sine and cosine of the two face normals; the gradients of
//    the edge vertices follow from translation and rotation invariance (each wing's gradient shared in proportion
//    to where the wing vertex projects onto the hinge edge).
//  * The seam block K (n n^T + (ext / d)(I - n n^T)) is K ((1 - t/d) I + (t/d) n n^T), t the target gap.
//  * Friction: E = mu lambda f0(s) of the tangential slip s, gradient mu lambda phi(s) u and Hessian
//    mu lambda (phi P + psi u u^T), with phi = f0'(s) / s and psi = phi'(s) / s.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace dk {

// ---- small vector algebra ----
__device__ inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ inline F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ inline F3 operator*(float s, F3 a) { return F3{s * a.x, s * a.y, s * a.z}; }
__device__ inline F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ inline F3& operator+=(F3& a, F3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
__device__ inline F3& operator-=(F3& a, F3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
__device__ inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline F3 cross(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline float norm2(F3 a) { return dot(a, a); }
__device__ inline float norm(F3 a) { return sqrtf(dot(a, a)); }
__device__ inline bool allZero(F3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ inline float axis(F3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Element a (0..3, a runtime value) of a small per-term array without indexing a private array (which would live
// in scratch memory): a chain of selects.
template <class T>
__device__ inline T pick(uint32_t a, T v0, T v1, T v2, T v3) {
  return a == 0u ? v0 : (a == 1u ? v1 : (a == 2u ? v2 : v3));
}

// A 3x3 block, row-major (element (r, c) at 3 r + c), the layout the matrix stores.
struct Blk {
  float e[9];
};
__device__ inline Blk blkZero() {
  Blk b;
#pragma unroll
  for (int k = 0; k < 9; ++k) b.e[k] = 0.0f;
  return b;
}
// s I + t w w^T
__device__ inline Blk blkIsoPlusRank1(float s, float t, F3 w) {
  const float ww[3] = {w.x, w.y, w.z};
  Blk b;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) b.e[3 * r + c] = t * (ww[r] * ww[c]) + (r == c ? s : 0.0f);
  return b;
}
// p q^T scaled by s
__device__ inline Blk blkOuter(F3 p, F3 q, float s) {
  const float pp[3] = {p.x, p.y, p.z}, qq[3] = {q.x, q.y, q.z};
  Blk b;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) b.e[3 * r + c] = s * (pp[r] * qq[c]);
  return b;
}
__device__ inline void blkAxpy(Blk& y, float s, const Blk& x) {
#pragma unroll
  for (int k = 0; k < 9; ++k) y.e[k] += s * x.e[k];
}
__device__ inline void blkAxpyT(Blk& y, float s, const Blk& x) {  // y += s x^T
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) y.e[3 * r + c] += s * x.e[3 * c + r];
}
__device__ inline F3 blkMul(const Blk& b, F3 v) {
  return f3(b.e[0] * v.x + b.e[1] * v.y + b.e[2] * v.z, b.e[3] * v.x + b.e[4] * v.y + b.e[5] * v.z,
            b.e[6] * v.x + b.e[7] * v.y + b.e[8] * v.z);
}

// ---- stretch (membrane) ----

// Shape coefficients of a triangle: f_u = sum_k cu[k] x_k, f_v = sum_k cv[k] x_k, from Dm^-1 stored as
// (0,0), (1,0), (0,1), (1,1).
struct Shape3 {
  float cu[3], cv[3];
};
__device__ inline Shape3 shapeOf(const float dmInv[4]) {
  Shape3 s;
  s.cu[1] = dmInv[0];
  s.cu[2] = dmInv[1];
  s.cu[0] = -(dmInv[0] + dmInv[1]);
  s.cv[1] = dmInv[2];
  s.cv[2] = dmInv[3];
  s.cv[0] = -(dmInv[2] + dmInv[3]);
  return s;
}

// A stretch term reduced to its derivatives with respect to the two deformation-gradient columns.
struct StretchForm {
  float energy;
  F3 gu, gv;          // dE/df_u, dE/df_v
  Blk huu, hvv, huv;  // d2E/df_u2, d2E/df_v2, d2E/(df_u df_v)   (Gauss-Newton for the bias, exact tension)
};

// One axial strain C = |f| - 1 of stiffness K (warp or weft) with its optional limit (strain above `limit`
// adds a stiffness limitScale K): adds its energy, its gradient along f^ and its block
// tau I + (K_total - tau) f^ f^T, tau = (K C [C > 0] + K_L (C - limit)) / |f| (the tension stiffness).
template <bool Deriv>
__device__ inline void axialStrain(F3 f, float K, float limit, float limitScale, float& energy, F3& g, Blk& h) {
  const float len = norm(f);
  if (!(len > 1e-12f)) return;  // a collapsed column carries no strain (C = 0)
  const float C = len - 1.0f;
  const F3 dir = (1.0f / len) * f;
  float force = K * C;       // dE/dC
  float curv = K;            // d2E/dC2
  float tension = C > 0.0f ? K * C : 0.0f;
  energy += 0.5f * K * C * C;
  if (limit > 0.0f && C > limit) {
    const float KL = limitScale * K, over = C - limit;
    energy += 0.5f * KL * over * over;
    force += KL * over;
    curv += KL;
    tension += KL * over;
  }
  if (Deriv) {
    g = force * dir;
    const float tau = tension / len;
    h = blkIsoPlusRank1(tau, curv - tau, dir);
  }
}

template <bool Deriv>
__device__ inline StretchForm stretchForm(const F3 x[3], const Shape3& s, float area, const Fabric32& fab,
                                          float limitScale) {
  StretchForm t;
  t.energy = 0.0f;
  t.gu = t.gv = f3(0, 0, 0);
  if (Deriv) t.huu = t.hvv = t.huv = blkZero();
  // f = cu[1] (x1 - x0) + cu[2] (x2 - x0): the columns from the edge vectors (cu[0] = -(cu[1] + cu[2])).
  const F3 e1 = x[1] - x[0], e2 = x[2] - x[0];
  const F3 fu = s.cu[1] * e1 + s.cu[2] * e2;
  const F3 fv = s.cv[1] * e1 + s.cv[2] * e2;
  if (fab.stretchWarp > 0.0f)
    axialStrain<Deriv>(fu, fab.stretchWarp * area, fab.strainLimitWarp, limitScale, t.energy, t.gu, t.huu);
  if (fab.stretchWeft > 0.0f)
    axialStrain<Deriv>(fv, fab.stretchWeft * area, fab.strainLimitWeft, limitScale, t.energy, t.gv, t.hvv);
  if (fab.stretchBias > 0.0f) {
    // Shear: C = cos of the angle between the columns, dC/df_u = (v^ - C u^) / |f_u|, dC/df_v likewise.
    const float lu = norm(fu), lv = norm(fv);
    if (lu > 1e-12f && lv > 1e-12f) {
      const float K = fab.stretchBias * area;
      const F3 du = (1.0f / lu) * fu, dv = (1.0f / lv) * fv;
      const float C = dot(du, dv);
      t.energy += 0.5f * K * C * C;
      if (Deriv) {
        const F3 pu = (1.0f / lu) * (dv - C * du), pv = (1.0f / lv) * (du - C * dv);
        t.gu += (K * C) * pu;
        t.gv += (K * C) * pv;
        blkAxpy(t.huu, 1.0f, blkOuter(pu, pu, K));
        blkAxpy(t.hvv, 1.0f, blkOuter(pv, pv, K));
        t.huv = blkOuter(pu, pv, K);
      }
    }
  }
  return t;
}

// Row a of a stretch term: gradient of vertex a and the factors R_a, Q_a of its blocks (block (a, b) =
// cu[b] R_a + cv[b] Q_a).
struct StretchRow {
  F3 g;
  Blk R, Q;
};
__device__ inline StretchRow stretchRow(const StretchForm& t, const Shape3& s, uint32_t a) {
  const float ca = pick(a, s.cu[0], s.cu[1], s.cu[2], 0.0f);
  const float da = pick(a, s.cv[0], s.cv[1], s.cv[2], 0.0f);
  StretchRow w;
  // Vertex 0's gradient is minus the others' (translation invariance), so a term's three gradients sum to zero.
  const F3 g1 = s.cu[1] * t.gu + s.cv[1] * t.gv, g2 = s.cu[2] * t.gu + s.cv[2] * t.gv;
  w.g = a == 0u ? -(g1 + g2) : (a == 1u ? g1 : g2);
  w.R = blkZero();
  w.Q = blkZero();
  blkAxpy(w.R, ca, t.huu);
  blkAxpyT(w.R, da, t.huv);
  blkAxpy(w.Q, da, t.hvv);
  blkAxpy(w.Q, ca, t.huv);
  return w;
}
__device__ inline Blk stretchBlock(const StretchRow& w, const Shape3& s, int b) {
  Blk out = blkZero();
  blkAxpy(out, s.cu[b], w.R);
  blkAxpy(out, s.cv[b], w.Q);
  return out;
}

// ---- hinge bending ----

// x[0], x[1] the wing vertices, x[2] -> x[3] the hinge edge. theta = atan2(sin, cos) of the angle between the
// face normals n1 = (x0 - x2) x (x0 - x3) and n2 = (x1 - x3) x (x1 - x2) about the edge direction (both
// arguments scaled by the positive |n1| |n2| |E|). Gradients (of the sign the reference uses): the wings
// -|E| n / |n|^2; the edge vertices share each wing's gradient by the wing vertex's projection s onto the edge
// (s = (x_w - x2) . E / |E|^2): start gets (s - 1) g_w, end gets -s g_w. Degenerate hinges give 0 and no gradient.
struct HingeForm {
  float theta;
  F3 g[4];
};
__device__ inline HingeForm hingeForm(const F3 x[4]) {
  HingeForm t;
  const F3 E = x[3] - x[2];
  const F3 n1 = cross(x[0] - x[2], x[0] - x[3]);
  const F3 n2 = cross(x[1] - x[3], x[1] - x[2]);
  const float q1 = norm2(n1), q2 = norm2(n2), e2 = norm2(E), el = sqrtf(e2);
  if (q1 < 1e-24f || q2 < 1e-24f || el < 1e-12f) {
    t.theta = 0.0f;
    t.g[0] = t.g[1] = t.g[2] = t.g[3] = f3(0, 0, 0);
    return t;
  }
  t.theta = atan2f(dot(cross(n1, n2), E), el * dot(n1, n2));
  const F3 w0 = (-el / q1) * n1, w1 = (-el / q2) * n2;
  const float s0 = dot(x[0] - x[2], E) / e2, s1 = dot(x[1] - x[2], E) / e2;
  t.g[0] = w0;
  t.g[1] = w1;
  t.g[2] = (s0 - 1.0f) * w0 + (s1 - 1.0f) * w1;
  t.g[3] = (-s0) * w0 + (-s1) * w1;
  return t;
}

// ---- seam spring ----

// One-sided spring between x0 and x1 above the target gap t: energy 1/2 K (d - t)^2, gradient +-K (d - t) n^
// and block S = K ((1 - t/d) I + (t/d) n^ n^T) at (a, a), -S at (a, 1 - a). A gap below 1e-12 m (only reachable
// with t < 1e-12) gives S = K I without energy.
struct SeamForm {
  bool active;
  float energy;
  F3 g0;  // gradient of x0 (x1 gets -g0)
  Blk S;
};
template <bool Deriv>
__device__ inline SeamForm seamForm(F3 x0, F3 x1, float target, float K) {
  SeamForm t;
  t.active = false;
  t.energy = 0.0f;
  t.g0 = f3(0, 0, 0);
  const F3 sep = x0 - x1;
  const float d = norm(sep);
  if (d <= target) return t;
  t.active = true;
  if (d < 1e-12f) {
    if (Deriv) t.S = blkIsoPlusRank1(K, 0.0f, f3(0, 0, 0));
    return t;
  }
  const float stretch = d - target;
  t.energy = 0.5f * K * stretch * stretch;
  if (Deriv) {
    const F3 dir = (1.0f / d) * sep;
    t.g0 = (K * stretch) * dir;
    const float slack = target / d;
    t.S = blkIsoPlusRank1(K * (1.0f - slack), K * slack, dir);
  }
  return t;
}

// ---- friction (smoothed Coulomb) ----

// Slip u = P dx with P = I - n n^T, s = |u|, eps the slip-speed scale: f0(s) = s for s >= eps, else the cubic
// -s^3/(3 eps^2) + s^2/eps + eps/3 (C1 at eps). Adds mu lambda f0, mu lambda phi u and mu lambda (phi P + psi u u^T).
template <bool Deriv>
__device__ inline float frictionAdd(F3 dx, F3 n, float muLambda, float eps, F3& g, Blk& H) {
  if (muLambda <= 0.0f || eps <= 0.0f) return 0.0f;
  const F3 u = dx - dot(n, dx) * n;
  const float s = norm(u);
  float f0, phi, psi;
  if (s < eps) {
    const float ie = 1.0f / eps;
    f0 = s * s * (ie - s * ie * ie / 3.0f) + eps / 3.0f;
    phi = 2.0f * ie - s * ie * ie;
    psi = s > 0.0f ? -(ie * ie) / s : 0.0f;
  } else {
    f0 = s;
    phi = 1.0f / s;
    psi = -phi * phi * phi;
  }
  if (Deriv) {
    g += (muLambda * phi) * u;
    blkAxpy(H, 1.0f, blkIsoPlusRank1(muLambda * phi, -muLambda * phi, n));
    blkAxpy(H, 1.0f, blkOuter(u, u, muLambda * psi));
  }
  return muLambda * f0;
}

// ---- 3x3 symmetric inverse ----

// Inverse of a symmetric 3x3 block (lower triangle of the row-major A read) by cofactors, returned as its upper
// triangle (00, 01, 02, 11, 12, 22). False when A is not positive definite by Sylvester's criterion (leading
// minors a00, a00 a11 - a10^2 and det A all positive), the case the block-Jacobi preconditioner replaces by I.
__device__ inline bool spdInverseUpper(const float* A, float inv[6]) {
  const float a = A[0], b = A[3], c = A[6], d = A[4], e = A[7], f = A[8];
  const float c00 = d * f - e * e, c01 = c * e - b * f, c02 = b * e - c * d;
  const float c11 = a * f - c * c, c12 = b * c - a * e, c22 = a * d - b * b;
  const float det = a * c00 + b * c01 + c * c02;
  if (!(a > 0.0f) || !(c22 > 0.0f) || !(det > 0.0f)) return false;
  const float r = 1.0f / det;
  inv[0] = c00 * r;
  inv[1] = c01 * r;
  inv[2] = c02 * r;
  inv[3] = c11 * r;
  inv[4] = c12 * r;
  inv[5] = c22 * r;
  return true;
}

}  // namespace dk
