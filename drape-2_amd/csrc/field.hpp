// field.hpp — device queries of a body's signed-distance grid and the post-solve contact corrections.
//
// Semantics of the reference's CollisionField (libs/core/src/collision_field.cpp:8-41) and collision helpers
// (libs/sim/src/collision.cpp:5-34; the landing on the body, :24-34): a sample is the trilinear interpolation of the grid (values at
// origin + (i, j, k) cell, x fastest); a point outside the sampled box (or NaN) reads +band; the gradient is the
// normalised central difference with step one cell, zero where the field is flat (|g| < 1e-4 in FP32, the
// reference GPU path's threshold: FP32 trilinear weights leave about 1e-6 of slope noise in a clamped region, a
// distance field's true slope is about 1).
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"
#include "terms.hpp"

namespace dk {

// Grid coordinate of a point along one axis of n samples: the cell (clamped to the last full cell, so the top
// sample is reached with fraction 1) and the fraction within it. False outside [0, n - 1] or for NaN.
template <class T>
__device__ inline bool gridCell(T coord, uint32_t n, uint32_t& cell, T& frac) {
  if (n < 2u || !(coord >= T(0) && coord <= T(n - 1u))) return false;
  const uint32_t c = uint32_t(floor(coord));
  cell = c < n - 2u ? c : n - 2u;
  frac = coord - T(cell);
  return true;
}

template <class T>
__device__ inline T lerpT(T a, T b, T t) {
  return a + t * (b - a);
}

// Trilinear sample in precision T at grid coordinates (gx, gy, gz) (position minus origin, over the cell size).
template <class T>
__device__ inline T gridSample(const float* values, const Field32& f, T gx, T gy, T gz, T outside) {
  uint32_t i, j, k;
  T tx, ty, tz;
  if (!gridCell(gx, f.dx, i, tx) || !gridCell(gy, f.dy, j, ty) || !gridCell(gz, f.dz, k, tz)) return outside;
  const float* base = values + f.offset + (size_t(k) * f.dy + j) * f.dx + i;
  const size_t sy = f.dx, sz = size_t(f.dx) * f.dy;
  T plane[2];
#pragma unroll
  for (int dz = 0; dz < 2; ++dz) {
    const float* p = base + dz * sz;
    const T lo = lerpT(T(p[0]), T(p[1]), tx);
    const T hi = lerpT(T(p[sy]), T(p[sy + 1]), tx);
    plane[dz] = lerpT(lo, hi, ty);
  }
  return lerpT(plane[0], plane[1], tz);
}

// FP32 signed distance at p (the band without a body).
__device__ inline float bodyDistance(const float* values, const Field32& f, F3 p) {
  if (!f.present) return f.band;
  const float inv = 1.0f / f.cell;
  return gridSample<float>(values, f, (p.x - f.ox) * inv, (p.y - f.oy) * inv, (p.z - f.oz) * inv, f.band);
}

// Unit outward direction of the field at p (zero where flat or outside the grid).
__device__ inline F3 bodyNormal(const float* values, const Field32& f, F3 p) {
  const float h = f.cell;
  float d[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const F3 step = f3(a == 0 ? h : 0.0f, a == 1 ? h : 0.0f, a == 2 ? h : 0.0f);
    d[a] = (bodyDistance(values, f, p + step) - bodyDistance(values, f, p - step)) / (2.0f * h);
  }
  const F3 g = f3(d[0], d[1], d[2]);
  const float len = norm(g);
  return len < 1e-4f ? f3(0, 0, 0) : (1.0f / len) * g;
}

// A particle at distance `dist` < radius from the body, where the field direction is n: moved out along n to the
// radius, then the tangential part of its motion since `from` is resisted by Coulomb friction with the push-out as
// the normal impulse (it keeps only the part beyond mu * depth) (collision.cpp:5-22).
__device__ inline void pushOut(float dist, F3 n, float radius, float mu, F3 from, F3& x) {
  const float depth = radius - dist;
  x += depth * n;
  const F3 moved = x - from;
  const F3 slide = moved - dot(moved, n) * n;
  const float len = norm(slide);
  const float grip = mu * depth;
  const float removed = len <= grip ? 1.0f : grip / len;
  x -= removed * slide;
}

// Where the field gives no direction: the last point of the segment from -> to still `radius` away from the body
// (bisection on the segment parameter, 40 halvings), or `from` itself when it is already closer.
__device__ inline F3 stopAtBody(const float* values, const Field32& f, float radius, F3 from, F3 to) {
  if (!(bodyDistance(values, f, from) >= radius)) return from;
  const F3 path = to - from;
  float clear = 0.0f, blocked = 1.0f;
  for (int i = 0; i < 40; ++i) {
    const float probe = 0.5f * (clear + blocked);
    const bool free = bodyDistance(values, f, from + probe * path) >= radius;
    clear = free ? probe : clear;
    blocked = free ? blocked : probe;
  }
  return from + clear * path;
}

}  // namespace dk
