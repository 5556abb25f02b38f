// bake.hpp — the body's signed distance field baked on the device from closed triangle meshes:
// bakeCollisionField (libs/avatar/include/drape/avatar/collision_bake.hpp:10-18, collision_bake.cpp:98-208).
//
// Semantics kept from the reference, per component (the field is the minimum over components):
//  - grid: origin = min vertex - padding, dims = ceil(extent / cell) + 1 per axis (extent = max - min + 2 padding);
//  - every triangle, in order, visits the samples of its bounding box grown by the band and keeps, per sample, the
//    smallest squared distance to the triangle (closest point: Ericson, RTCD 5.1.5), the first triangle winning a
//    tie; the sign is that of (p - q) . n, n the angle-weighted pseudonormal of the feature holding q (face, edge:
//    sum of its two faces, vertex: angle-weighted sum);
//  - samples beyond the band are outside when 6-connected to the grid boundary through such samples, else inside;
//  - value = clamp(sign * distance, -band, band) within the band, +-band beyond it, stored as float.
//
// Device form: FP64 throughout, no fused operations. The per-sample minimum is an atomic minimum over the
// distance's bit pattern (non-negative doubles order as their bits), then a second pass takes the smallest
// triangle index among those reaching it — the sequential loop's winner. The flood fill sweeps each axis both
// ways until nothing changes, which reaches the breadth-first search's fixpoint. Pseudonormals are computed on
// the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace bake {

struct Options {
  double cellSize = 0.004, padding = 0.05, band = 0.02;
};

struct Component {
  const double* vertices;  // 3 per vertex
  uint32_t nVertices;
  const uint32_t* triangles;  // 3 per triangle, counter-clockwise seen from outside
  uint32_t nTriangles;
};

struct Grid {
  double origin[3];
  int32_t dims[3];
  size_t count() const { return size_t(dims[0]) * size_t(dims[1]) * size_t(dims[2]); }
};

// Grid of the components' field (std::invalid_argument if a component is not a closed manifold: every
// undirected edge used by exactly two triangles, in opposite directions).
Grid gridOf(const Component* components, uint32_t n, const Options& options);

// Bakes into values (host, grid.count() floats, x fastest).
void bakeField(const Component* components, uint32_t n, const Options& options, float* values, hipStream_t st);

}  // namespace bake
