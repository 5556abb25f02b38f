// normals.hpp — per-vertex shading normals of the cloth on the device, the output side of the step for a live
// viewer: vertexNormals (libs/render/include/drape/render/normals.hpp:13-18, libs/render/src/normals.cpp:5-24).
//
// Semantics kept from the reference: each vertex sums the unnormalised cross products (x1 - x0) x (x2 - x0) of its
// triangles in triangle order (area weighting); then the weld pairs (a seam's two coincident particles) are taken
// in order, each setting both of its vertices to the sum of their two normals; then every normal is normalised, a
// vertex without area getting (0, 1, 0).
//
// Device form: a vertex's incident triangles are listed in increasing triangle order (CSR), so one thread per
// vertex sums them in the reference's order. The weld pairs are split, in order, into layers in which no vertex
// repeats: the pairs of a layer touch disjoint vertices and run in parallel, the layers run one after another, and
// the result equals the sequential pass. FP32 throughout, products and sums not fused.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_buffer.hpp"
#include "layout.hpp"

namespace nrm {

// Limits of the allocation-free form (its fixed working array: about 11.5 MB per GPU).
constexpr uint32_t kNoAllocMaxVertices = 1u << 18;   // 262,144
constexpr uint32_t kNoAllocMaxTriangles = 1u << 19;  // 524,288

// The allocation-free form: positions (3 floats per vertex), triangles, welds and out are device arrays; no
// memory is allocated. Calls that overlap on one GPU share its working array and race. Synchronises `st`.
// std::invalid_argument beyond the limits, std::out_of_range for an index >= n (the result is then undefined).
void vertexNormalsNoAlloc(const float* positions, uint32_t n, const uint32_t* triangles, uint32_t nTriangles,
                          const uint32_t* welds, uint32_t nWelds, float* out, hipStream_t st);

class Mesh {
 public:
  // n vertices; triangles: 3 vertex indices each; welds: (a, b) pairs. Indices are checked (std::out_of_range).
  Mesh(uint32_t n, const uint32_t* triangles, uint32_t nTriangles, const uint32_t* welds, uint32_t nWelds);
  uint32_t vertices() const { return n_; }
  // out[v] = unit normal of vertex v for positions pos (both device arrays of n entries), on stream st.
  void compute(const dk::F3* pos, dk::F3* out, hipStream_t st) const;

 private:
  uint32_t n_ = 0, nLayers_ = 0;
  std::vector<uint32_t> layerStart_;  // host copy: nLayers + 1
  dh::DeviceBuffer tri_, vtxStart_, vtxTri_, weldA_, weldB_, layerStartDev_, acc_;
};

}  // namespace nrm
