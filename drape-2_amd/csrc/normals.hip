// normals.hip — device vertex normals with seam welding (semantics in normals.hpp).
#include "normals.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace nrm {

namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("hip: ") + what + ": " + hipGetErrorString(e));
}

template <class T>
void upload(dh::DeviceBuffer& b, const std::vector<T>& v, hipStream_t st) {
  b.allocate(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (!v.empty()) check(hipMemcpyAsync(b.as<void>(), v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st), "copy");
}

}  // namespace

// (b - a) x (c - a) with every product and difference rounded separately, as the reference's FP32 Eigen code
// computes it without contraction.
__device__ inline dk::F3 triangleCross(dk::F3 a, dk::F3 b, dk::F3 c) {
#pragma clang fp contract(off)
  const float ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
  const float vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
  return dk::F3{uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx};
}

// Area-weighted sum per vertex, its triangles in increasing order (normals.cpp:9-12).
__global__ void __launch_bounds__(256) kAccumulate(uint32_t n, const dk::F3* __restrict__ pos,
                                                   const uint32_t* __restrict__ tri,
                                                   const uint32_t* __restrict__ vtxStart,
                                                   const uint32_t* __restrict__ vtxTri, dk::F3* acc) {
#pragma clang fp contract(off)
  const uint32_t v = blockIdx.x * 256u + threadIdx.x;
  if (v >= n) return;
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (uint32_t k = vtxStart[v]; k < vtxStart[v + 1]; ++k) {
    const uint32_t* t = tri + 3u * size_t(vtxTri[k]);
    const dk::F3 c = triangleCross(pos[t[0]], pos[t[1]], pos[t[2]]);
    sx = sx + c.x;
    sy = sy + c.y;
    sz = sz + c.z;
  }
  acc[v] = dk::F3{sx, sy, sz};
}

// Weld pairs layer by layer (normals.cpp:13-17): one workgroup; the pairs of a layer touch disjoint vertices.
__global__ void __launch_bounds__(256) kWeld(uint32_t nLayers, const uint32_t* __restrict__ layerStart,
                                             const uint32_t* __restrict__ weldA, const uint32_t* __restrict__ weldB,
                                             dk::F3* acc) {
#pragma clang fp contract(off)
  for (uint32_t L = 0; L < nLayers; ++L) {
    for (uint32_t i = layerStart[L] + threadIdx.x; i < layerStart[L + 1]; i += 256u) {
      const uint32_t a = weldA[i], b = weldB[i];
      const dk::F3 na = acc[a], nb = acc[b];
      const dk::F3 s{na.x + nb.x, na.y + nb.y, na.z + nb.z};
      acc[a] = s;
      acc[b] = s;
    }
    __syncthreads();  // the next layer reads what this one wrote
  }
}

// Normalise; no area -> (0, 1, 0) (normals.cpp:18-21).
__global__ void __launch_bounds__(256) kNormalize(uint32_t n, const dk::F3* __restrict__ acc, dk::F3* out) {
#pragma clang fp contract(off)
  const uint32_t v = blockIdx.x * 256u + threadIdx.x;
  if (v >= n) return;
  const dk::F3 a = acc[v];
  const float len = sqrtf((a.x * a.x + a.y * a.y) + a.z * a.z);
  out[v] = len > 0.0f ? dk::F3{a.x / len, a.y / len, a.z / len} : dk::F3{0.0f, 1.0f, 0.0f};
}

Mesh::Mesh(uint32_t n, const uint32_t* triangles, uint32_t nTriangles, const uint32_t* welds, uint32_t nWelds)
    : n_(n), nTri_(nTriangles) {
  for (size_t i = 0; i < 3 * size_t(nTriangles); ++i)
    if (triangles[i] >= n) throw std::out_of_range("normals: triangle vertex out of range");
  for (size_t i = 0; i < 2 * size_t(nWelds); ++i)
    if (welds[i] >= n) throw std::out_of_range("normals: weld vertex out of range");
  // Incident triangles per vertex, once per occurrence, in increasing triangle order.
  std::vector<uint32_t> start(size_t(n) + 1, 0), list(3 * size_t(nTriangles));
  for (size_t i = 0; i < 3 * size_t(nTriangles); ++i) ++start[triangles[i] + 1];
  for (uint32_t v = 0; v < n; ++v) start[v + 1] += start[v];
  std::vector<uint32_t> cursor(start.begin(), start.end() - 1);
  for (uint32_t t = 0; t < nTriangles; ++t)
    for (int k = 0; k < 3; ++k) list[cursor[triangles[3 * size_t(t) + k]]++] = t;
  // Weld layers: a new layer starts when a pair touches a vertex already in the current one.
  std::vector<uint32_t> a(nWelds), b(nWelds), seen(n, UINT32_MAX);
  layerStart_.assign(1, 0);
  uint32_t layer = 0;
  for (uint32_t i = 0; i < nWelds; ++i) {
    const uint32_t p = welds[2 * size_t(i)], q = welds[2 * size_t(i) + 1];
    if (seen[p] == layer || seen[q] == layer) {
      layerStart_.push_back(i);
      ++layer;
    }
    seen[p] = seen[q] = layer;
    a[i] = p;
    b[i] = q;
  }
  if (nWelds) layerStart_.push_back(nWelds);
  nLayers_ = static_cast<uint32_t>(layerStart_.size() - 1);
  std::vector<uint32_t> tri(triangles, triangles + 3 * size_t(nTriangles));
  hipStream_t st = nullptr;
  upload(tri_, tri, st);
  upload(vtxStart_, start, st);
  upload(vtxTri_, list, st);
  upload(weldA_, a, st);
  upload(weldB_, b, st);
  upload(layerStartDev_, layerStart_, st);
  acc_.allocate(sizeof(dk::F3) * std::max<uint32_t>(n, 1));
  check(hipStreamSynchronize(st), "sync");
}

void Mesh::compute(const dk::F3* pos, dk::F3* out, hipStream_t st) const {
  if (n_ == 0) return;
  const uint32_t blocks = (n_ + 255u) / 256u;
  kAccumulate<<<blocks, 256, 0, st>>>(n_, pos, tri_.as<uint32_t>(), vtxStart_.as<uint32_t>(), vtxTri_.as<uint32_t>(),
                                      acc_.as<dk::F3>());
  if (nLayers_) kWeld<<<1, 256, 0, st>>>(nLayers_, layerStartDev_.as<uint32_t>(), weldA_.as<uint32_t>(),
                                         weldB_.as<uint32_t>(), acc_.as<dk::F3>());
  kNormalize<<<blocks, 256, 0, st>>>(n_, acc_.as<dk::F3>(), out);
  check(hipGetLastError(), "normals launch");
}

}  // namespace nrm
