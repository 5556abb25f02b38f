// normals.hip — device vertex normals with seam welding (semantics in normals.hpp).
#include "normals.hpp"

#include <algorithm>
#include <mutex>
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
    : n_(n) {
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

// ---------------------------------------------------------------------------------------------------------------
// Allocation-free form for callers that cannot allocate (code under graph capture, serving loops): the caller's
// device arrays in and out, and every working array — the incident-triangle lists and the per-vertex sums — in
// one fixed device array of this file, `scratch`, held by each GPU until the process ends. Its address on a
// device comes from hipGetSymbolAddress (the array's name in host code is only a stand-in for it), looked up once
// per device with that device current.

namespace {

// Layout of `scratch` in 32-bit words: the error flag, per-vertex counts (n + 1), list starts (n + 1), the
// incident triangle of each corner (3 T), the per-vertex sums (3 n floats).
constexpr size_t kFlagWords = 64;
constexpr size_t kScratchWords =
    kFlagWords + 2 * (size_t(kNoAllocMaxVertices) + 1) + 3 * size_t(kNoAllocMaxTriangles) + 3 * size_t(kNoAllocMaxVertices);

}  // namespace

static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex gScratchLock;
uint32_t* gScratch[kMaxDevices] = {};

// This device's copy of `scratch` (looked up on the first call on the device).
uint32_t* deviceScratch() {
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("hip: device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(gScratchLock);
  if (!gScratch[dev]) {
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    gScratch[dev] = static_cast<uint32_t*>(p);
  }
  return gScratch[dev];
}

}  // namespace

__global__ void __launch_bounds__(256) kNaClear(uint32_t* words, uint32_t count) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < count) words[i] = 0u;
}

// Counts each vertex's corners; a corner index >= n sets the error flag and is left out.
__global__ void __launch_bounds__(256) kNaCount(const uint32_t* __restrict__ tri, uint32_t corners, uint32_t n,
                                                uint32_t* counts, uint32_t* flag) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= corners) return;
  const uint32_t v = tri[i];
  if (v >= n) {
    atomicOr(flag, 1u);
    return;
  }
  atomicAdd(counts + v, 1u);
}

// Exclusive scan of counts[0, n) into starts[0, n] (one workgroup).
__global__ void __launch_bounds__(1024) kNaScan(const uint32_t* counts, uint32_t* starts, uint32_t n) {
  __shared__ uint32_t buf[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < n; base += 1024u) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t own = i < n ? counts[i] : 0u;
    buf[threadIdx.x] = own;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
      const uint32_t add = threadIdx.x >= off ? buf[threadIdx.x - off] : 0u;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < n) starts[i] = carry + buf[threadIdx.x] - own;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) starts[n] = carry;
}

// Lists each corner's triangle under its vertex (slot order arbitrary; sorted per vertex below).
__global__ void __launch_bounds__(256) kNaFill(const uint32_t* __restrict__ tri, uint32_t corners, uint32_t n,
                                               const uint32_t* __restrict__ starts, uint32_t* cursor,
                                               uint32_t* list) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= corners) return;
  const uint32_t v = tri[i];
  if (v >= n) return;
  list[starts[v] + atomicAdd(cursor + v, 1u)] = i / 3u;
}

// Sorts a vertex's triangles into increasing order, then sums their cross products in that order — the order,
// and the arithmetic, of kAccumulate over the host-built lists.
__global__ void __launch_bounds__(256) kNaSortAccumulate(uint32_t n, const dk::F3* __restrict__ pos,
                                                         const uint32_t* __restrict__ tri,
                                                         const uint32_t* __restrict__ starts, uint32_t* list,
                                                         dk::F3* acc) {
#pragma clang fp contract(off)
  const uint32_t v = blockIdx.x * 256u + threadIdx.x;
  if (v >= n) return;
  const uint32_t b = starts[v], e = starts[v + 1];
  for (uint32_t i = b + 1; i < e; ++i) {
    const uint32_t key = list[i];
    uint32_t j = i;
    while (j > b && list[j - 1] > key) {
      list[j] = list[j - 1];
      --j;
    }
    list[j] = key;
  }
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (uint32_t k = b; k < e; ++k) {
    const uint32_t* t = tri + 3u * size_t(list[k]);
    // Every corner of a listed triangle was checked by kNaCount: a triangle with a bad corner still lists its
    // good corners, so its corners are checked again here and a bad one contributes nothing.
    if (t[0] >= n || t[1] >= n || t[2] >= n) continue;
    const dk::F3 c = triangleCross(pos[t[0]], pos[t[1]], pos[t[2]]);
    sx = sx + c.x;
    sy = sy + c.y;
    sz = sz + c.z;
  }
  acc[v] = dk::F3{sx, sy, sz};
}

// The weld pairs in order, one lane (the reference's sequential pass, normals.cpp:13-17); a pair with an index
// >= n sets the flag and is skipped.
__global__ void __launch_bounds__(64) kNaWeld(const uint32_t* __restrict__ welds, uint32_t nWelds, uint32_t n,
                                              dk::F3* acc, uint32_t* flag) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  for (uint32_t i = 0; i < nWelds; ++i) {
    const uint32_t a = welds[2 * size_t(i)], b = welds[2 * size_t(i) + 1];
    if (a >= n || b >= n) {
      atomicOr(flag, 1u);
      continue;
    }
    const dk::F3 na = acc[a], nb = acc[b];
    const dk::F3 s{na.x + nb.x, na.y + nb.y, na.z + nb.z};
    acc[a] = s;
    acc[b] = s;
  }
}

static void normalsNoAlloc(const dk::F3* pos, uint32_t n, const uint32_t* tri, uint32_t nTri, const uint32_t* welds,
                    uint32_t nWelds, dk::F3* out, hipStream_t st) {
  if (n > kNoAllocMaxVertices || nTri > kNoAllocMaxTriangles)
    throw std::invalid_argument("normals without allocation: at most " + std::to_string(kNoAllocMaxVertices) +
                                " vertices and " + std::to_string(kNoAllocMaxTriangles) + " triangles");
  if (n == 0) return;
  uint32_t* base = deviceScratch();
  uint32_t* flag = base;
  uint32_t* counts = base + kFlagWords;
  uint32_t* starts = counts + (n + 1);
  uint32_t* list = starts + (n + 1);
  dk::F3* acc = reinterpret_cast<dk::F3*>(list + 3 * size_t(nTri));
  const uint32_t corners = 3u * nTri;
  // Previous calls leave their contents: the flag and the counts start from zero (a fill kernel, not a memset).
  const uint32_t clear = uint32_t(kFlagWords) + n + 1;
  kNaClear<<<(clear + 255u) / 256u, 256, 0, st>>>(base, clear);
  if (corners) kNaCount<<<(corners + 255u) / 256u, 256, 0, st>>>(tri, corners, n, counts, flag);
  kNaScan<<<1, 1024, 0, st>>>(counts, starts, n);
  kNaClear<<<(n + 255u) / 256u, 256, 0, st>>>(counts, n);  // reused as the fill cursors
  if (corners) kNaFill<<<(corners + 255u) / 256u, 256, 0, st>>>(tri, corners, n, starts, counts, list);
  kNaSortAccumulate<<<(n + 255u) / 256u, 256, 0, st>>>(n, pos, tri, starts, list, acc);
  if (nWelds) kNaWeld<<<1, 64, 0, st>>>(welds, nWelds, n, acc, flag);
  kNormalize<<<(n + 255u) / 256u, 256, 0, st>>>(n, acc, out);
  check(hipGetLastError(), "normals launch");
  uint32_t bad = 0;
  check(hipMemcpyAsync(&bad, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, st), "copy");
  check(hipStreamSynchronize(st), "normals");
  if (bad) throw std::out_of_range("normals: triangle or weld vertex out of range");
}

void vertexNormalsNoAlloc(const float* positions, uint32_t n, const uint32_t* triangles, uint32_t nTriangles,
                          const uint32_t* welds, uint32_t nWelds, float* out, hipStream_t st) {
  normalsNoAlloc(reinterpret_cast<const dk::F3*>(positions), n, triangles, nTriangles, welds, nWelds,
                 reinterpret_cast<dk::F3*>(out), st);
}

}  // namespace nrm
