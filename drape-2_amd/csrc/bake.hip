// bake.hip — device SDF bake of closed triangle meshes (semantics in bake.hpp).
#include "bake.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>

#include "device_buffer.hpp"

namespace bake {

namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("hip: ") + what + ": " + hipGetErrorString(e));
}

struct V3 {
  double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 normalized(V3 a) {  // a unchanged when its norm is zero
  const double z = dot(a, a);
  if (!(z > 0.0)) return a;
  const double n = std::sqrt(z);
  return {a.x / n, a.y / n, a.z / n};
}

// Every undirected edge used by exactly two triangles, in opposite directions (triangle_mesh.hpp:16-17).
bool closedManifold(const Component& c) {
  std::vector<uint64_t> e;
  e.reserve(3 * size_t(c.nTriangles));
  for (uint32_t t = 0; t < c.nTriangles; ++t)
    for (int k = 0; k < 3; ++k) {
      const uint64_t a = c.triangles[3 * size_t(t) + k], b = c.triangles[3 * size_t(t) + (k + 1) % 3];
      e.push_back((a << 32) | b);
    }
  std::sort(e.begin(), e.end());
  if (std::adjacent_find(e.begin(), e.end()) != e.end()) return false;
  for (uint64_t k : e) {
    const uint64_t r = (k << 32) | (k >> 32);
    if (!std::binary_search(e.begin(), e.end(), r)) return false;
  }
  return !e.empty();
}

// Per triangle: corners a, b, c, then the pseudonormals of face, vertices a, b, c, edges ab, bc, ca (30 doubles).
constexpr int kTriDoubles = 30;

std::vector<double> triangleData(const Component& c) {
  auto X = [&](uint32_t v) { return V3{c.vertices[3 * size_t(v)], c.vertices[3 * size_t(v) + 1], c.vertices[3 * size_t(v) + 2]}; };
  std::vector<V3> face(c.nTriangles), vertex(c.nVertices, V3{0, 0, 0});
  std::map<std::pair<uint32_t, uint32_t>, V3> edge;
  for (uint32_t t = 0; t < c.nTriangles; ++t) {
    const uint32_t* tri = c.triangles + 3 * size_t(t);
    face[t] = normalized(cross(sub(X(tri[1]), X(tri[0])), sub(X(tri[2]), X(tri[0]))));
    for (int k = 0; k < 3; ++k) {
      const V3 p = X(tri[k]);
      const V3 e1 = normalized(sub(X(tri[(k + 1) % 3]), p));
      const V3 e2 = normalized(sub(X(tri[(k + 2) % 3]), p));
      const double angle = std::acos(std::clamp(dot(e1, e2), -1.0, 1.0));
      V3& vn = vertex[tri[k]];
      vn = {vn.x + angle * face[t].x, vn.y + angle * face[t].y, vn.z + angle * face[t].z};
      auto key = std::minmax(tri[k], tri[(k + 1) % 3]);
      V3& en = edge.try_emplace(key, V3{0, 0, 0}).first->second;
      en = {en.x + face[t].x, en.y + face[t].y, en.z + face[t].z};
    }
  }
  std::vector<double> out(kTriDoubles * size_t(c.nTriangles));
  for (uint32_t t = 0; t < c.nTriangles; ++t) {
    const uint32_t* tri = c.triangles + 3 * size_t(t);
    const V3 vals[10] = {X(tri[0]), X(tri[1]), X(tri[2]), face[t], vertex[tri[0]], vertex[tri[1]], vertex[tri[2]],
                         edge.at(std::minmax(tri[0], tri[1])), edge.at(std::minmax(tri[1], tri[2])),
                         edge.at(std::minmax(tri[2], tri[0]))};
    for (int k = 0; k < 10; ++k) {
      out[kTriDoubles * size_t(t) + 3 * k] = vals[k].x;
      out[kTriDoubles * size_t(t) + 3 * k + 1] = vals[k].y;
      out[kTriDoubles * size_t(t) + 3 * k + 2] = vals[k].z;
    }
  }
  return out;
}

}  // namespace

// ---- device ----

struct GridView {
  double ox, oy, oz, cell, band;
  int32_t nx, ny, nz, pad;
};

struct D3 {
  double x, y, z;
};
__device__ inline D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline double ddot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ inline D3 dscale(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline D3 dadd(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3 ld3(const double* p) { return {p[0], p[1], p[2]}; }

// Closest point on triangle abc to p and the feature holding it: 0 face, 1-3 vertices a, b, c, 4-6 edges ab, bc,
// ca (collision_bake.cpp:19-55, Ericson RTCD 5.1.5); the same operations in the same order.
__device__ inline D3 closestPoint(D3 p, D3 a, D3 b, D3 c, int& feature) {
#pragma clang fp contract(off)
  const D3 ab = dsub(b, a), ac = dsub(c, a), ap = dsub(p, a);
  const double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
  if (d1 <= 0 && d2 <= 0) {
    feature = 1;
    return a;
  }
  const D3 bp = dsub(p, b);
  const double d3 = ddot(ab, bp), d4 = ddot(ac, bp);
  if (d3 >= 0 && d4 <= d3) {
    feature = 2;
    return b;
  }
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    feature = 4;
    return dadd(a, dscale(d1 / (d1 - d3), ab));
  }
  const D3 cp = dsub(p, c);
  const double d5 = ddot(ab, cp), d6 = ddot(ac, cp);
  if (d6 >= 0 && d5 <= d6) {
    feature = 3;
    return c;
  }
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    feature = 6;
    return dadd(a, dscale(d2 / (d2 - d6), ac));
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    feature = 5;
    return dadd(b, dscale((d4 - d3) / ((d4 - d3) + (d5 - d6)), dsub(c, b)));
  }
  const double denom = 1.0 / (va + vb + vc);
  feature = 0;
  return dadd(dadd(a, dscale(vb * denom, ab)), dscale(vc * denom, ac));
}

// The samples a triangle visits: its bounding box grown by the band, clipped to the grid (cellRange,
// collision_bake.cpp:124-129).
struct Range {
  int32_t i0[3], i1[3];
  D3 tmin, tmax;
};
__device__ inline Range triangleRange(const GridView& g, D3 a, D3 b, D3 c) {
#pragma clang fp contract(off)
  Range r;
  r.tmin = {fmin(fmin(a.x, b.x), c.x), fmin(fmin(a.y, b.y), c.y), fmin(fmin(a.z, b.z), c.z)};
  r.tmax = {fmax(fmax(a.x, b.x), c.x), fmax(fmax(a.y, b.y), c.y), fmax(fmax(a.z, b.z), c.z)};
  const double lo[3] = {r.tmin.x, r.tmin.y, r.tmin.z}, hi[3] = {r.tmax.x, r.tmax.y, r.tmax.z};
  const double o[3] = {g.ox, g.oy, g.oz};
  const int32_t dims[3] = {g.nx, g.ny, g.nz};
  for (int a3 = 0; a3 < 3; ++a3) {
    r.i0[a3] = int32_t(fmax(0.0, ceil(((lo[a3] - g.band) - o[a3]) / g.cell)));
    r.i1[a3] = int32_t(fmin(double(dims[a3] - 1), floor(((hi[a3] + g.band) - o[a3]) / g.cell)));
  }
  return r;
}

__device__ inline D3 samplePos(const GridView& g, int32_t x, int32_t y, int32_t z) {
#pragma clang fp contract(off)
  return {g.ox + g.cell * double(x), g.oy + g.cell * double(y), g.oz + g.cell * double(z)};
}

// Pass 0: per-sample minimum squared distance (as bits). Pass 1: smallest triangle index reaching it.
template <int Pass>
__global__ void __launch_bounds__(256) kTriangleScan(GridView g, const double* __restrict__ tris,
                                                     unsigned long long* dist2, uint32_t* arg) {
#pragma clang fp contract(off)
  const uint32_t t = blockIdx.x;
  const double* T = tris + size_t(kTriDoubles) * t;
  const D3 a = ld3(T), b = ld3(T + 3), c = ld3(T + 6);
  const Range r = triangleRange(g, a, b, c);
  if (r.i1[0] < r.i0[0] || r.i1[1] < r.i0[1] || r.i1[2] < r.i0[2]) return;
  const uint32_t nxr = uint32_t(r.i1[0] - r.i0[0] + 1), nyr = uint32_t(r.i1[1] - r.i0[1] + 1);
  const uint32_t total = nxr * nyr * uint32_t(r.i1[2] - r.i0[2] + 1);
  for (uint32_t k = blockIdx.y * 256u + threadIdx.x; k < total; k += 256u * gridDim.y) {
    const int32_t x = r.i0[0] + int32_t(k % nxr), y = r.i0[1] + int32_t((k / nxr) % nyr), z = r.i0[2] + int32_t(k / (nxr * nyr));
    const size_t i = (size_t(z) * size_t(g.ny) + size_t(y)) * size_t(g.nx) + size_t(x);
    const D3 p = samplePos(g, x, y, z);
    const D3 gap = {fmax(fmax(r.tmin.x - p.x, p.x - r.tmax.x), 0.0), fmax(fmax(r.tmin.y - p.y, p.y - r.tmax.y), 0.0),
                    fmax(fmax(r.tmin.z - p.z, p.z - r.tmax.z), 0.0)};
    const double best = __longlong_as_double((long long)dist2[i]);
    const double gap2 = ddot(gap, gap);
    if (Pass == 0 ? gap2 >= best : gap2 > best) continue;  // cannot reach (pass 0) or tie (pass 1) the minimum
    int f;
    const D3 q = closestPoint(p, a, b, c, f);
    const D3 pq = dsub(p, q);
    const double d2 = ddot(pq, pq);
    if (Pass == 0) {
      if (d2 < best) atomicMin(dist2 + i, (unsigned long long)__double_as_longlong(d2));
    } else if (d2 == best) {
      atomicMin(arg + i, t);
    }
  }
}

// Sign of each sample with a nearest triangle, from the pseudonormal of the feature holding the closest point;
// samples beyond the band are "free" for the flood fill, and the grid boundary's free samples seed it.
__global__ void __launch_bounds__(256) kSignAndSeed(GridView g, const double* __restrict__ tris,
                                                    const unsigned long long* __restrict__ dist2,
                                                    const uint32_t* __restrict__ arg, int8_t* sign, uint8_t* state) {
#pragma clang fp contract(off)
  const size_t i = size_t(blockIdx.x) * 256u + threadIdx.x;
  const size_t count = size_t(g.nx) * size_t(g.ny) * size_t(g.nz);
  if (i >= count) return;
  const int32_t x = int32_t(i % size_t(g.nx)), y = int32_t((i / size_t(g.nx)) % size_t(g.ny)),
                z = int32_t(i / (size_t(g.nx) * size_t(g.ny)));
  const double d2 = __longlong_as_double((long long)dist2[i]);
  int8_t s = 0;
  if (arg[i] != UINT32_MAX) {
    const double* T = tris + size_t(kTriDoubles) * arg[i];
    int f;
    const D3 p = samplePos(g, x, y, z);
    const D3 q = closestPoint(p, ld3(T), ld3(T + 3), ld3(T + 6), f);
    const int slot = f == 0 ? 3 : f + 3;  // face 3; vertices a, b, c 4-6; edges ab, bc, ca 7-9
    s = ddot(dsub(p, q), ld3(T + 3 * slot)) >= 0 ? int8_t(1) : int8_t(-1);
  }
  sign[i] = s;
  const bool free = d2 > g.band * g.band;
  const bool boundary = x == 0 || y == 0 || z == 0 || x == g.nx - 1 || y == g.ny - 1 || z == g.nz - 1;
  state[i] = free ? uint8_t(boundary ? 2 : 1) : uint8_t(0);  // 0 within the band, 1 free, 2 outside
}

// One sweep of the flood fill along an axis: a free sample next (along the sweep) to an outside one becomes outside.
__global__ void __launch_bounds__(256) kSweep(GridView g, int axis, int forward, uint8_t* state, uint32_t* changed) {
  const int32_t n[3] = {g.nx, g.ny, g.nz};
  const int32_t u = axis == 0 ? 1 : 0, w = axis == 2 ? 1 : 2;  // the two other axes
  const uint32_t line = blockIdx.x * 256u + threadIdx.x;
  if (line >= uint32_t(n[u]) * uint32_t(n[w])) return;
  const int32_t cu = int32_t(line % uint32_t(n[u])), cw = int32_t(line / uint32_t(n[u]));
  const size_t stride[3] = {1, size_t(g.nx), size_t(g.nx) * size_t(g.ny)};
  const size_t base = size_t(cu) * stride[u] + size_t(cw) * stride[w];
  const int32_t len = n[axis];
  bool carry = false, any = false;
  for (int32_t k = 0; k < len; ++k) {
    const size_t i = base + size_t(forward ? k : len - 1 - k) * stride[axis];
    const uint8_t s = state[i];
    if (s == 2) {
      carry = true;
    } else if (s == 1) {
      if (carry) {
        state[i] = 2;
        any = true;
      }
    } else {
      carry = false;
    }
  }
  if (any) *changed = 1u;
}

// value = clamp(sign * distance) within the band, +-band beyond it; the field keeps the minimum over components.
__global__ void __launch_bounds__(256) kFinalize(GridView g, const unsigned long long* __restrict__ dist2,
                                                 const int8_t* __restrict__ sign, const uint8_t* __restrict__ state,
                                                 float* field) {
#pragma clang fp contract(off)
  const size_t i = size_t(blockIdx.x) * 256u + threadIdx.x;
  const size_t count = size_t(g.nx) * size_t(g.ny) * size_t(g.nz);
  if (i >= count) return;
  const double d2 = __longlong_as_double((long long)dist2[i]);
  double v;
  if (d2 <= g.band * g.band)
    v = double(sign[i]) * sqrt(d2);
  else
    v = state[i] == 2 ? g.band : -g.band;
  const float c = float(fmin(fmax(v, -g.band), g.band));
  field[i] = fminf(field[i], c);
}

__global__ void kFillField(float* field, size_t count, float v, unsigned long long* dist2, uint32_t* arg) {
  const size_t i = size_t(blockIdx.x) * 256u + threadIdx.x;
  if (i >= count) return;
  field[i] = v;
  dist2[i] = 0x7FF0000000000000ull;  // +infinity
  arg[i] = UINT32_MAX;
}

__global__ void kResetScan(size_t count, unsigned long long* dist2, uint32_t* arg) {
  const size_t i = size_t(blockIdx.x) * 256u + threadIdx.x;
  if (i >= count) return;
  dist2[i] = 0x7FF0000000000000ull;
  arg[i] = UINT32_MAX;
}

// ---- host ----

Grid gridOf(const Component* components, uint32_t n, const Options& options) {
  if (n == 0) throw std::invalid_argument("bake: no components");
  if (!(options.cellSize > 0.0) || !(options.band > 0.0) || !(options.padding >= 0.0))
    throw std::invalid_argument("bake: cell size and band must be positive, padding not negative");
  double lo[3] = {std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
                  std::numeric_limits<double>::infinity()};
  double hi[3] = {-lo[0], -lo[1], -lo[2]};
  for (uint32_t c = 0; c < n; ++c) {
    if (!closedManifold(components[c]))
      throw std::invalid_argument("collision component " + std::to_string(c) + " is not a closed manifold");
    for (uint32_t v = 0; v < components[c].nVertices; ++v)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], components[c].vertices[3 * size_t(v) + a]);
        hi[a] = std::max(hi[a], components[c].vertices[3 * size_t(v) + a]);
      }
  }
  Grid g{};
  for (int a = 0; a < 3; ++a) {
    g.origin[a] = lo[a] - options.padding;
    const double extent = (hi[a] - lo[a]) + 2 * options.padding;
    g.dims[a] = static_cast<int32_t>(std::ceil(extent / options.cellSize)) + 1;
  }
  return g;
}

void bakeField(const Component* components, uint32_t n, const Options& options, float* values, hipStream_t st) {
  const Grid grid = gridOf(components, n, options);
  const size_t count = grid.count();
  GridView g{grid.origin[0], grid.origin[1], grid.origin[2], options.cellSize, options.band,
             grid.dims[0], grid.dims[1], grid.dims[2], 0};
  dh::DeviceBuffer field, dist2, arg, sign, state, changed, tris;
  field.allocate(sizeof(float) * count);
  dist2.allocate(sizeof(unsigned long long) * count);
  arg.allocate(sizeof(uint32_t) * count);
  sign.allocate(count);
  state.allocate(count);
  changed.allocate(sizeof(uint32_t));
  const uint32_t cellBlocks = uint32_t((count + 255) / 256);
  kFillField<<<cellBlocks, 256, 0, st>>>(field.as<float>(), count, float(options.band), dist2.as<unsigned long long>(),
                                         arg.as<uint32_t>());
  for (uint32_t c = 0; c < n; ++c) {
    const std::vector<double> td = triangleData(components[c]);
    check(hipStreamSynchronize(st), "sync");  // the previous component's kernels are done with `tris`
    tris.allocate(sizeof(double) * std::max<size_t>(td.size(), 1));
    check(hipMemcpyAsync(tris.as<void>(), td.data(), sizeof(double) * td.size(), hipMemcpyHostToDevice, st), "copy");
    if (c > 0) kResetScan<<<cellBlocks, 256, 0, st>>>(count, dist2.as<unsigned long long>(), arg.as<uint32_t>());
    const uint32_t nt = components[c].nTriangles;
    // Large triangles visit many samples: split each triangle's samples over up to 128 workgroups.
    double maxCells = 1.0;
    for (uint32_t t = 0; t < nt; ++t) {
      double cells = 1.0;
      for (int a = 0; a < 3; ++a) {
        double lo = td[kTriDoubles * size_t(t) + a], hi = lo;
        for (int k = 1; k < 3; ++k) {
          lo = std::min(lo, td[kTriDoubles * size_t(t) + 3 * k + a]);
          hi = std::max(hi, td[kTriDoubles * size_t(t) + 3 * k + a]);
        }
        cells *= (hi - lo + 2 * options.band) / options.cellSize + 2.0;
      }
      maxCells = std::max(maxCells, cells);
    }
    const dim3 scanGrid(nt, uint32_t(std::clamp(maxCells / 8192.0, 1.0, 128.0)));
    kTriangleScan<0><<<scanGrid, 256, 0, st>>>(g, tris.as<double>(), dist2.as<unsigned long long>(), arg.as<uint32_t>());
    kTriangleScan<1><<<scanGrid, 256, 0, st>>>(g, tris.as<double>(), dist2.as<unsigned long long>(), arg.as<uint32_t>());
    kSignAndSeed<<<cellBlocks, 256, 0, st>>>(g, tris.as<double>(), dist2.as<unsigned long long>(), arg.as<uint32_t>(),
                                             sign.as<int8_t>(), state.as<uint8_t>());
    // Sweeps until a round changes nothing; a round that changes something marks at least one more sample, so at
    // most `count` rounds.
    for (size_t round = 0; round <= count; ++round) {
      check(hipMemsetAsync(changed.as<void>(), 0, sizeof(uint32_t), st), "memset");
      for (int axis = 0; axis < 3; ++axis) {
        const int32_t u = axis == 0 ? grid.dims[1] : grid.dims[0], w = axis == 2 ? grid.dims[1] : grid.dims[2];
        const uint32_t lines = uint32_t(u) * uint32_t(w);
        for (int forward = 1; forward >= 0; --forward)
          kSweep<<<(lines + 255) / 256, 256, 0, st>>>(g, axis, forward, state.as<uint8_t>(), changed.as<uint32_t>());
      }
      uint32_t hostChanged = 0;
      check(hipMemcpyAsync(&hostChanged, changed.as<void>(), sizeof(uint32_t), hipMemcpyDeviceToHost, st), "copy");
      check(hipStreamSynchronize(st), "sync");
      if (!hostChanged) break;
    }
    kFinalize<<<cellBlocks, 256, 0, st>>>(g, dist2.as<unsigned long long>(), sign.as<int8_t>(), state.as<uint8_t>(),
                                          field.as<float>());
  }
  check(hipGetLastError(), "bake launch");
  check(hipMemcpyAsync(values, field.as<void>(), sizeof(float) * count, hipMemcpyDeviceToHost, st), "copy");
  check(hipStreamSynchronize(st), "sync");
}

}  // namespace bake
