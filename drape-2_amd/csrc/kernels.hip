// kernels.hip — the HIP kernels of the implicit-Euler Newton–PCG cloth step on MI355X (gfx950).
//
// Execution model: one lane per particle row. Row kernels run 256-thread workgroups over 256-row tiles; the SELL-64
// matrix kernels give each wavefront one 64-row slice. Workgroup ids are remapped so that each of the eight XCDs
// (which receive workgroups round-robin) walks a contiguous range of tiles, keeping neighbouring rows — and the
// vector entries their block columns gather — in one XCD's L2.
//
// Determinism: no floating-point atomics. Every sum runs in a fixed order: a lane accumulates its row's terms in
// the reference's gather order (gpu_solver_core.cpp:184-224), FP64 partial sums per tile reduce in a fixed
// shuffle tree, and every workgroup that needs a segment total re-sums the same partials in the same order.
// Integer atomics appear only in the structure builder's counting pass, whose results are order independent (the
// lists they fill are sorted afterwards).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "field.hpp"
#include "terms.hpp"

namespace dk {

namespace {

void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("hip: ") + what + ": " + hipGetErrorString(e));
}

inline uint32_t xcdGrid(uint32_t count) { return 8u * ((count + 7u) >> 3); }

}  // namespace

// Workgroup -> tile such that XCD (bid % 8) covers the contiguous range [x * per, (x + 1) * per).
__device__ inline uint32_t xcdMap(uint32_t bid, uint32_t count) {
  const uint32_t per = (count + 7u) >> 3;
  return (bid & 7u) * per + (bid >> 3);
}

// Load of a workgroup-uniform value through the scalar cache (constant address space). Only for data no
// workgroup of the running kernel writes: the scalar cache is invalidated at every kernel start, so values
// written by earlier kernels are seen.
template <class T>
__device__ inline T ldUniform(const T* p) {
  return *reinterpret_cast<const __attribute__((address_space(4))) T*>(reinterpret_cast<uintptr_t>(p));
}

// Sum over a 256-thread workgroup, fixed order; the result is returned to every thread.
__device__ inline double blockSum256(double v, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) lds[4] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
  __syncthreads();
  return lds[4];
}

__device__ inline float blockMax256(float v, float* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_down(v, off, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) lds[4] = fmaxf(fmaxf(lds[0], lds[1]), fmaxf(lds[2], lds[3]));
  __syncthreads();
  return lds[4];
}

// Sum of part[begin, end) in a fixed order, identical in every workgroup that calls it.
__device__ inline double rangeSum256(const double* part, uint32_t begin, uint32_t end, double* lds) {
  double s = 0.0;
  for (uint32_t i = begin + threadIdx.x; i < end; i += 256u) s += part[i];
  return blockSum256(s, lds);
}

__device__ inline float rangeMax256(const float* part, uint32_t begin, uint32_t end, float* lds) {
  float s = 0.0f;
  for (uint32_t i = begin + threadIdx.x; i < end; i += 256u) s = fmaxf(s, part[i]);
  return blockMax256(s, lds);
}

__device__ inline double waveSum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;  // valid in lane 0
}

// Three sums over a 256-thread workgroup in one pass (fixed order), returned to every thread. lds >= 16 doubles.
__device__ inline void blockSum3x256(double& a, double& b, double& c, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
    c += __shfl_down(c, off, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
    lds[w] = a;
    lds[4 + w] = b;
    lds[8 + w] = c;
  }
  __syncthreads();
  a = ((lds[0] + lds[1]) + lds[2]) + lds[3];
  b = ((lds[4] + lds[5]) + lds[6]) + lds[7];
  c = ((lds[8] + lds[9]) + lds[10]) + lds[11];
}

// Column of slot t of a lane's SELL row (col points at the lane's slot 0), the dummy column past K. The load is
// unconditional — the column buffers carry kSellColSlack entries past their ends — and the select acts on the loaded
// value: a load of a selected address (this entry or the dummy index in the kernel arguments) would be a generic
// (flat) load, which retires out of order with global loads and makes every later wait a wait for all of them.
__device__ inline uint32_t sellColumn(const uint32_t* __restrict__ col, uint32_t t, uint32_t K, uint32_t dummy) {
  const uint32_t c = col[t * kSlice];
  return t < K ? c : dummy;
}

// Block times vector, and the transposed block times vector (a lower slot's block is stored as its transpose).
__device__ inline void blockMulAdd(const float* b, F3 x, float& a0, float& a1, float& a2) {
  a0 += b[0] * x.x + b[1] * x.y + b[2] * x.z;
  a1 += b[3] * x.x + b[4] * x.y + b[5] * x.z;
  a2 += b[6] * x.x + b[7] * x.y + b[8] * x.z;
}
__device__ inline void blockTMulAdd(const float* b, F3 x, float& a0, float& a1, float& a2) {
  a0 += b[0] * x.x + b[3] * x.y + b[6] * x.z;
  a1 += b[1] * x.x + b[4] * x.y + b[7] * x.z;
  a2 += b[2] * x.x + b[5] * x.y + b[8] * x.z;
}

// Row product of a slice of the symmetric SELL-64 matrix with a gathered vector: the lower slots (transposed
// blocks found through lowBlk) then the upper slots. `gather(c)` returns x_c (the dummy column nRows gives zero).
// Slots are taken eight at a time: the eight index loads, then the eight gathers, then the block products, so a
// row costs a few memory round trips instead of one per block.
template <class Gather>
__device__ inline F3 sellRowProduct(const View& v, uint32_t slice, uint32_t lane, Gather gather) {
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  if (v.lowK) {
    const uint32_t KL = v.lowK[slice];
    const uint32_t lb = v.lowBase[slice] + lane;
    for (uint32_t k0 = 0; k0 < KL; k0 += 8u) {
      uint32_t c[8], blk[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        c[t] = sellColumn(v.lowCol + lb, k0 + t, KL, v.nRows);
        blk[t] = sellColumn(v.lowBlk + lb, k0 + t, KL, v.zeroBlock);
      }
      F3 x[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) x[t] = gather(c[t]);
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (k0 + t < KL) blockTMulAdd(v.sellVal + size_t(blk[t]) * 9u, x[t], a0, a1, a2);
    }
  }
  const uint32_t K = v.sellK[slice];
  const uint32_t base = v.sellBase[slice];
  const float* __restrict__ val = v.sellVal + base + 9u * lane;
  const uint32_t* __restrict__ col = v.sellCol + base / 9u + lane;
  for (uint32_t k0 = 0; k0 < K; k0 += 8u) {
    uint32_t c[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) c[t] = sellColumn(col, k0 + t, K, v.nRows);
    F3 x[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) x[t] = gather(c[t]);
#pragma unroll
    for (int t = 0; t < 8; ++t)
      if (k0 + t < K) blockMulAdd(val + (k0 + t) * 9u * kSlice, x[t], a0, a1, a2);
  }
  return f3(a0, a1, a2);
}

// The same row product accumulated in FP64 (each FP32 block entry times an FP32 vector entry is exact in FP64, so
// the sum carries no FP32 rounding): the true residual of the refinement and of the solve report.
struct D3 {
  double x, y, z;
};
// With Sq, *sq also accumulates the square of every product term (b x)^2 of the row's three sums.
template <bool Sq = false, class Gather>
__device__ inline D3 sellRowProduct64(const View& v, uint32_t slice, uint32_t lane, Gather gather,
                                      double* sq = nullptr) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  auto squares = [&](const float* b, F3 x, int i0, int i1, int i2) {
    const double p0 = double(b[i0]) * x.x, p1 = double(b[i1]) * x.y, p2 = double(b[i2]) * x.z;
    return p0 * p0 + p1 * p1 + p2 * p2;
  };
  if (v.lowK) {
    const uint32_t KL = v.lowK[slice];
    const uint32_t lb = v.lowBase[slice] + lane;
    for (uint32_t k = 0; k < KL; ++k) {
      const uint32_t c = v.lowCol[lb + k * kSlice];
      const float* b = v.sellVal + size_t(v.lowBlk[lb + k * kSlice]) * 9u;  // transposed
      const F3 x = gather(c);
      a0 += double(b[0]) * x.x + double(b[3]) * x.y + double(b[6]) * x.z;
      a1 += double(b[1]) * x.x + double(b[4]) * x.y + double(b[7]) * x.z;
      a2 += double(b[2]) * x.x + double(b[5]) * x.y + double(b[8]) * x.z;
      if constexpr (Sq) *sq += squares(b, x, 0, 3, 6) + squares(b, x, 1, 4, 7) + squares(b, x, 2, 5, 8);
    }
  }
  const uint32_t K = v.sellK[slice];
  const uint32_t base = v.sellBase[slice];
  const float* __restrict__ val = v.sellVal + base + 9u * lane;
  const uint32_t* __restrict__ col = v.sellCol + base / 9u + lane;
  for (uint32_t k = 0; k < K; ++k) {
    const float* b = val + k * 9u * kSlice;
    const F3 x = gather(col[k * kSlice]);
    a0 += double(b[0]) * x.x + double(b[1]) * x.y + double(b[2]) * x.z;
    a1 += double(b[3]) * x.x + double(b[4]) * x.y + double(b[5]) * x.z;
    a2 += double(b[6]) * x.x + double(b[7]) * x.y + double(b[8]) * x.z;
    if constexpr (Sq) *sq += squares(b, x, 0, 1, 2) + squares(b, x, 3, 4, 5) + squares(b, x, 6, 7, 8);
  }
  return D3{a0, a1, a2};
}

// ---------------------------------------------------------------------------------------------------------------
// Contact planes and the per-row substep kernels.

// Contact radius of a particle: the segment's collision offset plus half its fabric's thickness (sim_model.hpp:71).
__device__ inline float fabricRadius(const SegParams& P, const Fabric32& f) {
  return P.collisionOffset + 0.5f * f.thickness;
}

// A held particle's vector with its component along the contact normal removed (the constraint filter of the
// solve, cpu_solver.cpp:69-73). The plane is stored as (normal, offset) in one float4.
__device__ inline F3 onPlane(uint32_t cflags, float4 plane, F3 w) {
  if (!(cflags & kContactActive)) return w;
  const F3 n = f3(plane.x, plane.y, plane.z);
  return w - dot(n, w) * n;
}

// filter(M w) with the block-Jacobi inverse M stored as its upper triangle (00 01 02 11 12 22)
// (cpu_solver.cpp:179-181).
__device__ inline F3 applyPrecond(const float* m, uint32_t cflags, float4 plane, F3 w) {
  const F3 y = f3(m[0] * w.x + m[1] * w.y + m[2] * w.z, m[1] * w.x + m[3] * w.y + m[4] * w.z,
                  m[2] * w.x + m[4] * w.y + m[5] * w.z);
  return onPlane(cflags, plane, y);
}

// Substep start (cpu_solver.cpp:234-263). Per row: the start position; the inertial target relative to it,
// y = h v + h^2 (g + f/m), with gravity only while the garment settles; the Newton guess, the reference's
// previous-velocity guess u = h v (cpu_solver.cpp:244), or with a learned warm start u = y + h d, the inertial
// target corrected by the network's velocity residual d (what the elastic and contact forces add to the free
// flight; near zero in free flight, so a network's error scales with a small quantity); and, for a particle
// within reach of the body, its contact plane: the field's tangent plane at the start position, offset by the
// contact radius, with friction weighted by the normal force of the previous solve. Pinned and padding rows have
// no target, no guess and no contact.
__global__ void __launch_bounds__(256) kSubstepBegin(View v, uint32_t learned) {
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  const SegParams& P = v.seg[s];
  if (tile == P.tileBegin && threadIdx.x == 0) {
    SegState* S = v.state + s;
    S->newtonDone = 0u;
    S->pcgConverged = 0u;
    S->lsDone = 0u;
    S->cntSubsteps += 1u;
  }
  const uint32_t row = tile * kTile + threadIdx.x;
  const F3 here = v.x[row];
  const float priorForce = v.cforce[row];
  v.xStart[row] = here;
  v.cforce[row] = 0.0f;
  F3 target = f3(0, 0, 0), guess = f3(0, 0, 0);
  uint32_t contact = 0u;
  float frictionWeight = 0.0f;
  if (!(v.rowFlags[row] & (kRowPinned | kRowPad))) {
    const float h = P.h;
    const F3 vel = v.v[row];
    const F3 g = v.garments[v.rowGarment[row]].phase == 1u ? f3(P.gx, P.gy, P.gz) : f3(0, 0, 0);
    const F3 f = v.fExt[row];
    const float invMass = 1.0f / v.mass[row];
    target = h * vel + (h * h) * (g + invMass * f);
    guess = learned ? target + h * v.warm[row] : h * vel;
    if (P.field.present) {
      const float radius = fabricRadius(P, v.fabrics[v.rowFabric[row]]);
      const float dist = bodyDistance(v.fieldValues, P.field, here);
      const F3 n = dist < radius + P.contactReach ? bodyNormal(v.fieldValues, P.field, here) : f3(0, 0, 0);
      if (!allZero(n)) {
        contact = kContactNear;
        v.cnormal[row] = make_float4(n.x, n.y, n.z, radius - dist);
        frictionWeight = priorForce;
      }
    }
  }
  v.yRel[row] = target;
  v.u[row] = guess;
  v.cflags[row] = contact;
  v.clambda[row] = frictionWeight;
}

// Newton iteration start (cpu_solver.cpp:217-223): a particle near the body is held on its contact plane once its
// guess is within the snap tolerance of it, and stays held until the assembly releases it; a held guess below the
// plane is moved onto it.
__global__ void __launch_bounds__(256) kHoldContacts(View v) {
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  if (v.state[s].newtonDone) return;
  const uint32_t row = tile * kTile + threadIdx.x;
  const uint32_t cf = v.cflags[row];
  if (!(cf & kContactNear)) return;
  const float4 plane = v.cnormal[row];
  const F3 n = f3(plane.x, plane.y, plane.z);
  const F3 guess = v.u[row];
  const float depth = plane.w - dot(n, guess);  // how far the guess lies below its plane
  const bool held = (cf & kContactActive) || depth > -v.seg[s].contactSnap;
  if (held && depth > 0.0f) v.u[row] = guess + depth * n;
  if (held) v.cflags[row] = cf | kContactActive;
}

// ---------------------------------------------------------------------------------------------------------------
// Energy, gradient and Hessian of the incremental potential (the terms of cpu_solver.cpp:99-138 and
// energies.cpp), assembled atomic-free by row owners. One lane per row evaluates its own term (inertia and
// friction) and then every term of its incidence list, in the list's order (kind, term, local vertex — the
// reference's gather order, gpu_solver_core.cpp:184-224), and keeps only its own row of the gradient and of the
// Hessian; a term's energy is counted by its local vertex 0.
//
// A term's vertices are placed relative to the term's first vertex at the substep start, x_k - xStart_0 =
// (xStart_k - xStart_0) + u_k (+ alpha step_k in a line-search trial): the energies are translation invariant,
// so no absolute coordinate enters the FP32 term arithmetic, and every row of a term evaluates it from identical
// inputs — the rows' shares of a term's force then cancel as they do in exact arithmetic (a per-row origin lets
// FP32 rounding differ between the rows and leaves a net force: a free-falling sheet drifts sideways).
//
// Entries are taken kIncBatch at a time: the batch's list words, then every record, slot list, fabric and vertex
// position it needs are issued before any term is evaluated, so a lane has a batch of independent loads in flight
// instead of one dependent chain per term. Each table is read through a selected base with a kind-dependent
// stride (no branch on the kind before the loads), and every per-term array is indexed by compile-time indices.

constexpr uint32_t kIncBatch = 4;
constexpr uint32_t kKindNone = 3;  // an empty batch position

struct IncTerm {
  uint32_t kind, a, t;
  uint4 w0, w1, w2;   // record words: stretch 48 B, hinge 32 B (w2 = w0), seam 16 B (w1 = w2 = w0)
  uint32_t slot[4];   // BSR slots of blocks (a, b), b < vertices of the term
  Fabric32 fab;       // stretch: its fabric
  Garment32 gar;      // seam: its garment
  F3 x[4];            // vertices relative to the row's start position
};

__device__ inline uint32_t incVertex(const IncTerm& it, int k) {
  // stretch: w0 = (v0, v1, v2, fabric); hinge: w0 = (v0, v1, v2, v3); seam: w0 = (a, b, garment, startGap)
  if (k == 0) return it.w0.x;
  if (k == 1) return it.w0.y;
  if (k == 2) return it.kind == kKindSeam ? it.w0.x : it.w0.z;
  return it.kind == kKindHinge ? it.w0.w : it.w0.x;
}

template <bool Trial>
__device__ inline void loadBatch(const View& v, uint32_t i, uint32_t end, float alpha, bool ownedOnly,
                                 IncTerm (&b)[kIncBatch]) {
#pragma unroll
  for (uint32_t j = 0; j < kIncBatch; ++j) {
    const uint32_t e = i + j < end ? v.inc[i + j] : 0u;
    b[j].kind = i + j < end ? incKind(e) : kKindNone;
    b[j].a = incLocal(e);
    b[j].t = incTerm(e);
    if (ownedOnly && b[j].a != 0u) b[j].kind = kKindNone;
  }
#pragma unroll
  for (uint32_t j = 0; j < kIncBatch; ++j) {
    const uint32_t k = b[j].kind, t = b[j].kind == kKindNone ? 0u : b[j].t;
    const uint4* head = k == kKindStretch ? reinterpret_cast<const uint4*>(v.stretch) + 3u * t
                        : k == kKindHinge ? reinterpret_cast<const uint4*>(v.hinges) + 2u * t
                                          : reinterpret_cast<const uint4*>(v.seams) + (k == kKindSeam ? t : 0u);
    b[j].w0 = head[0];
    b[j].w1 = head[k == kKindSeam || k == kKindNone ? 0 : 1];
    b[j].w2 = head[k == kKindStretch ? 2 : 0];
    const uint32_t m = k == kKindStretch ? 3u : k == kKindHinge ? 4u : 2u;
    const uint32_t* sl = k == kKindStretch ? v.stretchSlots + 9u * t + 3u * b[j].a
                         : k == kKindHinge ? v.hingeSlots + 16u * t + 4u * b[j].a
                                           : v.seamSlots + (k == kKindSeam ? 4u * t + 2u * b[j].a : 0u);
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) b[j].slot[q] = sl[q < m ? q : 0u];
  }
#pragma unroll
  for (uint32_t j = 0; j < kIncBatch; ++j) {
    const uint32_t k = b[j].kind;
    b[j].fab = v.fabrics[k == kKindStretch ? b[j].w0.w : 0u];
    b[j].gar = v.garments[k == kKindSeam ? b[j].w0.z : 0u];
    F3 start[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t vk = k == kKindNone ? 0u : incVertex(b[j], q);
      F3 d = v.u[vk];
      if (Trial) d += alpha * v.step[vk];
      start[q] = v.xStart[vk];
      b[j].x[q] = d;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) b[j].x[q] = (start[q] - start[0]) + b[j].x[q];
  }
}

// Seam target gap: closing from the start gap at the closing speed while the garment assembles, zero once it
// settles (sim_model.cpp:29-33).
__device__ inline float seamGap(const SegParams& P, const Garment32& g, float startGap) {
  return g.phase == 0u ? fmaxf(0.0f, startGap - P.closingSpeed * g.assemblyTime) : 0.0f;
}

// Adds block B to the row's upper slot (slot - kd) in the slice's LDS accumulator; blocks left of the diagonal are
// dropped (their transposes are assembled by the column's row).
__device__ inline void addUpper(float* acc, uint32_t slot, uint32_t kd, uint32_t lane, const Blk& B) {
  if (slot < kd) return;
  float* p = acc + ((slot - kd) * kSlice + lane) * 9u;
#pragma unroll
  for (int c = 0; c < 9; ++c) p[c] += B.e[c];
}

// Evaluates one loaded entry for its row: returns the term's energy when the row is its local vertex 0; with Deriv
// adds the row's gradient to g and its upper blocks to the accumulator.
template <bool Deriv>
__device__ inline float evalEntry(const View& v, const SegParams& P, const IncTerm& it, F3& g, float* acc,
                                  uint32_t kd, uint32_t lane) {
  float energy = 0.0f;
  if (it.kind == kKindStretch) {
    const float dmInv[4] = {__uint_as_float(it.w1.x), __uint_as_float(it.w1.y), __uint_as_float(it.w1.z),
                            __uint_as_float(it.w1.w)};
    const Shape3 sh = shapeOf(dmInv);
    const F3 xs[3] = {it.x[0], it.x[1], it.x[2]};
    const StretchForm form = stretchForm<Deriv>(xs, sh, __uint_as_float(it.w2.x), it.fab, P.limitScale);
    energy = form.energy;
    if (Deriv) {
      const StretchRow w = stretchRow(form, sh, it.a);
      g += w.g;
#pragma unroll
      for (int b = 0; b < 3; ++b) addUpper(acc, it.slot[b], kd, lane, stretchBlock(w, sh, b));
    }
  } else if (it.kind == kKindHinge) {
    const HingeForm form = hingeForm(it.x);
    const float K = __uint_as_float(it.w1.x);
    const float bend = form.theta - __uint_as_float(it.w1.y);
    energy = 0.5f * K * bend * bend;
    if (Deriv) {
      const F3 ga = pick(it.a, form.g[0], form.g[1], form.g[2], form.g[3]);
      g += (K * bend) * ga;
#pragma unroll
      for (int b = 0; b < 4; ++b) addUpper(acc, it.slot[b], kd, lane, blkOuter(ga, form.g[b], K));
    }
  } else if (it.kind == kKindSeam) {
    const float target = seamGap(P, it.gar, __uint_as_float(it.w0.w));
    const SeamForm form = seamForm<Deriv>(it.x[0], it.x[1], target, P.seamStiffness);
    energy = form.energy;
    if (Deriv && form.active) {
      g += it.a == 0u ? form.g0 : -form.g0;
      Blk neg = blkZero();
      blkAxpy(neg, -1.0f, form.S);
      addUpper(acc, it.slot[0], kd, lane, it.a == 0u ? form.S : neg);
      addUpper(acc, it.slot[1], kd, lane, it.a == 0u ? neg : form.S);
    }
  }
  return it.a == 0u ? energy : 0.0f;
}

// The row's own term: inertia 1/2 (m / h^2) |u - y|^2 for a free particle and smoothed Coulomb friction for one
// with a friction weight (kernels of cpu_solver.cpp:99-108, :127-138).
template <bool Deriv>
__device__ inline float ownEnergy(const View& v, const SegParams& P, uint32_t row, uint32_t flags, F3 disp, F3& g,
                                  Blk& H) {
  float e = 0.0f;
  if (!(flags & kRowPinned)) {
    const float k = v.mass[row] * P.invH2;
    const F3 off = disp - v.yRel[row];
    e = 0.5f * k * norm2(off);
    if (Deriv) {
      g = k * off;
      H = blkIsoPlusRank1(k, 0.0f, f3(0, 0, 0));
    }
  }
  const float lambda = v.clambda[row];
  if (lambda > 0.0f) {
    const float4 plane = v.cnormal[row];
    const float mu = v.fabrics[v.rowFabric[row]].friction;
    e += frictionAdd<Deriv>(disp, f3(plane.x, plane.y, plane.z), mu * lambda, P.frictionSlip * P.h, g, H);
  }
  return e;
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 4))) kAssemble(View v) {
  extern __shared__ float acc[];  // K * 9 * 64: the slice's upper blocks in SELL order
  const uint32_t slice = xcdMap(blockIdx.x, v.nSlices);
  if (slice >= v.nSlices) return;
  const uint32_t s = v.tileSeg[slice / kSlicesPerTile];
  if (v.state[s].newtonDone) return;
  const SegParams& P = v.seg[s];
  const uint32_t lane = threadIdx.x;
  const uint32_t row = slice * kSlice + lane;
  const uint32_t K = v.sellK[slice];
  for (uint32_t i = lane; i < K * 9u * kSlice; i += kSlice) acc[i] = 0.0f;
  __syncthreads();
  const uint32_t flags = v.rowFlags[row];
  double rowEnergy = 0.0;
  F3 g = f3(0, 0, 0);
  if (!(flags & kRowPad)) {
    const uint32_t kd = v.diagSlot[row];  // upper slot q of the row is BSR slot kd + q
    Blk own = blkZero();
    rowEnergy += double(ownEnergy<true>(v, P, row, flags, v.u[row], g, own));
    addUpper(acc, kd, kd, lane, own);
    const uint32_t end = v.incStart[row + 1];
    for (uint32_t i = v.incStart[row]; i < end; i += kIncBatch) {
      IncTerm batch[kIncBatch];
      loadBatch<false>(v, i, end, 0.0f, false, batch);
#pragma unroll
      for (uint32_t j = 0; j < kIncBatch; ++j) rowEnergy += double(evalEntry<true>(v, P, batch[j], g, acc, kd, lane));
    }
  }
  if (flags & kRowPinned) g = f3(0, 0, 0);
  v.grad[row] = g;
  // Contact release (cpu_solver.cpp:225-226): a held particle whose other forces pull it off the body is freed.
  const uint32_t cf = v.cflags[row];
  const float4 plane = v.cnormal[row];
  if ((cf & kContactActive) && dot(g, f3(plane.x, plane.y, plane.z)) < 0.0f) v.cflags[row] = cf & ~kContactActive;
  rowEnergy = waveSum(rowEnergy);
  if (lane == 0) v.partE[slice] = rowEnergy;
  __syncthreads();
  // Block-Jacobi preconditioner from the finished diagonal block: the inverse when it is positive definite, the
  // identity otherwise, zero for a pinned row (cpu_solver.cpp:171-178); stored as its upper triangle.
  float inv[kPrecondFloats] = {1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 1.0f};
  if (flags & kRowPinned) {
#pragma unroll
    for (uint32_t c = 0; c < kPrecondFloats; ++c) inv[c] = 0.0f;
  } else {
    float diag[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) diag[c] = acc[lane * 9u + c];
    float m[kPrecondFloats];
    if (spdInverseUpper(diag, m))
#pragma unroll
      for (uint32_t c = 0; c < kPrecondFloats; ++c) inv[c] = m[c];
  }
#pragma unroll
  for (uint32_t c = 0; c < kPrecondFloats; ++c) v.precond[kPrecondFloats * row + c] = inv[c];
  float* out = v.sellVal + v.sellBase[slice];
  for (uint32_t i = lane; i < K * 9u * kSlice; i += kSlice) out[i] = acc[i];
}

// e0 of this Newton iteration = sum of the assembly's energy partials (gpu_solver_core.cpp:344).
__global__ void __launch_bounds__(256) kNewtonBegin(View v) {
  __shared__ double lds[8];
  const uint32_t s = blockIdx.x;
  SegState* S = v.state + s;
  if (S->newtonDone) return;
  const SegParams& P = v.seg[s];
  const double e = rangeSum256(v.partE, P.tileBegin * kSlicesPerTile, P.tileEnd * kSlicesPerTile, lds);
  if (threadIdx.x == 0) {
    S->e0 = e;
    S->cntNewton += 1u;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Block-Jacobi PCG restricted to the contact planes (cpu_solver.cpp:169-213; kernels.h:237-347), FP32 vectors and
// FP64 dot products. Per iteration two kernels; the search-direction update p = z + beta p is folded into the
// next SpMV (each lane recomputes it for the columns it gathers), with p double-buffered by iteration parity.

__global__ void __launch_bounds__(256) kPcgInit(View v) {
  __shared__ double lds[8];
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  SegState* S = v.state + s;
  if (S->newtonDone) return;
  const SegParams& P = v.seg[s];
  if (tile == P.tileBegin && threadIdx.x == 0) {
    S->pcgConverged = 0u;
    S->pcgIter = 0u;
  }
  const uint32_t row = tile * kTile + threadIdx.x;
  const uint32_t cf = v.cflags[row];
  const float4 cn = v.cnormal[row];
  const F3 ri = onPlane(cf, cn, -v.grad[row]);
  const F3 zi = applyPrecond(v.precond + kPrecondFloats * row, cf, cn, ri);
  v.step[row] = f3(0, 0, 0);
  v.r[row] = ri;
  v.z[row] = zi;
  v.pA[row] = zi;
  const double rz = double(ri.x) * zi.x + double(ri.y) * zi.y + double(ri.z) * zi.z;
  const double rr = double(ri.x) * ri.x + double(ri.y) * ri.y + double(ri.z) * ri.z;
  const double srz = blockSum256(rz, lds);
  const double srr = blockSum256(rr, lds);
  if (threadIdx.x == 0) {
    v.partA[tile] = srz;
    v.partB[tile] = srr;
  }
}

// Iteration k, first half: convergence test of the previous update, beta, p = z + beta p, ap = filter(A p), p.ap.
__global__ void __launch_bounds__(256) kPcgSpmv(View v, uint32_t k, const F3* __restrict__ pOld, F3* pNew) {
  __shared__ double lds[8];
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  SegState* S = v.state + s;
  const SegParams& P = v.seg[s];
  const double rzNew = rangeSum256(v.partA, P.tileBegin, P.tileEnd, lds);
  const double rr = rangeSum256(v.partB, P.tileBegin, P.tileEnd, lds);
  if (S->newtonDone) return;
  const bool first = tile == P.tileBegin && threadIdx.x == 0;
  uint32_t conv;
  double beta = 0.0;
  if (k == 0) {
    const double r0 = sqrt(rr);
    conv = r0 > 0.0 ? 0u : 1u;
    if (first) {
      S->rzPar[0] = rzNew;
      S->r0 = r0;
      S->rr = rr;
      S->beta = 0.0;
      if (conv) {
        S->pcgConverged = 1u;
        S->pcgExit = kPcgExitZero;
      }
    }
  } else {
    conv = S->pcgConverged;
    if (!conv) {
      if (sqrt(rr) <= P.cgTolerance * S->r0) {
        conv = 1u;
        if (first) S->pcgExit = kPcgExitTolerance;
      } else {
        beta = rzNew / S->rzPar[(k + 1u) & 1u];
      }
    }
    if (first) {
      S->rzPar[k & 1u] = rzNew;
      S->rr = rr;
      S->beta = beta;
      if (conv) S->pcgConverged = 1u;
    }
  }
  if (k >= P.maxCg) {
    if (first && !conv) {
      S->pcgConverged = 1u;
      S->pcgExit = kPcgExitCap;
    }
    conv = 1u;
  }
  if (conv) return;
  const float b = float(beta);
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t slice = tile * kSlicesPerTile + wave;
  const uint32_t row = slice * kSlice + lane;
  const F3* __restrict__ z = v.z;
  const F3 prod = sellRowProduct(v, slice, lane, [&](uint32_t c) { return z[c] + b * pOld[c]; });
  const uint32_t flags = v.rowFlags[row];
  F3 acc = (flags & kRowPinned) ? f3(0, 0, 0) : prod;
  const F3 pi = z[row] + b * pOld[row];
  pNew[row] = pi;
  const F3 api = onPlane(v.cflags[row], v.cnormal[row], acc);
  v.ap[row] = api;
  const double pap = double(pi.x) * api.x + double(pi.y) * api.y + double(pi.z) * api.z;
  const double sp = blockSum256(pap, lds);
  if (threadIdx.x == 0) v.partC[tile] = sp;
}

// Iteration k, second half: pAp, alpha, step += alpha p, r -= alpha ap, z = filter(M r); partials r.r and r.z.
__global__ void __launch_bounds__(256) kPcgUpdate(View v, uint32_t k, const F3* __restrict__ p) {
  __shared__ double lds[8];
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  SegState* S = v.state + s;
  const SegParams& P = v.seg[s];
  const double pAp = rangeSum256(v.partC, P.tileBegin, P.tileEnd, lds);
  if (S->newtonDone || S->pcgConverged || k >= P.maxCg) return;
  const bool first = tile == P.tileBegin && threadIdx.x == 0;
  if (!(pAp > 0.0)) {
    if (first) {
      S->pcgConverged = 1u;
      S->pcgExit = kPcgExitCurvature;
    }
    return;
  }
  const double alpha = S->rzPar[k & 1u] / pAp;
  if (first) {
    S->pAp = pAp;
    S->alpha = alpha;
    S->pcgIter += 1u;
    S->cntPcg += 1u;
  }
  const float a = float(alpha);
  const uint32_t row = tile * kTile + threadIdx.x;
  const F3 pi = p[row];
  v.step[row] = v.step[row] + a * pi;
  const F3 ri = v.r[row] - a * v.ap[row];
  v.r[row] = ri;
  const F3 zi = applyPrecond(v.precond + kPrecondFloats * row, v.cflags[row], v.cnormal[row], ri);
  v.z[row] = zi;
  const double rr = double(ri.x) * ri.x + double(ri.y) * ri.y + double(ri.z) * ri.z;
  const double rz = double(ri.x) * zi.x + double(ri.y) * zi.y + double(ri.z) * zi.z;
  const double srr = blockSum256(rr, lds);
  const double srz = blockSum256(rz, lds);
  if (threadIdx.x == 0) {
    v.partA[tile] = srz;
    v.partB[tile] = srr;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// The same block-Jacobi PCG with one kernel and one global reduction per iteration.
//
// Kernel k applies update k of the standard iteration (step += alpha_{k-1} p_{k-1}, r_k = r_{k-1} -
// alpha_{k-1} ap_{k-1}), tests convergence, forms p_k = z_k + beta_k p_{k-1} and ap_k = S A p_k, and writes the
// partial dot products the next kernel needs. Standard PCG needs a second global reduction after its update, for
// r_k.z_k (beta) and r_k.r_k (the stopping test); here both are expanded over the previous iteration's vectors,
//   r_k.z_k = r.zr - 2 alpha r.t + alpha^2 ap.t,    r_k.r_k = r.r - 2 alpha r.ap + alpha^2 ap.ap,
// with zr = Mf r and t = Mf ap computed directly by each row's owner (Mf x = S M x: the block-Jacobi inverse, then
// the contact filter). Every term is a dot product of stored vectors, so no scalar is carried by recurrence and no
// orthogonality is assumed; p.ap is reduced directly, as in the standard form. A gathered column's z_k is rebuilt
// from its owner's zr and t (z_k = zr - alpha t = Mf r_k). Stopping rules of cpu_solver.cpp:188-213: r0 = 0,
// p.Ap <= 0 before an update, |r| <= tol |r0| after one, at most maxCg updates.
//
// Matrix rows: W wavefronts share each 64-row SELL slice (column t goes to wavefront t mod W) and combine their
// row sums through LDS in a fixed order that does not depend on W, so one 256-row tile runs as 4 W wavefronts. Every load that does not
// depend on the scalars (column indices, block values and the gathered zr, t, p of the first batch of columns) is
// issued before the segment reduction.

namespace {
constexpr int kPcgParts = 7;  // per tile: p.ap, r.zr, r.t, ap.t, r.ap, ap.ap, r.r (kPartStride doubles apart)
static_assert(kPcgParts <= int(kPartStride), "a tile's partial sums overlap the next tile's");
enum : int { kPpAp = 0, kPrZ, kPrT, kPapT, kPrAp, kPapAp, kPrR };
}  // namespace

// A double moved between lanes by a DPP pattern (both halves by v_mov_b32_dpp); lanes the pattern leaves without
// a source (rows outside RowMask) read +0.
template <int Ctrl, int RowMask = 0xf>
__device__ inline double dppMove(double v) {
  const uint64_t u = uint64_t(__double_as_longlong(v));
  const uint32_t lo = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(u)), Ctrl, RowMask, 0xf, false));
  const uint32_t hi = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(u >> 32)), Ctrl, RowMask, 0xf, false));
  return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}

// Sum over the 64 lanes of a wavefront in a fixed order, returned to every lane, without LDS: pairs and quads
// (quad_perm), half rows and rows of 16 (mirrors), then the rows (row_bcast15 into rows 1 and 3, row_bcast31 into
// rows 2 and 3); lane 63 then holds the total.
__device__ inline double waveTotal(double v) {
  v += dppMove<0xB1>(v);        // quad_perm [1,0,3,2]
  v += dppMove<0x4E>(v);        // quad_perm [2,3,0,1]
  v += dppMove<0x141>(v);       // row_half_mirror
  v += dppMove<0x140>(v);       // row_mirror
  v += dppMove<0x142, 0xA>(v);  // row_bcast15
  v += dppMove<0x143, 0xC>(v);  // row_bcast31
  const uint64_t u = uint64_t(__double_as_longlong(v));
  const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u)), 63));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u >> 32)), 63));
  return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}

// One level of waveTotal applied to N independent values, each move issued before any add: the N dependency chains
// interleave instead of running one after another.
template <int Ctrl, int RowMask, int N>
__device__ inline void dppAddN(double (&x)[N]) {
  double t[N];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = dppMove<Ctrl, RowMask>(x[i]);
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] += t[i];
}

// waveTotal of N values at once (the same moves and adds per value, so the same sums).
template <int N>
__device__ inline void waveTotalN(double (&x)[N]) {
  dppAddN<0xB1, 0xf>(x);
  dppAddN<0x4E, 0xf>(x);
  dppAddN<0x141, 0xf>(x);
  dppAddN<0x140, 0xf>(x);
  dppAddN<0x142, 0xA>(x);
  dppAddN<0x143, 0xC>(x);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint64_t u = uint64_t(__double_as_longlong(x[i]));
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u)), 63));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u >> 32)), 63));
    x[i] = __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
  }
}

// N sums over a workgroup of NW wavefronts (fixed order), returned to every thread. lds >= N * NW doubles.
// Fresh: `lds` has not been read earlier in this kernel, so the barrier that protects its previous readers is
// left out.
template <int N, int NW, bool Fresh = false>
__device__ inline void blockSumN(double (&x)[N], double* lds) {
  waveTotalN(x);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if constexpr (!Fresh) __syncthreads();
  if (l == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) lds[i * NW + w] = x[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NW; ++j) s += lds[i * NW + j];
    x[i] = s;
  }
}

// The search direction of a row from what its owner stored: p_k = (zr - alpha t) + beta p_{k-1}. The owner and
// every row gathering it evaluate this same expression.
__device__ inline F3 pcgDirection(F3 zr, F3 t, F3 p, float alpha, float beta) {
  return (zr - alpha * t) + beta * p;
}

__device__ inline double dot64(F3 a, F3 b) { return double(a.x) * b.x + double(a.y) * b.y + double(a.z) * b.z; }

// Init: r_0 = S(-grad), zr_0 = Mf r_0, written as the outputs of a kernel "-1" (slot 1) with ap = p = t = 0;
// partials r.zr and r.r.
__global__ void __launch_bounds__(256) kPcgStepInit(View v) {
  __shared__ double lds[kPcgParts * 4];
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  SegState* S = v.state + s;
  const uint32_t done = S->newtonDone;
  if (threadIdx.x == 0) v.tileDone[0][tile] = done;  // read by the first kPcgStep of this solve
  if (done) return;
  const SegParams& P = v.seg[s];
  if (tile == P.tileBegin && threadIdx.x == 0) {
    S->pcgConverged = 0u;
    S->pcgIter = 0u;
    S->kIter[0] = 0u;
    S->restartActive = 0u;
    S->kBase = 0u;
    S->restartsThisSolve = 0u;
  }
  const uint32_t row = tile * kTile + threadIdx.x;
  const uint32_t cf = v.cflags[row];
  const float4 cn = v.cnormal[row];
  const F3 r = onPlane(cf, cn, -v.grad[row]);
  const F3 zr = applyPrecond(v.precond + kPrecondFloats * row, cf, cn, r);
  const F3 zero = f3(0, 0, 0);
  v.cgR[1][row] = r;
  v.cgAp[1][row] = zero;
  v.cgP[1][row] = zero;
  v.cgZr[1][row] = zr;
  v.cgT[1][row] = zero;
  v.step[row] = zero;
  double d[kPcgParts] = {0.0, dot64(r, zr), 0.0, 0.0, 0.0, 0.0, dot64(r, r)};
  blockSumN<kPcgParts, 4>(d, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < kPcgParts; ++i) v.cgPart[1][kPartStride * tile + i] = d[i];
}

// A vector store the kernel never reads back (the next launch reads it from beyond the L2): marked nontemporal.
__device__ inline void storeStreaming(F3* p, F3 x) {
  __builtin_nontemporal_store(x.x, &p->x);
  __builtin_nontemporal_store(x.y, &p->y);
  __builtin_nontemporal_store(x.z, &p->z);
}

// A vector load of data no other lane of the launch reads: marked nontemporal, so that it does not push out of
// the L2 the lines the gathers and the transposed-block loads reuse.
__device__ inline F3 loadStreaming(const F3* p) {
  return f3(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y), __builtin_nontemporal_load(&p->z));
}

// One of a View's parity-buffered pointers by a select between the two, which are kernel arguments: indexing the
// pair with the parity would turn into a second, dependent load from the argument segment.
template <class Ptr>
__device__ inline Ptr par(Ptr const (&pair)[2], uint32_t parity) {
  return parity ? pair[1] : pair[0];
}

// Launch shapes: W wavefronts share each 64-row slice (each takes every W-th slot), and a batch covers NB slots of
// the lower part and NB of the upper part (B = NB / W per wavefront). NB = 8 keeps a whole sheet row in one batch
// (about 390 registers: one workgroup per CU); NB = 4 fits two workgroups per CU, which pays when a launch has
// many more tiles than CUs.
//
// H (halo, W = 1 only): when every tile's columns lie in one window of at most kHaloRows rows (a banded matrix such
// as a grid sheet's), the tile loads that window's (zr, t, p) with coalesced loads in its first round trip, keeps
// it in LDS, and gathers from there: the second round trip then carries only the transposed lower blocks.
//
// U (1x8 only, for a launch of one segment in the uniform layout): every slice has the same upper and lower slot
// counts (at most 8: one batch), so its matrix sits at slice times those counts, and the segment's tiles are all
// tiles: the partial sums and the first round trip (indices and block values) are issued at once, without
// waiting for the slice headers.
template <int W, int NB, bool H, bool U = false>
__global__ void __launch_bounds__(256 * W) __attribute__((amdgpu_waves_per_eu(NB == 4 ? 2 : 1, 8)))
kPcgStep(View v, uint32_t q) {
  static_assert(!H || W == 1, "the halo form has one wavefront per slice");
  static_assert(!U || (W == 1 && NB == int(kUniformSlots) && !H), "the uniform form is the 1x8 shape");
  // (In the uniform form K and KL are at most NB: the first batch is the only one.)
  constexpr int NW = 4 * W;
  constexpr uint32_t B = NB / W;  // lower slots, and upper slots, per wavefront in one batch
  constexpr uint32_t HR = H ? kHaloRows / 256u : 1u;  // halo rows staged per thread
  __shared__ double red[kPcgParts * NW], red2[kPcgParts * NW];  // one buffer per reduction: no barrier before either
  __shared__ float comb[(W > 1 ? 4 : 1) * kSlicesPerTile * 3 * kSlice];
  __shared__ F3 hz[H ? kHaloRows : 1], ht[H ? kHaloRows : 1], hp[H ? kHaloRows : 1];
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t o = 1u - q;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t sl = wave & 3u, part = W == 1 ? 0u : wave >> 2;
  // Slice-uniform values through the scalar cache, issued with the segment index: neither depends on the other.
  const uint32_t slice = __builtin_amdgcn_readfirstlane(tile * kSlicesPerTile + sl);
  const uint32_t row = slice * kSlice + lane;
  const F3* __restrict__ Zr = par(v.cgZr, o);
  const F3* __restrict__ T = par(v.cgT, o);
  const F3* __restrict__ Pp = par(v.cgP, o);
  // The owner row's data depends on nothing loaded. The 8-slot shapes (one workgroup per CU, for launches of at
  // most about one tile per CU, whose tiles rarely find their solve ended) issue it first, so that it streams in
  // while the slice headers make their scalar round trip; the 4-slot shapes, for launches with many tiles per CU,
  // issue it after the solve-ended test, so that the tiles of converged segments exit without loading it.
  constexpr bool kOwnFirst = NB == 8;
  F3 oR, oAp, oZ, oT, oP, oStep;
  float m[kPrecondFloats];
  uint32_t oFlags = 0u, oCf = 0u;
  float4 oCn = make_float4(0.f, 0.f, 0.f, 0.f);
  // (zr, t and p are gathered by the neighbours as well; the rest is read by this row only: nontemporal.)
  auto loadOwner = [&] {
    if (part == 0) {
      oR = loadStreaming(par(v.cgR, o) + row);
      oAp = loadStreaming(par(v.cgAp, o) + row);
      if constexpr (!H) {
        oZ = Zr[row];
        oT = T[row];
        oP = Pp[row];
      }
      oStep = loadStreaming(v.step + row);
#pragma unroll
      for (uint32_t e = 0; e < kPrecondFloats; ++e)
        m[e] = __builtin_nontemporal_load(v.precond + kPrecondFloats * row + e);
      oFlags = __builtin_nontemporal_load(v.rowFlags + row);
      oCf = __builtin_nontemporal_load(v.cflags + row);
      const float* cn = reinterpret_cast<const float*>(v.cnormal + row);
      oCn = make_float4(__builtin_nontemporal_load(cn), __builtin_nontemporal_load(cn + 1),
                        __builtin_nontemporal_load(cn + 2), __builtin_nontemporal_load(cn + 3));
    }
  };
  if constexpr (kOwnFirst) loadOwner();
  __builtin_amdgcn_sched_barrier(0);
  const uint32_t K = U ? v.uniformK : ldUniform(v.sellK + slice);
  const uint32_t base = U ? slice * v.uniformK * 9u * kSlice : ldUniform(v.sellBase + slice);
  const uint32_t KL = U ? v.uniformKL : ldUniform(v.lowK + slice);
  const uint32_t lbase = U ? slice * v.uniformKL * kSlice : ldUniform(v.lowBase + slice);

  // Round trip 1 (everything here depends only on the slice and the parity): the first batch's columns (lower and
  // upper), lower block indices and upper block values. Index loads first (unconditional, see sellColumn), block
  // loads after, and nothing that uses an index before both are issued: the wait before the gathers is then one
  // for the indices only (the memory counter retires in issue order), not for the 36 B per slot of block values
  // still in flight. Issued at once in the uniform layout, else after the slice headers and the solve-ended test.
  const uint32_t* __restrict__ col = v.sellCol + base / 9u + lane;
  const float* __restrict__ val = v.sellVal + base + 9u * lane;
  const uint32_t* __restrict__ lcol = v.lowCol + lbase + lane;
  const uint32_t* __restrict__ lblk = v.lowBlk + lbase + lane;
  constexpr uint32_t HR0 = HR;  // (named for the lambda below)
  uint32_t c[B], cl[B], bl[B];
  float sv[HR0][9];
  float bv[B][9];
  uint32_t hb = 0u, hc = 0u;
  auto issueMatrix = [&] {
#pragma unroll
    for (uint32_t j = 0; j < B; ++j) {
      cl[j] = __builtin_nontemporal_load(lcol + (part + W * j) * kSlice);
      bl[j] = __builtin_nontemporal_load(lblk + (part + W * j) * kSlice);
    }
#pragma unroll
    for (uint32_t j = 0; j < B; ++j) c[j] = __builtin_nontemporal_load(col + (part + W * j) * kSlice);
    // The halo window, coalesced, before the block values: the wait before it goes to LDS then covers the indices
    // and the window only. (Threads past the window's end re-read its last row and store it past the end, where
    // no gather looks: loads and stores stay unconditional, so the staging arrays live in registers.)
    if constexpr (H) {
#pragma unroll
      for (uint32_t r = 0; r < HR0; ++r) {
        const uint32_t i = hb + min(r * 256u + threadIdx.x, hc - 1u);
        const F3 z = Zr[i], tt = T[i], pp = Pp[i];
        sv[r][0] = z.x, sv[r][1] = z.y, sv[r][2] = z.z;
        sv[r][3] = tt.x, sv[r][4] = tt.y, sv[r][5] = tt.z;
        sv[r][6] = pp.x, sv[r][7] = pp.y, sv[r][8] = pp.z;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    // The first batch's upper block values. Slots past K are read unconditionally — the value buffer carries
    // kSellSlack zero floats past its end — because their column is the dummy one, whose gathered vectors are
    // zero, so they add exactly nothing.
#pragma unroll
    for (uint32_t j = 0; j < B; ++j) {
      const uint32_t t = part + W * j;
#pragma unroll
      for (int e = 0; e < 9; ++e) bv[j][e] = val[t * 9u * kSlice + e];
    }
    if constexpr (!kOwnFirst) loadOwner();
    __builtin_amdgcn_sched_barrier(0);
  };
  // The previous kernel's partial sums go before the matrix: the reduction over them then runs while this
  // kernel's block values are still streaming in. The segment's tile range comes per tile, with the slice headers
  // (in the uniform form the launch is one segment: all tiles).
  const double* __restrict__ prev = par(v.cgPart, o);
  double sum[kPcgParts];
  uint32_t ti = 0u, tileBegin = 0u, tileEnd = v.nTiles;
  auto loadPartials = [&] {
    // The first 256 threads take tiles t, t + 256, ... whatever the shape, and the wavefronts past the fourth add
    // zeros: the sums round the same way in every shape.
    ti = threadIdx.x < 256u ? tileBegin + threadIdx.x : tileEnd;
#pragma unroll
    for (int n = 0; n < kPcgParts; ++n) sum[n] = ti < tileEnd ? prev[kPartStride * ti + n] : 0.0;
    __builtin_amdgcn_sched_barrier(0);
  };
  if constexpr (U) {
    loadPartials();
    issueMatrix();
  }

  const uint32_t s = U ? 0u : ldUniform(v.tileSeg + tile);
  if constexpr (!U) {
    tileBegin = ldUniform(v.tileRange + 2u * tile);
    tileEnd = ldUniform(v.tileRange + 2u * tile + 1u);
  }
  if constexpr (H) {
    hb = ldUniform(v.tileHalo + 2u * tile);
    hc = ldUniform(v.tileHalo + 2u * tile + 1u);
  }
  // A tile whose solve has ended (the segment's PCG converged, or its Newton loop finished) passes the flag on and
  // loads nothing more. The flag is per tile, so it arrives with the slice headers: no extra round trip for the
  // tiles that iterate, and none of the slice's matrix loads for a tile of a batch whose segment converged early.
  if (ldUniform(par(v.tileDone, q) + tile)) {
    if (threadIdx.x == 0) par(v.tileDone, o)[tile] = 1u;
    return;
  }
  if constexpr (!U) {
    loadPartials();
    issueMatrix();
  }

  // Per-segment scalars through the scalar cache: none of them is written by this kernel (it writes only the
  // other parity slot of kIter, and r0 when k == 0, where it does not read it).
  SegState* S = v.state + s;
  const SegParams* P = v.seg + s;
  const uint32_t maxCg = ldUniform(&P->maxCg);
  const double cgTolerance = ldUniform(&P->cgTolerance);
  const uint32_t k = ldUniform(&S->kIter[q]);
  const double r0Prev = ldUniform(&S->r0);
  // A continuation from the FP64 true residual keeps the solve's r0 and its iteration count.
  const uint32_t restarted = ldUniform(&S->restartActive);
  const uint32_t kDone = k + ldUniform(&S->kBase);
  // (zr, t, p) of a column: from the halo window in LDS, or gathered from memory; the dummy column gives zero.
  auto gather = [&](uint32_t cc, F3& z, F3& tt, F3& pp) {
    if constexpr (H) {
      const bool dummy = cc == v.nRows;
      const uint32_t i = dummy ? 0u : cc - hb;
      z = dummy ? f3(0, 0, 0) : hz[i];
      tt = dummy ? f3(0, 0, 0) : ht[i];
      pp = dummy ? f3(0, 0, 0) : hp[i];
    } else {
      z = Zr[cc];
      tt = T[cc];
      pp = Pp[cc];
    }
  };
  // Round trip 2: the transposed lower blocks, and (without the halo) the gathered (zr, t, p) of the lower and
  // upper columns.
#pragma unroll
  for (uint32_t j = 0; j < B; ++j) {
    const uint32_t t = part + W * j;
    cl[j] = t < KL ? cl[j] : v.nRows;
    bl[j] = t < KL ? bl[j] : v.zeroBlock;
    c[j] = t < K ? c[j] : v.nRows;
  }
  float bt[B][9];
#pragma unroll
  for (uint32_t j = 0; j < B; ++j)
#pragma unroll
    for (int e = 0; e < 9; ++e) bt[j][e] = v.sellVal[size_t(bl[j]) * 9u + e];  // 64-bit index: one wide load
  // With the halo, the window goes to LDS only once the transposed blocks are in flight: its wait then overlaps
  // their round trip.
  if constexpr (H) {
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (uint32_t r = 0; r < HR; ++r) {
      const uint32_t i = r * 256u + threadIdx.x;
      hz[i] = f3(sv[r][0], sv[r][1], sv[r][2]);
      ht[i] = f3(sv[r][3], sv[r][4], sv[r][5]);
      hp[i] = f3(sv[r][6], sv[r][7], sv[r][8]);
    }
    __syncthreads();
    gather(row, oZ, oT, oP);
  }
  constexpr uint32_t GB = H ? 1u : B;  // gathered values held in registers (none with the halo)
  F3 lz[GB], lt[GB], lp[GB], gz[GB], gt[GB], gp[GB];
  if constexpr (!H) {
#pragma unroll
    for (uint32_t j = 0; j < B; ++j) gather(cl[j], lz[j], lt[j], lp[j]);
#pragma unroll
    for (uint32_t j = 0; j < B; ++j) gather(c[j], gz[j], gt[j], gp[j]);
  }

  // Scalars of this kernel (identical in every workgroup of the segment), while the gathers are in flight.
  for (ti += 256u; ti < tileEnd; ti += 256u)
#pragma unroll
    for (int n = 0; n < kPcgParts; ++n) sum[n] += prev[kPartStride * ti + n];
  blockSumN<kPcgParts, NW, true>(sum, red);
  uint32_t conv = 0u;
  bool update = false;
  double alpha = 0.0, beta = 0.0, r0 = 0.0, rrNext = sum[kPrR];
  if (k == 0) {
    r0 = restarted ? r0Prev : sqrt(sum[kPrR]);
    if (!(sum[kPrR] > 0.0))
      conv = restarted ? kPcgExitTolerance : kPcgExitZero;
    else if (kDone >= maxCg)
      conv = kPcgExitCap;
  } else if (!(sum[kPpAp] > 0.0)) {
    conv = kPcgExitCurvature;
  } else {
    update = true;
    alpha = sum[kPrZ] / sum[kPpAp];
    rrNext = sum[kPrR] - 2.0 * alpha * sum[kPrAp] + alpha * alpha * sum[kPapAp];
    const double rzNext = sum[kPrZ] - 2.0 * alpha * sum[kPrT] + alpha * alpha * sum[kPapT];
    if (sqrt(fmax(rrNext, 0.0)) <= cgTolerance * r0Prev)
      conv = kPcgExitTolerance;
    else if (kDone >= maxCg)
      conv = kPcgExitCap;
    else
      beta = rzNext / sum[kPrZ];
  }
  if (threadIdx.x == 0) par(v.tileDone, o)[tile] = conv ? 1u : 0u;  // every workgroup of the segment decides the same
  if (tile == tileBegin && threadIdx.x == 0) {
    if (k == 0) S->r0 = r0;
    S->kIter[o] = k + 1u;
    if (update) {
      S->alpha = alpha;
      S->rr = rrNext;
      S->pcgIter += 1u;
      S->cntPcg += 1u;
    }
    if (conv) {
      S->pcgConverged = 1u;
      S->pcgExit = conv;
    }
  }
  if (!update && conv) return;
  const float a = float(alpha), b = float(beta);
  if (part == 0 && update) storeStreaming(v.step + row, oStep + a * oP);
  if (conv) return;

  // Row sums of A p_k. Slots are summed in four groups (slot index mod 4); each group takes its lower slots, then
  // its upper slots, in increasing slot order, and the groups are added as ((g0 + g1) + g2) + g3: every launch shape
  // gives bit-identical results.
  constexpr uint32_t G = 4u / W;  // groups held by one wavefront: group part + W lg
  float acc[G][3];
#pragma unroll
  for (uint32_t lg = 0; lg < G; ++lg) acc[lg][0] = acc[lg][1] = acc[lg][2] = 0.0f;
#pragma unroll
  for (uint32_t j = 0; j < B; ++j) {
    F3 z, tt, pp;
    if constexpr (H) gather(cl[j], z, tt, pp);
    else { z = lz[j]; tt = lt[j]; pp = lp[j]; }
    blockTMulAdd(bt[j], pcgDirection(z, tt, pp, a, b), acc[j % G][0], acc[j % G][1], acc[j % G][2]);
  }
  for (uint32_t j0 = B; !U && part + W * j0 < KL; j0 += B) {  // rows with more than NB lower slots
#pragma unroll
    for (uint32_t j = 0; j < B; ++j) {
      const uint32_t t = part + W * (j0 + j);
      cl[j] = sellColumn(lcol, t, KL, v.nRows);
      bl[j] = sellColumn(lblk, t, KL, v.zeroBlock);
    }
#pragma unroll
    for (uint32_t j = 0; j < B; ++j) {
      const uint32_t t = part + W * (j0 + j);
      F3 z, tt, pp;
      gather(cl[j], z, tt, pp);
      const F3 x = pcgDirection(z, tt, pp, a, b);
      if (t < KL) blockTMulAdd(v.sellVal + size_t(bl[j]) * 9u, x, acc[j % G][0], acc[j % G][1], acc[j % G][2]);
    }
  }
#pragma unroll
  for (uint32_t j = 0; j < B; ++j) {
    F3 z, tt, pp;
    if constexpr (H) gather(c[j], z, tt, pp);
    else { z = gz[j]; tt = gt[j]; pp = gp[j]; }
    blockMulAdd(bv[j], pcgDirection(z, tt, pp, a, b), acc[j % G][0], acc[j % G][1], acc[j % G][2]);
  }
  for (uint32_t j0 = B; !U && part + W * j0 < K; j0 += B) {  // rows with more than NB upper slots
#pragma unroll
    for (uint32_t j = 0; j < B; ++j) c[j] = sellColumn(col, part + W * (j0 + j), K, v.nRows);
#pragma unroll
    for (uint32_t j = 0; j < B; ++j) {
      const uint32_t t = part + W * (j0 + j);
      F3 z, tt, pp;
      gather(c[j], z, tt, pp);
      const F3 x = pcgDirection(z, tt, pp, a, b);
      if (t < K) blockMulAdd(val + t * 9u * kSlice, x, acc[j % G][0], acc[j % G][1], acc[j % G][2]);
    }
  }
  if constexpr (W > 1) {
    if (part > 0) {
#pragma unroll
      for (uint32_t lg = 0; lg < G; ++lg) {
        float* cb = comb + ((part + W * lg) * kSlicesPerTile + sl) * 3u * kSlice + lane;
        cb[0] = acc[lg][0];
        cb[kSlice] = acc[lg][1];
        cb[2 * kSlice] = acc[lg][2];
      }
    }
    __syncthreads();
  }
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  if (part == 0) {
    float grp[4][3];
#pragma unroll
    for (uint32_t g = 0; g < 4; ++g) {
      if (g % W == 0) {
        grp[g][0] = acc[g / W][0];
        grp[g][1] = acc[g / W][1];
        grp[g][2] = acc[g / W][2];
      } else {
        const float* cb = comb + (g * kSlicesPerTile + sl) * 3u * kSlice + lane;
        grp[g][0] = cb[0];
        grp[g][1] = cb[kSlice];
        grp[g][2] = cb[2 * kSlice];
      }
    }
    a0 = ((grp[0][0] + grp[1][0]) + grp[2][0]) + grp[3][0];
    a1 = ((grp[0][1] + grp[1][1]) + grp[2][1]) + grp[3][1];
    a2 = ((grp[0][2] + grp[1][2]) + grp[2][2]) + grp[3][2];
  }

  // Owner rows: r_k, p_k, ap_k, zr_k = Mf r_k, t_k = Mf ap_k; partials for the next kernel.
  double d[kPcgParts] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (part == 0) {
    const F3 rk = oR - a * oAp;
    const F3 pk = pcgDirection(oZ, oT, oP, a, b);
    const F3 apk = (oFlags & kRowPinned) ? f3(0, 0, 0) : onPlane(oCf, oCn, f3(a0, a1, a2));
    const F3 zr = applyPrecond(m, oCf, oCn, rk);
    const F3 tk = applyPrecond(m, oCf, oCn, apk);
    storeStreaming(par(v.cgR, q) + row, rk);
    storeStreaming(par(v.cgAp, q) + row, apk);
    storeStreaming(par(v.cgP, q) + row, pk);
    storeStreaming(par(v.cgZr, q) + row, zr);
    storeStreaming(par(v.cgT, q) + row, tk);
    d[kPpAp] = dot64(pk, apk);
    d[kPrZ] = dot64(rk, zr);
    d[kPrT] = dot64(rk, tk);
    d[kPapT] = dot64(apk, tk);
    d[kPrAp] = dot64(rk, apk);
    d[kPapAp] = dot64(apk, apk);
    d[kPrR] = dot64(rk, rk);
  }
  blockSumN<kPcgParts, NW, true>(d, red2);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < kPcgParts; ++i) par(v.cgPart, q)[kPartStride * tile + i] = d[i];
}

// ---------------------------------------------------------------------------------------------------------------
// Persistent solve (kPcgSolve): one launch runs the one-reduction PCG from the state kPcgStepInit (or a refinement's
// kTrueResidual) left in parity slot 1 until every segment's solve has ended. Iteration k computes exactly what the
// k-th kPcgStep<1, 8, true> launch computes (same expressions, same summation orders), so the two forms give
// bit-identical results.
//
// One workgroup owns one 256-row tile for the whole solve, one wavefront per 64-row slice, each wavefront with a
// SIMD to itself, the gathered window in registers. Where the tiles outnumber the CUs, kPcgSolvePair (below) runs
// two tiles per 512-thread workgroup, one workgroup per CU, with the window staged in LDS. The slice's matrix (upper block values
// and columns, the lower slots' transposed blocks and columns: the first 8 + 8 slots; with Extra, a seam row's slots
// past the eighth, at most 16 + 16, are read from memory every iteration) and its rows' private PCG data (r, ap,
// step, the preconditioner, the row and contact flags) are loaded once and stay in registers. An iteration moves only
// the (zr, t, p) window the tile gathers from (up to kSolveWindows row ranges), the tile's own new (zr, t, p) and 7
// FP64 partial sums: about 80 B per row against kPcgStep's 500.
//
// Residency. A segment's tiles wait for each other every iteration, so all of them must be resident together. The
// host launches either every tile at once (one or two per CU), or, when the tiles outnumber the resident
// workgroups, a queue: workgroup i takes the tile order[i], each segment's tiles consecutive. Workgroups are
// dispatched in order, round-robin over the eight XCDs, so a segment of at most 512 tiles has at most 64 on any XCD
// (32 CUs, two each), and the earliest segment
// not yet finished has all its tiles resident (everything dispatched before it has left): it finishes and frees its
// CUs for the segments behind it. No wait relies on this for termination: every wait is bounded (below).
//
// Exchange between workgroups. The L2 of each XCD is not coherent with the others, and an agent-scope release
// fence writes back a whole L2 (buffer_wbl2), which costs microseconds when every workgroup of an XCD issues one.
// So every value another workgroup reads during the launch is stored with device-scope stores (sc1: written
// through to the device-coherent level) and loaded with device-scope loads (sc1), and a workgroup announces its
// stores only after every wavefront has waited for their acknowledgement (s_waitcnt vmcnt(0)). The announcement is
// a device-scope atomic on the segment's counters, and the waits poll a generation line with device-scope loads.
// Nothing that changes during the launch is read through the scalar cache.
//
// Segment barrier: a two-level tree over the segment's tiles (groups of kBarrierGroup): the last arrival of a
// group increments the segment's top counter, and the last arrival at the top publishes the generation on one line
// per group, which that group's tiles poll. Only workgroups of one segment wait for each other; a segment whose
// solve has ended leaves the launch at once. Every wait is bounded: after spinCap polls, or once the abort flag is
// set (by the host or by another workgroup that gave up), the workgroup sets the abort flag and leaves, writing
// nothing; the host then re-runs the solve with kPcgStep launches.
namespace {
constexpr int kScDevice = 16;  // cache-policy operand of the buffer builtins: sc1 (device scope)
typedef uint32_t U3 __attribute__((ext_vector_type(3)));

__device__ inline __amdgpu_buffer_rsrc_t bufferOf(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0x7fffffff, 0x00020000);
}
__device__ inline F3 loadDevice(__amdgpu_buffer_rsrc_t r, uint32_t row) {
  const U3 u = __builtin_amdgcn_raw_buffer_load_b96(r, row * 12u, 0, kScDevice);
  return f3(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z));
}
__device__ inline void storeDevice(__amdgpu_buffer_rsrc_t r, uint32_t row, F3 x) {
  U3 u;
  u.x = __float_as_uint(x.x);
  u.y = __float_as_uint(x.y);
  u.z = __float_as_uint(x.z);
  __builtin_amdgcn_raw_buffer_store_b96(u, r, row * 12u, 0, kScDevice);
}
// A tile's kPcgParts FP64 partial sums (56 consecutive bytes) with three 16-byte and one 8-byte device-scope loads.
typedef uint32_t U4 __attribute__((ext_vector_type(4)));
typedef uint32_t U2 __attribute__((ext_vector_type(2)));
__device__ inline double asDouble(uint32_t lo, uint32_t hi) {
  return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}
__device__ inline void loadPartsDevice(__amdgpu_buffer_rsrc_t r, uint32_t tile, double (&x)[7]) {
  const uint32_t off = tile * (8u * kPartStride);
  const U4 a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kScDevice);
  const U4 b = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16u, 0, kScDevice);
  const U4 c = __builtin_amdgcn_raw_buffer_load_b128(r, off + 32u, 0, kScDevice);
  const U2 d = __builtin_amdgcn_raw_buffer_load_b64(r, off + 48u, 0, kScDevice);
  x[0] = asDouble(a.x, a.y);
  x[1] = asDouble(a.z, a.w);
  x[2] = asDouble(b.x, b.y);
  x[3] = asDouble(b.z, b.w);
  x[4] = asDouble(c.x, c.y);
  x[5] = asDouble(c.z, c.w);
  x[6] = asDouble(d.x, d.y);
}
__device__ inline void storeDevice(double* p, double x) {
  __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t loadDevice(uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t addDevice(uint32_t* p) {
  return __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Arrival of a tile at the segment barrier of iteration `it` (0-based) and the wait for the segment's other tiles.
// Counter lines of a segment with G groups: [0] top, [1 + g] group g's arrivals, [1 + G + g] group g's generation.
// The last arrival of a group adds one to the top counter; the last arrival at the top publishes the generation on
// every group's line, and each tile polls its own group's line (at most kBarrierGroup pollers per line). Called by
// every thread of the workgroup; returns false if the workgroup gave up (abort flag set).
__device__ inline bool segmentBarrier(const SolveCtl& c, uint32_t s, uint32_t localTile, uint32_t nTiles, uint32_t it,
                                      uint32_t* quit) {
  __builtin_amdgcn_s_waitcnt(0);  // this wavefront's device-scope stores are acknowledged
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t* base = c.ctr + 64u * c.segCtr[s];
    const uint32_t g = localTile / kBarrierGroup, nGroups = (nTiles + kBarrierGroup - 1u) / kBarrierGroup;
    const uint32_t gsz = min(kBarrierGroup, nTiles - g * kBarrierGroup);
    const uint32_t want = it + 1u;
    if (addDevice(base + 64u * (1u + g)) == want * gsz - 1u && addDevice(base) == want * nGroups - 1u)
      for (uint32_t k = 0; k < nGroups; ++k)
        __hip_atomic_store(base + 64u * (1u + nGroups + k), want, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t* gen = base + 64u * (1u + nGroups + g);
    uint32_t spins = 0, gaveUp = 0;
    while (loadDevice(gen) < want) {
      if (++spins > c.spinCap || ((spins & 63u) == 0u && loadDevice(c.abortFlag))) {
        gaveUp = 1u;
        __hip_atomic_store(c.abortFlag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *quit = gaveUp;
  }
  __syncthreads();
  return *quit == 0u;
}
}  // namespace

// The window of each exchanged vector: kPcgSolve loads it into registers (3 rows a thread on a sheet) and only the
// rows' search directions go to LDS. kPcgSolvePair, whose two wavefronts per SIMD leave 256 registers a lane, copies
// it to LDS byte for byte, loaded straight into LDS (buffer_load ... lds: 16 bytes a lane, device scope) in 16-byte
// chunks from the chunk that holds the window's first row: row i of the window then starts `delta` floats into the
// copy, plus 3 i.

__device__ inline __amdgpu_buffer_rsrc_t bufferOfBytes(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, int(bytes), 0x00020000);
}
__device__ inline F3 ldsF3(const float* p) { return f3(p[0], p[1], p[2]); }
// 16 bytes a lane from a device-scope buffer at byte `offset` straight into LDS at lds + 16 * lane (lds uniform over
// the wavefront).
__device__ inline void loadChunkToLds(__amdgpu_buffer_rsrc_t r, float* lds, uint32_t offset) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, lds, 16, offset, 0, 0, kScDevice);
}

// Extra: some slice has more than 8 upper or lower slots (seam rows); the slots past the eighth go to an LDS list.
// The arithmetic is the same in every instance.
template <bool Extra>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) kPcgSolve(View v, SolveCtl c) {
  __shared__ double red[kPcgParts * 4], red2[kPcgParts * 4];
  __shared__ F3 hpk[kHaloRows + 1];  // p_k of the window's rows; entry kHaloRows: the dummy column
  __shared__ uint32_t quit;
  // Queued launch (more tiles than the device holds at once): workgroup i takes tile order[i]; each segment's tiles
  // are consecutive in that order, so the XCDs (which receive workgroups round-robin, each in order) hold every
  // segment's tiles spread over them, and the earliest segment not yet finished always has all its tiles resident.
  const uint32_t tile = c.order ? (blockIdx.x < v.nTiles ? c.order[blockIdx.x] : v.nTiles) : xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t slice = tile * kSlicesPerTile + wave;
  const uint32_t row = slice * kSlice + lane;
  if (v.tileDone[0][tile]) {  // the segment's solve has ended (or its Newton loop): pass the flag on
    if (threadIdx.x == 0) v.tileDone[1][tile] = 1u;
    return;
  }
  const uint32_t s = v.tileSeg[tile];
  const uint32_t tileBegin = v.tileRange[2u * tile], tileEnd = v.tileRange[2u * tile + 1u];
  SegState* S = v.state + s;
  const SegParams* P = v.seg + s;
  const uint32_t maxCg = P->maxCg;
  const double cgTolerance = P->cgTolerance;
  const uint32_t k0 = S->kIter[0];
  const uint32_t restarted = S->restartActive;
  const uint32_t kBase = S->kBase;
  double r0 = S->r0;
  // The tile's window: up to kSolveWindows row ranges, concatenated in LDS (range w from entry off[w]). The table
  // lives in LDS (read at uniform addresses): in registers it would crowd out the matrix.
  __shared__ uint32_t wf[kSolveWindows], wc[kSolveWindows], off[kSolveWindows], totals[1];
  if (threadIdx.x == 0) {
    uint32_t entries = 0;
    for (uint32_t w = 0; w < kSolveWindows; ++w) {
      wf[w] = v.tileWindows[2u * (kSolveWindows * tile + w)];
      wc[w] = v.tileWindows[2u * (kSolveWindows * tile + w) + 1u];
      off[w] = entries;
      entries += wc[w];
    }
    totals[0] = entries;
  }
  __syncthreads();
  const uint32_t hc = totals[0];
  const bool oneRange = wc[1] == 0u;  // (most tiles: every column in one range)
  const uint32_t first0 = wf[0];
  // Window entry of a column (the dummy column: the zero entry at kHaloRows), and the range of a window entry.
  const auto entryOf = [&](uint32_t col) {
    uint32_t e = kHaloRows;
    for (uint32_t w = 0; w < kSolveWindows; ++w)
      if (col - wf[w] < wc[w]) e = off[w] + (col - wf[w]);
    return col == v.nRows ? kHaloRows : e;
  };
  const auto rangeOf = [&](uint32_t e) {
    uint32_t r = 0;
    for (uint32_t w = 1; w < kSolveWindows; ++w)
      if (wc[w] && e >= off[w]) r = w;
    return r;
  };

  // The slice's matrix, as kPcgStep<1, 8, true> loads it in its first round trip; columns become window offsets
  // (the dummy column: the zero entry at kHaloRows).
  const uint32_t K = v.sellK[slice], base = v.sellBase[slice], KL = v.lowK[slice], lbase = v.lowBase[slice];
  uint32_t ent[8];  // window entries of slot j's lower (low 16 bits) and upper (high 16 bits) column
  float bt[8][9], bv[8][9];
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    const uint32_t cl = j < KL ? v.lowCol[lbase + j * kSlice + lane] : v.nRows;
    const uint32_t bl = j < KL ? v.lowBlk[lbase + j * kSlice + lane] : v.zeroBlock;
    const uint32_t cu = j < K ? v.sellCol[base / 9u + j * kSlice + lane] : v.nRows;
    ent[j] = entryOf(cl) | (entryOf(cu) << 16);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      bt[j][e] = v.sellVal[size_t(bl) * 9u + e];
      bv[j][e] = v.sellVal[base + j * 9u * kSlice + 9u * lane + e];
    }
  }
  // Extra: the slots past the eighth (seam rows), read once into an LDS list: per lane its real lower slots 8..,
  // then its real upper slots 8.. (xb their blocks as stored, xe their window entries), from entry xAt on. A real
  // slot of a pinned column keeps its block and gathers the zero entry, as kPcgStep does; the slice's padding slots
  // past a row's own count are added as the zero products they are.
  // Each lane's first entry and counts stay in LDS too (xinfo: first | lower << 8 | upper << 12), not in registers
  // the matrix needs.
  __shared__ float xb[Extra ? kSolveExtraCap * 9u : 1u];
  __shared__ uint32_t xe[Extra ? kSolveExtraCap : 1u], xinfo[Extra ? 256 : 1], xn;
  if constexpr (Extra) {
    if (threadIdx.x == 0) xn = 0u;
    __syncthreads();
    const uint32_t kd = v.diagSlot[row], lower = kd - v.rowStart[row], upper = v.rowStart[row + 1u] - kd;
    const uint32_t xl = lower > 8u ? lower - 8u : 0u, xu = upper > 8u ? upper - 8u : 0u;
    const uint32_t xAt = xl + xu ? atomicAdd(&xn, xl + xu) : 0u;
    xinfo[threadIdx.x] = (xAt & 0xffu) | xl << 8 | xu << 12;
    __syncthreads();
    if (xn > kSolveExtraCap) {  // the host admits a matrix only when every tile's list fits: give up, as a wait would
      if (threadIdx.x == 0) __hip_atomic_store(c.abortFlag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    for (uint32_t i = 0; i < xl; ++i) {
      const uint32_t at = lbase + (8u + i) * kSlice + lane;
      xe[xAt + i] = entryOf(v.lowCol[at]);
      const float* b = v.sellVal + size_t(v.lowBlk[at]) * 9u;
      for (uint32_t e = 0; e < 9u; ++e) xb[(xAt + i) * 9u + e] = b[e];
    }
    for (uint32_t i = 0; i < xu; ++i) {
      xe[xAt + xl + i] = entryOf(v.sellCol[base / 9u + (8u + i) * kSlice + lane]);
      const float* b = v.sellVal + base + (8u + i) * 9u * kSlice + 9u * lane;
      for (uint32_t e = 0; e < 9u; ++e) xb[(xAt + xl + i) * 9u + e] = b[e];
    }
  }
  // The rows' private PCG state of iteration -1.
  F3 oR = v.cgR[1][row], oAp = v.cgAp[1][row];
  F3 oZ = v.cgZr[1][row], oT = v.cgT[1][row], oP = v.cgP[1][row];
  F3 oStep = v.step[row];
  float m[kPrecondFloats];
#pragma unroll
  for (uint32_t e = 0; e < kPrecondFloats; ++e) m[e] = v.precond[kPrecondFloats * row + e];
  const uint32_t oFlags = v.rowFlags[row], oCf = v.cflags[row];
  const float4 oCn = v.cnormal[row];
  if (threadIdx.x == 0) hpk[kHaloRows] = f3(0, 0, 0);

  // The vectors' sizes: nRows + 1 entries and 16 zero bytes past them, which cover the last window chunk.
  const uint32_t vecBytes = 12u * (v.nRows + 1u) + 16u;
  const __amdgpu_buffer_rsrc_t zrBuf[2] = {bufferOfBytes(v.cgZr[0], vecBytes), bufferOfBytes(v.cgZr[1], vecBytes)};
  const __amdgpu_buffer_rsrc_t tBuf[2] = {bufferOfBytes(v.cgT[0], vecBytes), bufferOfBytes(v.cgT[1], vecBytes)};
  const __amdgpu_buffer_rsrc_t pBuf[2] = {bufferOfBytes(v.cgP[0], vecBytes), bufferOfBytes(v.cgP[1], vecBytes)};
  const __amdgpu_buffer_rsrc_t partBuf[2] = {bufferOf(v.cgPart[0]), bufferOf(v.cgPart[1])};
  uint32_t updates = 0, conv = 0, k = k0, lastQ = 0;
  double alphaLast = 0.0, rrLast = 0.0;
  bool wroteR0 = false;
  for (uint32_t it = 0;; ++it) {
    const uint32_t q = it & 1u, o = q ^ 1u;
    k = k0 + it;
    lastQ = q;
    // The window's (zr, t, p) and the segment's partial sums of the previous iteration (at most 256 tiles: one a
    // thread). Loads complete in the order they are issued, and the sums are needed first: the window rows past the
    // third of each thread (wide windows only) go before them, the first three rows (every thread, the index clamped
    // to the window) after them, so that the sums wait for no window row.
    constexpr uint32_t HR = kHaloRows / 256u, HR0 = 3u;
    F3 wz[HR], wt[HR], wp[HR];
    const uint32_t hr = (hc + 255u) / 256u;  // window rows per thread (uniform over the workgroup)
    const auto loadRow = [&](uint32_t r) {
      const uint32_t e = min(r * 256u + threadIdx.x, hc - 1u);
      uint32_t i = first0 + e;
      if (!oneRange) {
        const uint32_t w = rangeOf(e);
        i = wf[w] + (e - off[w]);
      }
      wz[r] = loadDevice(o ? zrBuf[1] : zrBuf[0], i);
      wt[r] = loadDevice(o ? tBuf[1] : tBuf[0], i);
      wp[r] = loadDevice(o ? pBuf[1] : pBuf[0], i);
    };
#pragma unroll
    for (uint32_t r = HR0; r < HR; ++r)
      if (r < hr) loadRow(r);
    double sum[kPcgParts] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const uint32_t ti = tileBegin + threadIdx.x;
    if (ti < tileEnd) loadPartsDevice(o ? partBuf[1] : partBuf[0], ti, sum);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (uint32_t r = 0; r < HR0; ++r) loadRow(r);
    blockSumN<kPcgParts, 4, true>(sum, red);

    // The decision of kPcgStep, from the same sums.
    const uint32_t kDone = k + kBase;
    bool update = false;
    double alpha = 0.0, beta = 0.0, rrNext = sum[kPrR];
    if (k == 0) {
      r0 = restarted ? r0 : sqrt(sum[kPrR]);
      wroteR0 = true;
      if (!(sum[kPrR] > 0.0))
        conv = restarted ? kPcgExitTolerance : kPcgExitZero;
      else if (kDone >= maxCg)
        conv = kPcgExitCap;
    } else if (!(sum[kPpAp] > 0.0)) {
      conv = kPcgExitCurvature;
    } else {
      update = true;
      alpha = sum[kPrZ] / sum[kPpAp];
      rrNext = sum[kPrR] - 2.0 * alpha * sum[kPrAp] + alpha * alpha * sum[kPapAp];
      const double rzNext = sum[kPrZ] - 2.0 * alpha * sum[kPrT] + alpha * alpha * sum[kPapT];
      if (sqrt(fmax(rrNext, 0.0)) <= cgTolerance * r0)
        conv = kPcgExitTolerance;
      else if (kDone >= maxCg)
        conv = kPcgExitCap;
      else
        beta = rzNext / sum[kPrZ];
    }
    if (update) {
      ++updates;
      alphaLast = alpha;
      rrLast = rrNext;
    }
    const float a = float(alpha), b = float(beta);
    if (update) oStep = oStep + a * oP;
    if (conv) break;

    // Each window row's search direction once (kPcgStep rebuilds it per gathered column from the same values with
    // the same expression), to LDS.
#pragma unroll
    for (uint32_t r = 0; r < HR; ++r) {
      const uint32_t i = r * 256u + threadIdx.x;
      if (i < hc) hpk[i] = pcgDirection(wz[r], wt[r], wp[r], a, b);
    }
    __syncthreads();

    // Row sums of A p_k in kPcgStep's order: four groups (slot mod 4), each its lower then its upper slots.
    float acc[4][3];
#pragma unroll
    for (uint32_t g = 0; g < 4; ++g) acc[g][0] = acc[g][1] = acc[g][2] = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
      blockTMulAdd(bt[j], hpk[ent[j] & 0xffffu], acc[j % 4][0], acc[j % 4][1], acc[j % 4][2]);
    if constexpr (Extra) {  // lower slots 8 + i of the slice (a slice-uniform bound): the LDS list, or a zero product
      const uint32_t info = xinfo[threadIdx.x], xAt = info & 0xffu, xl = (info >> 8) & 0xfu;
#pragma unroll
      for (uint32_t i = 0; i < 8u; ++i) {
        if (8u + i >= KL) break;
        if (i < xl) {
          blockTMulAdd(&xb[(xAt + i) * 9u], hpk[xe[xAt + i]], acc[i % 4][0], acc[i % 4][1], acc[i % 4][2]);
        } else {
          acc[i % 4][0] += 0.0f;
          acc[i % 4][1] += 0.0f;
          acc[i % 4][2] += 0.0f;
        }
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) blockMulAdd(bv[j], hpk[ent[j] >> 16], acc[j % 4][0], acc[j % 4][1], acc[j % 4][2]);
    if constexpr (Extra) {  // upper slots 8 + i
      const uint32_t info = xinfo[threadIdx.x], xAt = info & 0xffu, xl = (info >> 8) & 0xfu, xu = info >> 12;
#pragma unroll
      for (uint32_t i = 0; i < 8u; ++i) {
        if (8u + i >= K) break;
        if (i < xu) {
          blockMulAdd(&xb[(xAt + xl + i) * 9u], hpk[xe[xAt + xl + i]], acc[i % 4][0], acc[i % 4][1], acc[i % 4][2]);
        } else {
          acc[i % 4][0] += 0.0f;
          acc[i % 4][1] += 0.0f;
          acc[i % 4][2] += 0.0f;
        }
      }
    }
    const float a0 = ((acc[0][0] + acc[1][0]) + acc[2][0]) + acc[3][0];
    const float a1 = ((acc[0][1] + acc[1][1]) + acc[2][1]) + acc[3][1];
    const float a2 = ((acc[0][2] + acc[1][2]) + acc[2][2]) + acc[3][2];

    const F3 rk = oR - a * oAp;
    const F3 pk = pcgDirection(oZ, oT, oP, a, b);
    const F3 apk = (oFlags & kRowPinned) ? f3(0, 0, 0) : onPlane(oCf, oCn, f3(a0, a1, a2));
    const F3 zr = applyPrecond(m, oCf, oCn, rk);
    const F3 tk = applyPrecond(m, oCf, oCn, apk);
    storeDevice(q ? zrBuf[1] : zrBuf[0], row, zr);
    storeDevice(q ? tBuf[1] : tBuf[0], row, tk);
    storeDevice(q ? pBuf[1] : pBuf[0], row, pk);
    double d[kPcgParts] = {dot64(pk, apk), dot64(rk, zr), dot64(rk, tk), dot64(apk, tk),
                           dot64(rk, apk), dot64(apk, apk), dot64(rk, rk)};
    blockSumN<kPcgParts, 4, true>(d, red2);
    if (threadIdx.x < uint32_t(kPcgParts)) {
      double x = d[0];
#pragma unroll
      for (int n = 1; n < kPcgParts; ++n) x = threadIdx.x == uint32_t(n) ? d[n] : x;
      storeDevice(v.cgPart[q] + kPartStride * tile + threadIdx.x, x);
    }
    oR = rk;
    oAp = apk;
    oZ = zr;
    oT = tk;
    oP = pk;
    if (!segmentBarrier(c, s, tile - tileBegin, tileEnd - tileBegin, it, &quit)) return;
  }

  // The solve has ended in every workgroup of the segment at the same iteration: the step, the per-tile flags,
  // and (first tile) the segment's scalars as the sequence of kPcgStep launches leaves them.
  v.step[row] = oStep;
  if (threadIdx.x == 0) {
    v.tileDone[0][tile] = 1u;
    v.tileDone[1][tile] = 1u;
  }
  if (tile == tileBegin && threadIdx.x == 0) {
    if (wroteR0) S->r0 = r0;
    S->kIter[lastQ ^ 1u] = k + 1u;
    if (k > k0) S->kIter[lastQ] = k;
    if (updates) {
      S->alpha = alphaLast;
      S->rr = rrLast;
      S->pcgIter += updates;
      S->cntPcg += updates;
    }
    S->pcgConverged = 1u;
    S->pcgExit = conv;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Persistent solve by tile pairs (kPcgSolvePair): where a scene's tiles outnumber the CUs, one 512-thread workgroup
// per CU takes two consecutive tiles of a segment (waves 0-3 the first, 4-7 the second; the last workgroup of a
// segment with an odd tile count has one), instead of two 256-thread workgroups per CU taking one each. The pair
// stages one window, the union of its tiles' columns (up to kPairWindows row ranges, kPairRows rows), forms p once per
// row of it, reads the segment's partial sums once and meets the segment's barrier once. Per row and per tile it
// computes what kPcgSolve<Extra> computes for a tile, in the same order: the tiles' partial sums stay per tile, each summed
// over its own four wavefronts, and the segment totals are the sums of the same per-tile partials in the same order.

// N sums over one half of a 512-thread workgroup (wavefronts 4 half .. 4 half + 3), in blockSumN's order, returned
// to every thread of the workgroup; every wavefront calls it. lds >= 2 * N * 4 doubles. Fresh: as in blockSumN.
template <int N, bool Fresh = false>
__device__ inline void halfSumN(double (&x)[N], double* lds, uint32_t half) {
  waveTotalN(x);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if constexpr (!Fresh) __syncthreads();
  if (l == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) lds[(w >> 2) * N * 4 + i * 4 + (w & 3)] = x[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) t += lds[half * N * 4 + i * 4 + j];
    x[i] = t;
  }
}

constexpr uint32_t kPairChunks = (kPairRows * 12u + kPairWindows * 24u) / 16u;  // as kHaloChunks

template <bool Extra>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) kPcgSolvePair(View v, SolveCtl c) {
  __shared__ double red[2 * kPcgParts * 4], red2[2 * kPcgParts * 4];
  __shared__ __attribute__((aligned(16))) float hraw[3][kPairChunks * 4];  // zr, t, p copies of the pair's window
  __shared__ F3 hpk[kPairRows + 1];  // p_k of the window's rows; entry kPairRows: the dummy column
  __shared__ uint32_t quit;
  // The workgroup's pair: its first tile and tile count (queued launch: workgroup i takes pair order[i]).
  const uint32_t wg = c.order ? (blockIdx.x < v.nPairs ? c.order[blockIdx.x] : v.nPairs) : xcdMap(blockIdx.x, v.nPairs);
  if (wg >= v.nPairs) return;
  const uint32_t pt = v.pairTiles[wg], first = pt & 0x7fffffffu, count = 1u + (pt >> 31);
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u, half = wave >> 2;
  const bool has = half < count;  // (uniform over each wavefront) this half has a tile; else it mirrors the first
  const uint32_t tile = first + (has ? half : 0u);
  const uint32_t slice = tile * kSlicesPerTile + (wave & 3u);
  const uint32_t row = slice * kSlice + lane;
  if (v.tileDone[0][first]) {  // the segment's solve has ended (or its Newton loop): pass the flags on
    if (threadIdx.x < count) v.tileDone[1][first + threadIdx.x] = 1u;
    return;
  }
  const uint32_t s = v.tileSeg[first];
  const uint32_t tileBegin = v.tileRange[2u * first], tileEnd = v.tileRange[2u * first + 1u];
  const uint32_t localWg = (first - tileBegin) / 2u, nWg = (tileEnd - tileBegin + 1u) / 2u;
  SegState* S = v.state + s;
  const SegParams* P = v.seg + s;
  const uint32_t maxCg = P->maxCg;
  const double cgTolerance = P->cgTolerance;
  const uint32_t k0 = S->kIter[0];
  const uint32_t restarted = S->restartActive;
  const uint32_t kBase = S->kBase;
  double r0 = S->r0;
  // The pair's window: up to kPairWindows row ranges, concatenated (range w from entry off[w], its bytes from chunk
  // ch[w] of the copies, its first row delta[w] floats into that chunk).
  __shared__ uint32_t wf[kPairWindows], wc[kPairWindows], off[kPairWindows], ch[kPairWindows], delta[kPairWindows],
      totals[2];
  if (threadIdx.x == 0) {
    uint32_t entries = 0, chunks = 0;
    for (uint32_t w = 0; w < kPairWindows; ++w) {
      wf[w] = v.pairWindows[2u * (kPairWindows * wg + w)];
      wc[w] = v.pairWindows[2u * (kPairWindows * wg + w) + 1u];
      off[w] = entries;
      ch[w] = chunks;
      delta[w] = (12u * wf[w] - ((12u * wf[w]) & ~15u)) / 4u;
      entries += wc[w];
      chunks += wc[w] ? (12u * wc[w] + 4u * delta[w] + 15u) / 16u : 0u;
    }
    totals[0] = entries;
    totals[1] = chunks;
  }
  __syncthreads();
  const uint32_t hc = totals[0];
  const bool oneRange = wc[1] == 0u;
  const auto entryOf = [&](uint32_t col) {
    uint32_t e = kPairRows;
    for (uint32_t w = 0; w < kPairWindows; ++w)
      if (col - wf[w] < wc[w]) e = off[w] + (col - wf[w]);
    return col == v.nRows ? kPairRows : e;
  };
  const auto rangeOf = [&](uint32_t e) {
    uint32_t r = 0;
    for (uint32_t w = 1; w < kPairWindows; ++w)
      if (wc[w] && e >= off[w]) r = w;
    return r;
  };
  const uint32_t ownEntry = entryOf(row);
  const uint32_t ownRange = rangeOf(ownEntry);
  const uint32_t own = 4u * ch[ownRange] + delta[ownRange] + 3u * (ownEntry - off[ownRange]);

  const uint32_t K = v.sellK[slice], base = v.sellBase[slice], KL = v.lowK[slice], lbase = v.lowBase[slice];
  uint32_t ent[8];
  float bt[8][9], bv[8][9];
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    const uint32_t cl = j < KL ? v.lowCol[lbase + j * kSlice + lane] : v.nRows;
    const uint32_t bl = j < KL ? v.lowBlk[lbase + j * kSlice + lane] : v.zeroBlock;
    const uint32_t cu = j < K ? v.sellCol[base / 9u + j * kSlice + lane] : v.nRows;
    ent[j] = entryOf(cl) | (entryOf(cu) << 16);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      bt[j][e] = v.sellVal[size_t(bl) * 9u + e];
      bv[j][e] = v.sellVal[base + j * 9u * kSlice + 9u * lane + e];
    }
  }
  // Extra: the slots past the eighth in the pair's LDS list, as in kPcgSolve (2 kSolveExtraCap entries a pair).
  constexpr uint32_t kCap = 2u * kSolveExtraCap;
  __shared__ float xb[Extra ? kCap * 9u : 1u];
  __shared__ uint32_t xe[Extra ? kCap : 1u], xinfo[Extra ? 512 : 1], xn;
  if constexpr (Extra) {
    if (threadIdx.x == 0) xn = 0u;
    __syncthreads();
    const uint32_t kd = v.diagSlot[row], lower = kd - v.rowStart[row], upper = v.rowStart[row + 1u] - kd;
    const uint32_t xl = has && lower > 8u ? lower - 8u : 0u, xu = has && upper > 8u ? upper - 8u : 0u;
    const uint32_t xAt = xl + xu ? atomicAdd(&xn, xl + xu) : 0u;
    xinfo[threadIdx.x] = (xAt & 0xffu) | xl << 8 | xu << 12;
    __syncthreads();
    if (xn > kCap) {  // the host admits a matrix only when every pair's list fits: give up, as a wait would
      if (threadIdx.x == 0) __hip_atomic_store(c.abortFlag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    for (uint32_t i = 0; i < xl; ++i) {
      const uint32_t at = lbase + (8u + i) * kSlice + lane;
      xe[xAt + i] = entryOf(v.lowCol[at]);
      const float* b = v.sellVal + size_t(v.lowBlk[at]) * 9u;
      for (uint32_t e = 0; e < 9u; ++e) xb[(xAt + i) * 9u + e] = b[e];
    }
    for (uint32_t i = 0; i < xu; ++i) {
      xe[xAt + xl + i] = entryOf(v.sellCol[base / 9u + (8u + i) * kSlice + lane]);
      const float* b = v.sellVal + base + (8u + i) * 9u * kSlice + 9u * lane;
      for (uint32_t e = 0; e < 9u; ++e) xb[(xAt + xl + i) * 9u + e] = b[e];
    }
  }
  F3 oR = v.cgR[1][row], oAp = v.cgAp[1][row], oP = f3(0, 0, 0);
  F3 oStep = v.step[row];
  float m[kPrecondFloats];
#pragma unroll
  for (uint32_t e = 0; e < kPrecondFloats; ++e) m[e] = v.precond[kPrecondFloats * row + e];
  const uint32_t oFlags = v.rowFlags[row], oCf = v.cflags[row];
  const float4 oCn = v.cnormal[row];
  if (threadIdx.x == 0) hpk[kPairRows] = f3(0, 0, 0);

  const uint32_t vecBytes = 12u * (v.nRows + 1u) + 16u;
  const __amdgpu_buffer_rsrc_t zrBuf[2] = {bufferOfBytes(v.cgZr[0], vecBytes), bufferOfBytes(v.cgZr[1], vecBytes)};
  const __amdgpu_buffer_rsrc_t tBuf[2] = {bufferOfBytes(v.cgT[0], vecBytes), bufferOfBytes(v.cgT[1], vecBytes)};
  const __amdgpu_buffer_rsrc_t pBuf[2] = {bufferOfBytes(v.cgP[0], vecBytes), bufferOfBytes(v.cgP[1], vecBytes)};
  const __amdgpu_buffer_rsrc_t partBuf[2] = {bufferOf(v.cgPart[0]), bufferOf(v.cgPart[1])};
  uint32_t updates = 0, conv = 0, k = k0, lastQ = 0;
  double alphaLast = 0.0, rrLast = 0.0;
  bool wroteR0 = false;
  for (uint32_t it = 0;; ++it) {
    const uint32_t q = it & 1u, o = q ^ 1u;
    k = k0 + it;
    lastQ = q;
    // The first half's wavefronts load the segment's partial sums (the 256-thread order) and reduce them; meanwhile
    // the second half's copy the window's (zr, t, p) straight into LDS, needed only after the decision. (With the
    // extra-slot list, Extra, the split costs spilled registers and measured slower: there every wavefront copies its
    // share of the window, the first half's also load the sums, and the window is waited for before the reduction.)
    double sum[kPcgParts] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if constexpr (Extra) {
      for (uint32_t w = 0; w < kPairWindows; ++w) {
        const uint32_t firstByte = (12u * wf[w]) & ~15u;
        const uint32_t n = wc[w] ? (12u * wc[w] + 4u * delta[w] + 15u) / 16u : 0u;
        for (uint32_t j0 = wave * 64u; j0 < n; j0 += 512u) {  // uniform over the wavefront
          const uint32_t j = j0 + lane;
          if (j < n) {
            loadChunkToLds(o ? zrBuf[1] : zrBuf[0], &hraw[0][4u * (ch[w] + j0)], firstByte + 16u * j);
            loadChunkToLds(o ? tBuf[1] : tBuf[0], &hraw[1][4u * (ch[w] + j0)], firstByte + 16u * j);
            loadChunkToLds(o ? pBuf[1] : pBuf[0], &hraw[2][4u * (ch[w] + j0)], firstByte + 16u * j);
          }
        }
      }
      if (half == 0u) {
        uint32_t ti = tileBegin + threadIdx.x;
        if (ti < tileEnd) loadPartsDevice(o ? partBuf[1] : partBuf[0], ti, sum);
        for (ti += 256u; ti < tileEnd; ti += 256u) {
          double more[kPcgParts];
          loadPartsDevice(o ? partBuf[1] : partBuf[0], ti, more);
#pragma unroll
          for (int n = 0; n < kPcgParts; ++n) sum[n] += more[n];
        }
      }
      __builtin_amdgcn_s_waitcnt(0);  // this wavefront's chunks are in LDS; the barrier below shows them
      halfSumN<kPcgParts, true>(sum, red, 0u);  // the first half's threads hold the 256-thread order's sums
    } else if (half == 0u) {
      uint32_t ti = tileBegin + threadIdx.x;
      if (ti < tileEnd) loadPartsDevice(o ? partBuf[1] : partBuf[0], ti, sum);
      for (ti += 256u; ti < tileEnd; ti += 256u) {
        double more[kPcgParts];
        loadPartsDevice(o ? partBuf[1] : partBuf[0], ti, more);
#pragma unroll
        for (int n = 0; n < kPcgParts; ++n) sum[n] += more[n];
      }
      waveTotalN(sum);
      if (lane == 0)
#pragma unroll
        for (int i = 0; i < kPcgParts; ++i) red[i * 4 + wave] = sum[i];
    } else {
      for (uint32_t w = 0; w < kPairWindows; ++w) {
        const uint32_t firstByte = (12u * wf[w]) & ~15u;
        const uint32_t n = wc[w] ? (12u * wc[w] + 4u * delta[w] + 15u) / 16u : 0u;
        for (uint32_t j0 = (wave - 4u) * 64u; j0 < n; j0 += 256u) {  // uniform over the wavefront
          const uint32_t j = j0 + lane;
          if (j < n) {
            loadChunkToLds(o ? zrBuf[1] : zrBuf[0], &hraw[0][4u * (ch[w] + j0)], firstByte + 16u * j);
            loadChunkToLds(o ? tBuf[1] : tBuf[0], &hraw[1][4u * (ch[w] + j0)], firstByte + 16u * j);
            loadChunkToLds(o ? pBuf[1] : pBuf[0], &hraw[2][4u * (ch[w] + j0)], firstByte + 16u * j);
          }
        }
      }
    }
    if constexpr (!Extra) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kPcgParts; ++i) {  // blockSumN's order over the first half's four wavefronts
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) t += red[i * 4 + j];
        sum[i] = t;
      }
    }

    const uint32_t kDone = k + kBase;
    bool update = false;
    double alpha = 0.0, beta = 0.0, rrNext = sum[kPrR];
    if (k == 0) {
      r0 = restarted ? r0 : sqrt(sum[kPrR]);
      wroteR0 = true;
      if (!(sum[kPrR] > 0.0))
        conv = restarted ? kPcgExitTolerance : kPcgExitZero;
      else if (kDone >= maxCg)
        conv = kPcgExitCap;
    } else if (!(sum[kPpAp] > 0.0)) {
      conv = kPcgExitCurvature;
    } else {
      update = true;
      alpha = sum[kPrZ] / sum[kPpAp];
      rrNext = sum[kPrR] - 2.0 * alpha * sum[kPrAp] + alpha * alpha * sum[kPapAp];
      const double rzNext = sum[kPrZ] - 2.0 * alpha * sum[kPrT] + alpha * alpha * sum[kPapT];
      if (sqrt(fmax(rrNext, 0.0)) <= cgTolerance * r0)
        conv = kPcgExitTolerance;
      else if (kDone >= maxCg)
        conv = kPcgExitCap;
      else
        beta = rzNext / sum[kPrZ];
    }
    if (update) {
      ++updates;
      alphaLast = alpha;
      rrLast = rrNext;
    }
    const float a = float(alpha), b = float(beta);
    if constexpr (!Extra) {
      __builtin_amdgcn_s_waitcnt(0);  // (second half) this wavefront's chunks are in LDS; the barrier shows them
      __syncthreads();
    }
    oP = ldsF3(&hraw[2][own]);
    if (update) oStep = oStep + a * oP;
    if (conv) break;

    for (uint32_t i = threadIdx.x; i < hc; i += 512u) {
      uint32_t at = delta[0] + 3u * i;
      if (!oneRange) {
        const uint32_t w = rangeOf(i);
        at = 4u * ch[w] + delta[w] + 3u * (i - off[w]);
      }
      hpk[i] = pcgDirection(ldsF3(&hraw[0][at]), ldsF3(&hraw[1][at]), ldsF3(&hraw[2][at]), a, b);
    }
    __syncthreads();

    float acc[4][3];
#pragma unroll
    for (uint32_t g = 0; g < 4; ++g) acc[g][0] = acc[g][1] = acc[g][2] = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
      blockTMulAdd(bt[j], hpk[ent[j] & 0xffffu], acc[j % 4][0], acc[j % 4][1], acc[j % 4][2]);
    if constexpr (Extra) {
      const uint32_t info = xinfo[threadIdx.x], xAt = info & 0xffu, xl = (info >> 8) & 0xfu;
#pragma unroll
      for (uint32_t i = 0; i < 8u; ++i) {
        if (8u + i >= KL) break;
        if (i < xl) {
          blockTMulAdd(&xb[(xAt + i) * 9u], hpk[xe[xAt + i]], acc[i % 4][0], acc[i % 4][1], acc[i % 4][2]);
        } else {
          acc[i % 4][0] += 0.0f;
          acc[i % 4][1] += 0.0f;
          acc[i % 4][2] += 0.0f;
        }
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) blockMulAdd(bv[j], hpk[ent[j] >> 16], acc[j % 4][0], acc[j % 4][1], acc[j % 4][2]);
    if constexpr (Extra) {
      const uint32_t info = xinfo[threadIdx.x], xAt = info & 0xffu, xl = (info >> 8) & 0xfu, xu = info >> 12;
#pragma unroll
      for (uint32_t i = 0; i < 8u; ++i) {
        if (8u + i >= K) break;
        if (i < xu) {
          blockMulAdd(&xb[(xAt + xl + i) * 9u], hpk[xe[xAt + xl + i]], acc[i % 4][0], acc[i % 4][1], acc[i % 4][2]);
        } else {
          acc[i % 4][0] += 0.0f;
          acc[i % 4][1] += 0.0f;
          acc[i % 4][2] += 0.0f;
        }
      }
    }
    const float a0 = ((acc[0][0] + acc[1][0]) + acc[2][0]) + acc[3][0];
    const float a1 = ((acc[0][1] + acc[1][1]) + acc[2][1]) + acc[3][1];
    const float a2 = ((acc[0][2] + acc[1][2]) + acc[2][2]) + acc[3][2];

    const F3 rk = oR - a * oAp;
    const F3 pk = hpk[ownEntry];
    const F3 apk = (oFlags & kRowPinned) ? f3(0, 0, 0) : onPlane(oCf, oCn, f3(a0, a1, a2));
    const F3 zr = applyPrecond(m, oCf, oCn, rk);
    const F3 tk = applyPrecond(m, oCf, oCn, apk);
    if (has) {
      storeDevice(q ? zrBuf[1] : zrBuf[0], row, zr);
      storeDevice(q ? tBuf[1] : tBuf[0], row, tk);
      storeDevice(q ? pBuf[1] : pBuf[0], row, pk);
    }
    double d[kPcgParts] = {dot64(pk, apk), dot64(rk, zr), dot64(rk, tk), dot64(apk, tk),
                           dot64(rk, apk), dot64(apk, apk), dot64(rk, rk)};
    halfSumN<kPcgParts, true>(d, red2, half);
    const uint32_t t256 = threadIdx.x & 255u;
    if (has && t256 < uint32_t(kPcgParts)) {
      double x = d[0];
#pragma unroll
      for (int n = 1; n < kPcgParts; ++n) x = t256 == uint32_t(n) ? d[n] : x;
      storeDevice(v.cgPart[q] + kPartStride * tile + t256, x);
    }
    oR = rk;
    oAp = apk;
    if (!segmentBarrier(c, s, localWg, nWg, it, &quit)) return;
  }

  if (has) v.step[row] = oStep;
  if (threadIdx.x < count) {
    v.tileDone[0][first + threadIdx.x] = 1u;
    v.tileDone[1][first + threadIdx.x] = 1u;
  }
  if (first == tileBegin && threadIdx.x == 0) {
    if (wroteR0) S->r0 = r0;
    S->kIter[lastQ ^ 1u] = k + 1u;
    if (k > k0) S->kIter[lastQ] = k;
    if (updates) {
      S->alpha = alphaLast;
      S->rr = rrLast;
      S->pcgIter += updates;
      S->cntPcg += updates;
    }
    S->pcgConverged = 1u;
    S->pcgExit = conv;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// FP64 true-residual refinement of the FP32 PCG (mixed-precision iterative refinement). The FP32 recursion for r
// drifts from the true residual b - A step (b = S(-grad)) as the iteration proceeds, so a solve the recursion calls
// converged can be short of |b - A step| <= tol |r0| (the reference's FP64 solve meets its test on a residual that
// stays close to the true one, cpu_solver.cpp:188-204). Per row: r = S(-grad - A step) from an FP64 product, and the
// continuation's first state written as the output of a kernel "-1" (r, zr = S M r, p = ap = t = 0; step kept);
// per tile: |r|^2 and the partial sums the next kPcgStep reads, and in partC the rows' sum of (A_ij step_j)^2 over
// every product term, from which kRefineDecide estimates the part of |r| that rounding step to FP32 leaves.
__global__ void __launch_bounds__(256) kTrueResidual(View v, const uint32_t* abortFlag) {
  __shared__ double lds[kPcgParts * 4];
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles || (abortFlag && *abortFlag)) return;
  const uint32_t s = v.tileSeg[tile];
  const SegState* S = v.state + s;
  if (S->newtonDone || S->pcgExit != kPcgExitTolerance) return;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t slice = tile * kSlicesPerTile + wave;
  const uint32_t row = slice * kSlice + lane;
  const F3* __restrict__ stp = v.step;
  double sq = 0.0;
  const D3 as = sellRowProduct64<true>(v, slice, lane, [&](uint32_t c) { return stp[c]; }, &sq);
  const F3 g = v.grad[row];
  const F3 resid = f3(float(-double(g.x) - as.x), float(-double(g.y) - as.y), float(-double(g.z) - as.z));
  const uint32_t cf = v.cflags[row];
  const float4 plane = v.cnormal[row];
  const bool pinned = v.rowFlags[row] & kRowPinned;
  const F3 r = pinned ? f3(0, 0, 0) : onPlane(cf, plane, resid);
  const F3 zr = applyPrecond(v.precond + kPrecondFloats * row, cf, plane, r);
  const F3 zero = f3(0, 0, 0);
  v.cgR[1][row] = r;
  v.cgAp[1][row] = zero;
  v.cgP[1][row] = zero;
  v.cgZr[1][row] = zr;
  v.cgT[1][row] = zero;
  double d[kPcgParts] = {0.0, dot64(r, zr), 0.0, 0.0, 0.0, 0.0, dot64(r, r)};
  double f[1] = {pinned ? 0.0 : sq};
  blockSumN<kPcgParts, 4>(d, lds);
  blockSumN<1, 4>(f, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kPcgParts; ++i) v.cgPart[1][kPartStride * tile + i] = d[i];
    v.partB[tile] = d[kPrR];
    v.partC[tile] = f[0];
  }
}

// A continuation pays when the recursion has drifted well away from the true residual, not for a true residual a
// few percent above the tolerance (FP32 recursions end there routinely): it starts above kRefineFactor tol r0.
// Nor does it pay once the true residual is down to what rounding step to FP32 leaves: step is stored in FP32, and
// each component's rounding error, uniform within half a unit in the last place (2^-24 |step_j| relative), puts
// about 2^-24 sqrt(sum (A_ij step_j)^2 / 3) into |b - A step| however exact the solve (the sum over every product
// term of every row). Continuations go on while the true residual is above kRefineFactor times that floor as well;
// below it, a continuation moves step by a fraction of its FP32 rounding.
constexpr double kRefineFactor = 4.0;

// Per segment: a solve that stopped on the recursion's tolerance test continues from its true residual when that
// is above kRefineFactor max(tol r0, the FP32 rounding floor) and iterations remain (at most `maxRestarts`
// continuations per solve). Every tile of the segment gets its solve-ended flag for the next kPcgStep; the host
// looks at pcgConverged afterwards.
__global__ void __launch_bounds__(256) kRefineDecide(View v, uint32_t maxRestarts, const uint32_t* abortFlag) {
  __shared__ double lds[8];
  const uint32_t s = blockIdx.x;
  SegState* S = v.state + s;
  const SegParams& P = v.seg[s];
  const bool aborted = abortFlag && *abortFlag;  // the persistent solve before gave up: no continuation
  const bool candidate = !aborted && !S->newtonDone && S->pcgExit == kPcgExitTolerance;
  const double trueSq = candidate ? rangeSum256(v.partB, P.tileBegin, P.tileEnd, lds) : 0.0;
  const double termSq = candidate ? rangeSum256(v.partC, P.tileBegin, P.tileEnd, lds) : 0.0;
  const double floor32 = 0x1p-24 * sqrt(termSq / 3.0);
  const bool restart = candidate && sqrt(trueSq) > kRefineFactor * fmax(P.cgTolerance * S->r0, floor32) &&
                       S->pcgIter < P.maxCg && S->restartsThisSolve < maxRestarts;
  for (uint32_t t = P.tileBegin + threadIdx.x; t < P.tileEnd; t += 256u) v.tileDone[0][t] = restart ? 0u : 1u;
  if (threadIdx.x == 0) {
    S->restartActive = restart ? 1u : 0u;
    if (restart) {
      S->pcgConverged = 0u;
      S->kIter[0] = 0u;
      S->kBase = S->pcgIter;
      S->restartsThisSolve += 1u;
      S->cntRestarts += 1u;
    }
  }
}

// After the solve, from one product A step over every row: each held particle's normal force from the unfiltered
// residual -grad - A step (cpu_solver.cpp:210-212: the reference keeps that residual as a running vector; here it is
// formed explicitly), the FP64 norm partials of the filtered true residual S(-grad - A step) and of its FP32 rounding
// floor's terms (pinned rows excluded) for the solve report, and the Newton slope grad . step partials.
__global__ void __launch_bounds__(256) kContactForceSlope(View v) {
  __shared__ double lds[8];
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  if (v.state[s].newtonDone) return;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t slice = tile * kSlicesPerTile + wave;
  const uint32_t row = slice * kSlice + lane;
  const uint32_t cf = v.cflags[row];
  const F3 gi = v.grad[row];
  const F3 si = v.step[row];
  const F3* __restrict__ stp = v.step;
  double sq = 0.0;
  const D3 as = sellRowProduct64<true>(v, slice, lane, [&](uint32_t c) { return stp[c]; }, &sq);
  const F3 residual = f3(float(-double(gi.x) - as.x), float(-double(gi.y) - as.y), float(-double(gi.z) - as.z));
  const float4 plane = v.cnormal[row];
  v.cforce[row] = (cf & kContactActive) ? fmaxf(0.0f, -dot(residual, f3(plane.x, plane.y, plane.z))) : 0.0f;
  const bool pinned = v.rowFlags[row] & kRowPinned;
  const F3 filtered = pinned ? f3(0, 0, 0) : onPlane(cf, plane, residual);
  const double slope = blockSum256(dot64(gi, si), lds);
  const double trueSq = blockSum256(dot64(filtered, filtered), lds);
  const double termSq = blockSum256(pinned ? 0.0 : sq, lds);
  if (threadIdx.x == 0) {
    v.partA[tile] = slope;
    v.partB[tile] = trueSq;
    v.partC[tile] = termSq;
  }
}

// slope = grad . step; no line search when it is not negative (cpu_solver.cpp:267-268). Also closes the solve's
// report: its exit reason and its true relative residual.
__global__ void __launch_bounds__(256) kNewtonSlope(View v) {
  __shared__ double lds[8];
  const uint32_t s = blockIdx.x;
  SegState* S = v.state + s;
  if (S->newtonDone) return;
  const SegParams& P = v.seg[s];
  const double slope = rangeSum256(v.partA, P.tileBegin, P.tileEnd, lds);
  const double trueSq = rangeSum256(v.partB, P.tileBegin, P.tileEnd, lds);
  const double termSq = rangeSum256(v.partC, P.tileBegin, P.tileEnd, lds);
  if (threadIdx.x == 0) {
    S->slope = slope;
    const double rel = S->r0 > 0.0 ? sqrt(trueSq) / S->r0 : 0.0;
    const double floorRel = S->r0 > 0.0 ? 0x1p-24 * sqrt(termSq / 3.0) / S->r0 : 0.0;
    S->cntSolves += 1u;
    S->cntExitTolerance += S->pcgExit == kPcgExitTolerance ? 1u : 0u;
    S->cntExitCap += S->pcgExit == kPcgExitCap ? 1u : 0u;
    S->cntExitCurvature += S->pcgExit == kPcgExitCurvature ? 1u : 0u;
    S->cntExitZero += S->pcgExit == kPcgExitZero ? 1u : 0u;
    S->trueRelLast = rel;
    S->trueRelSum += rel;
    S->trueRelMax = fmax(S->trueRelMax, rel);
    if (S->pcgExit == kPcgExitTolerance) {
      S->trueRelMaxConverged = fmax(S->trueRelMaxConverged, rel);
      S->trueExcessMax = fmax(S->trueExcessMax, rel / fmax(P.cgTolerance, floorRel));
    }
    if (!(slope < 0.0)) {
      S->newtonDone = 1u;
    } else {
      S->lsAlpha = 1.0;
      S->lsDone = 0u;
      S->lsTrial = 0u;
    }
  }
}

// Line-search trial (cpu_solver.cpp:269-274): the incremental potential at u + alpha step, energy only. Each term
// is evaluated once, by its local vertex 0's row, at the same relative placement the assembly uses.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) kEnergyTrial(View v) {
  __shared__ double lds[8];
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  const SegState* S = v.state + s;
  if (S->newtonDone || S->lsDone) return;
  const SegParams& P = v.seg[s];
  const float alpha = float(S->lsAlpha);
  const uint32_t row = tile * kTile + threadIdx.x;
  const uint32_t flags = v.rowFlags[row];
  double rowEnergy = 0.0;
  if (!(flags & kRowPad)) {
    F3 g = f3(0, 0, 0);
    Blk unused;
    const F3 disp = v.u[row] + alpha * v.step[row];
    rowEnergy += double(ownEnergy<false>(v, P, row, flags, disp, g, unused));
    const uint32_t end = v.incStart[row + 1];
    for (uint32_t i = v.incStart[row]; i < end; i += kIncBatch) {
      IncTerm batch[kIncBatch];
      loadBatch<true>(v, i, end, alpha, true, batch);
#pragma unroll
      for (uint32_t j = 0; j < kIncBatch; ++j)
        rowEnergy += double(evalEntry<false>(v, P, batch[j], g, nullptr, 0u, 0u));
    }
  }
  const double total = blockSum256(rowEnergy, lds);
  if (threadIdx.x == 0) v.partC[tile] = total;
}

// Armijo test E(trial) <= e0 + 1e-4 alpha slope; else halve, at most 20 trials (cpu_solver.cpp:269-274).
__global__ void __launch_bounds__(256) kLineSearchDecide(View v) {
  __shared__ double lds[8];
  const uint32_t s = blockIdx.x;
  SegState* S = v.state + s;
  if (S->newtonDone || S->lsDone) return;
  const SegParams& P = v.seg[s];
  const double e = rangeSum256(v.partC, P.tileBegin, P.tileEnd, lds);
  if (threadIdx.x == 0) {
    const double alpha = S->lsAlpha;
    S->eTrial = e;
    S->cntLs += 1u;
    if (e <= S->e0 + 1e-4 * alpha * S->slope) {
      S->lsDone = 1u;
    } else {
      S->lsAlpha = alpha * 0.5;
      S->lsTrial += 1u;
      if (S->lsTrial >= 20u) S->lsDone = 1u;
    }
  }
}

// Accepts the line search's step (cpu_solver.cpp:275-282): u += alpha step, the largest component of the move
// into the tile's partial, and the friction weight of the next Newton iteration from this solve's normal force
// (zero for a particle no longer held). Acts once per Newton iteration, on a segment whose line search has
// decided (lsDone == 1) and whose step is not yet taken (kNewtonEnd sets lsDone = 2), so a captured sequence may
// launch it behind a fixed number of trials and again after further trials.
__global__ void __launch_bounds__(256) kApply(View v) {
  __shared__ float lds[8];
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  const SegState* S = v.state + s;
  if (S->newtonDone || S->lsDone != 1u) return;
  const uint32_t row = tile * kTile + threadIdx.x;
  const F3 move = float(S->lsAlpha) * v.step[row];
  v.u[row] += move;
  v.clambda[row] = (v.cflags[row] & kContactActive) ? v.cforce[row] : 0.0f;
  const float largest = blockMax256(fmaxf(fmaxf(fabsf(move.x), fabsf(move.y)), fabsf(move.z)), lds);
  if (threadIdx.x == 0) v.partMax[tile] = largest;
}

// Per segment: the Newton loop ends once no particle moved by 1e-9 m (cpu_solver.cpp:283); the step is marked
// taken.
__global__ void __launch_bounds__(256) kNewtonEnd(View v) {
  __shared__ float lds[8];
  const uint32_t s = blockIdx.x;
  SegState* S = v.state + s;
  if (S->newtonDone || S->lsDone != 1u) return;
  const SegParams& P = v.seg[s];
  const float largest = rangeMax256(v.partMax, P.tileBegin, P.tileEnd, lds);
  if (threadIdx.x == 0) {
    S->maxMove = largest;
    S->lsDone = 2u;
    S->newtonDone = double(largest) < 1e-9 ? 1u : 0u;
  }
}

// Substep end (cpu_solver.cpp:285-310): x = xStart + u. A free particle that still ends deeper in the body than
// the snap tolerance is moved out (stopped where its path met the body when the field gives no direction there);
// pins hold; the velocity is the substep's displacement over h, damped by max(0, 1 - c h) with c the fabric's
// damping rate (plus the assembly damping while the garment assembles). The velocity change the substep realised
// (u / h - v) is kept as a feature of the learned warm start.
__global__ void __launch_bounds__(256) kFinish(View v) {
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const SegParams& P = v.seg[v.tileSeg[tile]];
  const uint32_t row = tile * kTile + threadIdx.x;
  const uint32_t flags = v.rowFlags[row];
  if (flags & kRowPad) return;
  const Fabric32 fab = v.fabrics[v.rowFabric[row]];
  const F3 from = v.xStart[row];
  const F3 disp = v.u[row];
  F3 end = from + disp;
  if (flags & kRowPinned) {
    end = v.pinPos[row];
  } else if (P.field.present) {
    const float radius = fabricRadius(P, fab);
    const float dist = bodyDistance(v.fieldValues, P.field, end);
    if (dist < radius - P.contactSnap) {
      const F3 n = bodyNormal(v.fieldValues, P.field, end);
      if (allZero(n))
        end = stopAtBody(v.fieldValues, P.field, radius, from, end);
      else
        pushOut(dist, n, radius, fab.friction, from, end);
    }
  }
  const float rate = P.dampingScale * fab.damping +
                     (v.garments[v.rowGarment[row]].phase == 0u ? P.assemblyExtraDamping : 0.0f);
  const float invH = 1.0f / P.h;
  v.dvLast[row] = (flags & kRowPinned) ? f3(0, 0, 0) : invH * disp - v.v[row];
  v.x[row] = end;
  v.v[row] = (invH * fmaxf(0.0f, 1.0f - rate * P.h)) * (end - from);
}

// ---------------------------------------------------------------------------------------------------------------
// Stats on the device (SimModel::stats, sim_model.cpp:222-241): per-tile partials, then one workgroup per segment
// sums its tiles in a fixed order. Only the per-segment results leave the device.

// The body distance in FP64 (CollisionField::sample, collision_field.cpp:8-27) from the grid's FP64 origin and
// cell, for the settle-loop stats.
__device__ inline double bodyDistance64(const float* values, const SegParams& P, F3 p) {
  const double inv = 1.0 / P.fieldCell;
  return gridSample<double>(values, P.field, (double(p.x) - P.fieldOrigin[0]) * inv,
                            (double(p.y) - P.fieldOrigin[1]) * inv, (double(p.z) - P.fieldOrigin[2]) * inv,
                            P.fieldBand);
}

__device__ inline bool finite3(F3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

namespace {
// Per tile: kinetic energy, sum of v^2 and moving count (sums); non-finite flag, penetration and seam gap (maxima).
constexpr int kStatSums = 3, kStatMaxes = 3, kStatParts = kStatSums + kStatMaxes;
}  // namespace

template <int N>
__device__ inline void blockMaxN256(double (&x)[N], double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = fmax(x[i], __shfl_down(x[i], off, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) lds[i * 4 + w] = x[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = fmax(fmax(lds[i * 4], lds[i * 4 + 1]), fmax(lds[i * 4 + 2], lds[i * 4 + 3]));
}

__global__ void __launch_bounds__(256) kStatsTile(View v, const double* massF64, double* part) {
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  __shared__ double lds[kStatParts * 4];
  const SegParams& P = v.seg[v.tileSeg[tile]];
  const uint32_t row = tile * kTile + threadIdx.x;
  const uint32_t flags = v.rowFlags[row];
  double sum[kStatSums] = {0.0, 0.0, 0.0};
  double mx[kStatMaxes] = {0.0, 0.0, 0.0};
  if (!(flags & kRowPad)) {
    const F3 x = v.x[row], vel = v.v[row];
    if (!finite3(x) || !finite3(vel)) mx[0] = 1.0;
    if (!(flags & kRowPinned)) {
      const double v2 = double(vel.x) * vel.x + double(vel.y) * vel.y + double(vel.z) * vel.z;
      sum[0] = 0.5 * massF64[row] * v2;
      sum[1] = v2;
      sum[2] = 1.0;
    }
    if (P.field.present) {
      const double pen = -bodyDistance64(v.fieldValues, P, x);
      if (pen > mx[1]) mx[1] = pen;
    }
    // Seams whose first vertex is this row: seam entries end the row's sorted incidence list.
    for (uint32_t k = v.incStart[row + 1]; k > v.incStart[row]; --k) {
      const uint32_t e = v.inc[k - 1];
      if ((e >> 30) != kKindSeam) break;
      if ((e & 3u) != 0u) continue;
      const Seam32 sp = v.seams[(e >> 2) & 0x0FFFFFFFu];
      const F3 d = v.x[sp.a] - v.x[sp.b];
      const float gap = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(d.x, d.x), __fmul_rn(d.y, d.y)), __fmul_rn(d.z, d.z)));
      if (double(gap) > mx[2]) mx[2] = gap;  // a NaN gap never wins, as with std::max
    }
  }
  blockSumN<kStatSums, 4>(sum, lds);
  blockMaxN256<kStatMaxes>(mx, lds);
  if (threadIdx.x == 0) {
    double* o = part + size_t(tile) * kStatParts;
    for (int i = 0; i < kStatSums; ++i) o[i] = sum[i];
    for (int i = 0; i < kStatMaxes; ++i) o[kStatSums + i] = mx[i];
  }
}

// One workgroup per segment: its tiles' partials in tile order -> out[segment * 6].
__global__ void __launch_bounds__(256) kStatsSegment(View v, const double* part, double* out) {
  const uint32_t s = blockIdx.x;
  const SegParams& P = v.seg[s];
  __shared__ double lds[kStatParts * 4];
  double sum[kStatSums] = {0.0, 0.0, 0.0};
  double mx[kStatMaxes] = {0.0, 0.0, 0.0};
  for (uint32_t t = P.tileBegin + threadIdx.x; t < P.tileEnd; t += 256u) {
    const double* o = part + size_t(t) * kStatParts;
    for (int i = 0; i < kStatSums; ++i) sum[i] += o[i];
    for (int i = 0; i < kStatMaxes; ++i) mx[i] = fmax(mx[i], o[kStatSums + i]);
  }
  blockSumN<kStatSums, 4>(sum, lds);
  blockMaxN256<kStatMaxes>(mx, lds);
  if (threadIdx.x == 0) {
    double* o = out + size_t(s) * kStatParts;
    for (int i = 0; i < kStatSums; ++i) o[i] = sum[i];
    for (int i = 0; i < kStatMaxes; ++i) o[kStatSums + i] = mx[i];
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Learned warm start support.

// SDF distance and normal at the current position (distance = band and zero normal without a body).
__global__ void __launch_bounds__(256) kWarmFeatures(View v) {
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const SegParams& P = v.seg[v.tileSeg[tile]];
  const uint32_t row = tile * kTile + threadIdx.x;
  float4 pn = make_float4(0.f, 0.f, 0.f, P.field.present ? P.field.band : 0.02f);
  if (!(v.rowFlags[row] & kRowPad) && P.field.present) {
    const F3 x = v.x[row];
    const F3 n = bodyNormal(v.fieldValues, P.field, x);
    pn = make_float4(n.x, n.y, n.z, bodyDistance(v.fieldValues, P.field, x));
  }
  v.phiN[row] = pn;
}

// Per row: step = the move from the learned guess (in u) to the previous-velocity guess h v; each segment
// evaluates its energy at the learned guess first (lsAlpha = 0).
__global__ void __launch_bounds__(256) kWarmPrepare(View v) {
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  const SegParams& P = v.seg[s];
  if (tile == P.tileBegin && threadIdx.x == 0) v.state[s].lsAlpha = 0.0;
  const uint32_t row = tile * kTile + threadIdx.x;
  v.step[row] = (v.rowFlags[row] & (kRowPinned | kRowPad)) ? f3(0, 0, 0) : P.h * v.v[row] - v.u[row];
}

// Per segment: energy of the learned guess into eTrial, then evaluate the previous-velocity guess (lsAlpha = 1).
__global__ void __launch_bounds__(256) kWarmLearnedEnergy(View v) {
  __shared__ double lds[8];
  const uint32_t s = blockIdx.x;
  SegState* S = v.state + s;
  const SegParams& P = v.seg[s];
  const double e = rangeSum256(v.partC, P.tileBegin, P.tileEnd, lds);
  if (threadIdx.x == 0) {
    S->eTrial = e;
    S->lsAlpha = 1.0;
  }
}

// Per segment: keep the learned guess unless its energy is above the previous-velocity guess's.
__global__ void __launch_bounds__(256) kWarmDecide(View v) {
  __shared__ double lds[8];
  const uint32_t s = blockIdx.x;
  SegState* S = v.state + s;
  const SegParams& P = v.seg[s];
  const double ePlain = rangeSum256(v.partC, P.tileBegin, P.tileEnd, lds);
  if (threadIdx.x == 0) {
    const bool reject = !(S->eTrial <= ePlain);
    S->warmRejected = reject ? 1u : 0u;
    if (reject)
      S->cntWarmRejected += 1u;
    else
      S->cntWarmAccepted += 1u;
    S->lsAlpha = 1.0;
  }
}

// Rows of a segment that rejected the learned guess go back to u = h v, computed as the predict kernel computes
// it with dv = 0, so a rejected guess leaves no trace in the trajectory.
__global__ void __launch_bounds__(256) kWarmApply(View v) {
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t s = v.tileSeg[tile];
  if (!v.state[s].warmRejected) return;
  const uint32_t row = tile * kTile + threadIdx.x;
  const F3 zero = f3(0, 0, 0);
  v.u[row] = (v.rowFlags[row] & (kRowPinned | kRowPad)) ? zero : v.seg[s].h * (v.v[row] + zero);
}

__global__ void kFill(uint32_t* p, uint32_t value, uint64_t count) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = value;
}

// ---------------------------------------------------------------------------------------------------------------
// Debug / standalone kernels.

// out = A in over free rows and columns, no filter (multiply, cpu_solver.cpp:147-159).
__global__ void __launch_bounds__(256) kSpmvPlain(View v, const F3* __restrict__ in, F3* out) {
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t slice = tile * kSlicesPerTile + wave;
  const uint32_t row = slice * kSlice + lane;
  const F3 prod = sellRowProduct(v, slice, lane, [&](uint32_t c) { return in[c]; });
  out[row] = (v.rowFlags[row] & kRowPinned) ? f3(0, 0, 0) : prod;
}

// The same product for a general matrix (every block stored, lowK null) whose rows have at most 16 blocks: column
// indices and block values of all 16 slots in one memory round trip — slots past K are read unconditionally (the
// value buffer carries kSellSlack zero floats past its end) and point at the dummy column, so they add nothing —
// then the gathered x. Row sums in four groups (slot index mod 4).
// One wavefront per SIMD is all a tile-per-workgroup launch of this size gets, so the kernel may use the whole
// register file: the 16 columns, 144 block values and 48 gathered values of a lane are all in flight at once.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
kSpmv16(View v, const F3* __restrict__ in, F3* out) {
  const uint32_t tile = xcdMap(blockIdx.x, v.nTiles);
  if (tile >= v.nTiles) return;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t slice = __builtin_amdgcn_readfirstlane(tile * kSlicesPerTile + wave);
  const uint32_t row = slice * kSlice + lane;
  const uint32_t K = ldUniform(v.sellK + slice);
  const uint32_t base = ldUniform(v.sellBase + slice);
  const uint32_t* __restrict__ col = v.sellCol + base / 9u + lane;
  const float* __restrict__ val = v.sellVal + base + 9u * lane;
  // Columns (unconditional loads, see sellColumn) issued before the block values and used only after them, so the
  // gathers wait for the columns alone.
  uint32_t c[16];
#pragma unroll
  for (uint32_t t = 0; t < 16; ++t) c[t] = col[t * kSlice];
  __builtin_amdgcn_sched_barrier(0);
  float bv[16][9];
#pragma unroll
  for (uint32_t t = 0; t < 16; ++t)
#pragma unroll
    for (int e = 0; e < 9; ++e) bv[t][e] = val[t * 9u * kSlice + e];
  __builtin_amdgcn_sched_barrier(0);
  F3 x[16];
#pragma unroll
  for (uint32_t t = 0; t < 16; ++t) x[t] = in[t < K ? c[t] : v.nRows];
  __builtin_amdgcn_sched_barrier(0);  // every gather issued before the first product waits on memory
  float g[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q) g[q][0] = g[q][1] = g[q][2] = 0.0f;
#pragma unroll
  for (uint32_t t = 0; t < 16; ++t) {
    float* a = g[t & 3u];
    a[0] += bv[t][0] * x[t].x + bv[t][1] * x[t].y + bv[t][2] * x[t].z;
    a[1] += bv[t][3] * x[t].x + bv[t][4] * x[t].y + bv[t][5] * x[t].z;
    a[2] += bv[t][6] * x[t].x + bv[t][7] * x[t].y + bv[t][8] * x[t].z;
  }
  const F3 prod = f3(((g[0][0] + g[1][0]) + g[2][0]) + g[3][0], ((g[0][1] + g[1][1]) + g[2][1]) + g[3][1],
                     ((g[0][2] + g[1][2]) + g[2][2]) + g[3][2]);
  out[row] = (v.rowFlags[row] & kRowPinned) ? f3(0, 0, 0) : prod;
}

// SELL values -> canonical slot order (9 floats per slot), for parity checks of the assembled matrix.
__global__ void __launch_bounds__(256) kSellToSlots(View v, float* out) {
  const uint32_t row = blockIdx.x * 256u + threadIdx.x;
  if (row >= v.nRows) return;
  const uint32_t slice = row / kSlice, lane = row % kSlice;
  const uint32_t rs = v.rowStart[row], re = v.rowStart[row + 1], kd = v.diagSlot[row];
  const float* base = v.sellVal + v.sellBase[slice] + 9u * lane;
  for (uint32_t sl = rs; sl < kd; ++sl) {  // below the diagonal: the transpose of the column's upper block
    const float* b = v.sellVal + 9u * v.lowBlk[v.lowBase[slice] + (sl - rs) * kSlice + lane];
    for (int c = 0; c < 9; ++c) out[9u * sl + c] = b[(c % 3) * 3 + c / 3];
  }
  for (uint32_t sl = kd; sl < re; ++sl)
    for (int c = 0; c < 9; ++c) out[9u * sl + c] = base[(sl - kd) * 9u * kSlice + c];
}

// Standalone term evaluation for the FP32-vs-FP64 checks: energy, every vertex's gradient and every block (a, b)
// of one term per thread, from the same term forms the assembly uses. kind 0 stretch (params dmInv[4], area),
// 1 hinge (K, rest angle), 2 seam (target gap, K), 3 friction (x = slip, params normal[3], mu lambda, eps).
__device__ inline void putTerm(float* grad, float* hess, uint32_t i, int N, int a, F3 ga, const Blk* rowBlocks) {
  float* gp = grad + (size_t(i) * N + a) * 3;
  gp[0] = ga.x;
  gp[1] = ga.y;
  gp[2] = ga.z;
  for (int b = 0; b < N; ++b)
    for (int c = 0; c < 9; ++c) hess[(size_t(i) * N * N + a * N + b) * 9 + c] = rowBlocks[b].e[c];
}

__global__ void kEvalTerms(int kind, uint32_t n, const float* x, const float* params, const Fabric32* fabrics,
                           float limitScale, float* energy, float* grad, float* hess) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int N = kind == 0 ? 3 : kind == 1 ? 4 : kind == 2 ? 2 : 1;
  const int np = kind == 0 ? 5 : kind == 1 ? 2 : kind == 2 ? 2 : 5;
  F3 xs[4];
  for (int a = 0; a < N; ++a) xs[a] = f3(x[(i * N + a) * 3], x[(i * N + a) * 3 + 1], x[(i * N + a) * 3 + 2]);
  const float* pr = params + size_t(i) * np;
  Blk rows[4];
  if (kind == 0) {
    const float dmInv[4] = {pr[0], pr[1], pr[2], pr[3]};
    const Shape3 sh = shapeOf(dmInv);
    const StretchForm form = stretchForm<true>(xs, sh, pr[4], fabrics[i], limitScale);
    energy[i] = form.energy;
    for (int a = 0; a < 3; ++a) {
      const StretchRow w = stretchRow(form, sh, uint32_t(a));
      for (int b = 0; b < 3; ++b) rows[b] = stretchBlock(w, sh, b);
      putTerm(grad, hess, i, 3, a, w.g, rows);
    }
  } else if (kind == 1) {
    const HingeForm form = hingeForm(xs);
    const float K = pr[0], bend = form.theta - pr[1];
    energy[i] = 0.5f * K * bend * bend;
    for (int a = 0; a < 4; ++a) {
      for (int b = 0; b < 4; ++b) rows[b] = blkOuter(form.g[a], form.g[b], K);
      putTerm(grad, hess, i, 4, a, (K * bend) * form.g[a], rows);
    }
  } else if (kind == 2) {
    const SeamForm form = seamForm<true>(xs[0], xs[1], pr[0], pr[1]);
    energy[i] = form.energy;
    Blk neg = blkZero();
    if (form.active) blkAxpy(neg, -1.0f, form.S);
    const Blk pos = form.active ? form.S : blkZero();
    rows[0] = pos;
    rows[1] = neg;
    putTerm(grad, hess, i, 2, 0, form.g0, rows);
    rows[0] = neg;
    rows[1] = pos;
    putTerm(grad, hess, i, 2, 1, -form.g0, rows);
  } else {
    F3 g = f3(0, 0, 0);
    Blk H = blkZero();
    energy[i] = frictionAdd<true>(xs[0], f3(pr[0], pr[1], pr[2]), pr[3], pr[4], g, H);
    rows[0] = H;
    putTerm(grad, hess, i, 1, 0, g, rows);
  }
}

__global__ void kFieldQuery(const float* values, Field32 f, uint32_t n, const float* p, float* phi, float* normal) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const F3 q = f3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
  phi[i] = bodyDistance(values, f, q);
  const F3 g = bodyNormal(values, f, q);
  normal[3 * i] = g.x;
  normal[3 * i + 1] = g.y;
  normal[3 * i + 2] = g.z;
}

// ---------------------------------------------------------------------------------------------------------------
// Structure builder: incidence lists, block-CSR pattern, diagonal slots, term slot tables and the SELL layout, all
// on the device, reproducing sim_model.cpp:41-74 exactly.

__device__ inline bool termItem(const BuildView& b, uint32_t i, uint32_t& kind, uint32_t& t, uint32_t& a,
                                uint32_t& vtx) {
  const uint32_t nS = 3u * b.nStretch, nH = 4u * b.nHinge, nK = 2u * b.nSeam;
  if (i < nS) {
    kind = kKindStretch;
    t = i / 3u;
    a = i % 3u;
    vtx = b.stretch[t].v[a];
    return true;
  }
  i -= nS;
  if (i < nH) {
    kind = kKindHinge;
    t = i / 4u;
    a = i % 4u;
    vtx = b.hinges[t].v[a];
    return true;
  }
  i -= nH;
  if (i < nK) {
    kind = kKindSeam;
    t = i / 2u;
    a = i % 2u;
    vtx = a == 0 ? b.seams[t].a : b.seams[t].b;
    return true;
  }
  return false;
}

__global__ void kIncCount(BuildView b, uint32_t items) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= items) return;
  uint32_t kind, t, a, vtx;
  if (termItem(b, i, kind, t, a, vtx)) atomicAdd(b.incCount + vtx, 1u);
}

__global__ void kIncFill(BuildView b, uint32_t items, uint32_t* cursor) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= items) return;
  uint32_t kind, t, a, vtx;
  if (!termItem(b, i, kind, t, a, vtx)) return;
  const uint32_t pos = atomicAdd(cursor + vtx, 1u);
  b.inc[b.incStart[vtx] + pos] = incEncode(kind, t, a);
}

// Exclusive scan of in[0, n) into out[0, n] (one 1024-thread workgroup; the builder runs once per build).
__global__ void __launch_bounds__(1024) kScan(const uint32_t* in, uint32_t* out, uint32_t n) {
  __shared__ uint32_t buf[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < n; base += 1024u) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t val = i < n ? in[i] : 0u;
    buf[threadIdx.x] = val;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
      const uint32_t add = threadIdx.x >= off ? buf[threadIdx.x - off] : 0u;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < n) out[i] = carry + buf[threadIdx.x] - val;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}

// Sorted, duplicate-free column list of a row: {row} and every vertex of every incident term.
__device__ inline uint32_t rowColumns(const BuildView& b, uint32_t row, uint32_t* cols, uint32_t* err) {
  uint32_t n = 0;
  auto insert = [&](uint32_t c) {
    uint32_t pos = 0;
    while (pos < n && cols[pos] < c) ++pos;
    if (pos < n && cols[pos] == c) return;
    if (n >= kMaxRow) {
      *err = 2u;
      return;
    }
    for (uint32_t q = n; q > pos; --q) cols[q] = cols[q - 1];
    cols[pos] = c;
    ++n;
  };
  insert(row);
  for (uint32_t i = b.incStart[row]; i < b.incStart[row + 1]; ++i) {
    const uint32_t e = b.inc[i];
    const uint32_t kind = incKind(e), t = incTerm(e);
    if (kind == kKindStretch) {
      for (int q = 0; q < 3; ++q) insert(b.stretch[t].v[q]);
    } else if (kind == kKindHinge) {
      for (int q = 0; q < 4; ++q) insert(b.hinges[t].v[q]);
    } else {
      insert(b.seams[t].a);
      insert(b.seams[t].b);
    }
  }
  return n;
}

// Sorts each row's incidences by (kind, term, local) and counts its columns.
__global__ void kRowSortCount(BuildView b) {
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= b.nRows) return;
  const uint32_t i0 = b.incStart[row], i1 = b.incStart[row + 1];
  if (i1 - i0 > kMaxIncidence) {
    *b.error = 1u;
    b.rowLen[row] = 1;
    return;
  }
  for (uint32_t i = i0 + 1; i < i1; ++i) {
    const uint32_t key = b.inc[i];
    uint32_t j = i;
    while (j > i0 && b.inc[j - 1] > key) {
      b.inc[j] = b.inc[j - 1];
      --j;
    }
    b.inc[j] = key;
  }
  uint32_t cols[kMaxRow];
  b.rowLen[row] = rowColumns(b, row, cols, b.error);
}

__global__ void kPatternFill(BuildView b) {
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= b.nRows) return;
  uint32_t cols[kMaxRow];
  const uint32_t n = rowColumns(b, row, cols, b.error);
  const uint32_t rs = b.rowStart[row];
  for (uint32_t k = 0; k < n; ++k) {
    b.colIndex[rs + k] = cols[k];
    if (cols[k] == row) b.diagSlot[row] = rs + k;
  }
}

__device__ inline uint32_t lowerBound(const uint32_t* colIndex, uint32_t lo, uint32_t hi, uint32_t c) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (colIndex[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// slot(v_a, v_b) for every term and (a, b) (sim_model.cpp:35-39, :62-72).
__global__ void kSlots(BuildView b) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t total = b.nStretch + b.nHinge + b.nSeam;
  if (i >= total) return;
  uint32_t vs[4];
  uint32_t m;
  uint32_t* out;
  if (i < b.nStretch) {
    m = 3;
    for (int q = 0; q < 3; ++q) vs[q] = b.stretch[i].v[q];
    out = b.stretchSlots + 9u * i;
  } else if (i < b.nStretch + b.nHinge) {
    const uint32_t t = i - b.nStretch;
    m = 4;
    for (int q = 0; q < 4; ++q) vs[q] = b.hinges[t].v[q];
    out = b.hingeSlots + 16u * t;
  } else {
    const uint32_t t = i - b.nStretch - b.nHinge;
    m = 2;
    vs[0] = b.seams[t].a;
    vs[1] = b.seams[t].b;
    out = b.seamSlots + 4u * t;
  }
  for (uint32_t a = 0; a < m; ++a)
    for (uint32_t c = 0; c < m; ++c)
      out[a * m + c] = lowerBound(b.colIndex, b.rowStart[vs[a]], b.rowStart[vs[a] + 1], vs[c]);
}

// Slots of each slice: upper (the diagonal and the columns above it) and lower (the columns below it). Every row
// holds its diagonal, so a row's lower count is its diagonal's position in its sorted column list.
__global__ void kSellK(BuildView b) {
  const uint32_t slice = blockIdx.x * blockDim.x + threadIdx.x;
  if (slice >= b.nSlices) return;
  uint32_t K = 0, KL = 0;
  for (uint32_t l = 0; l < kSlice; ++l) {
    const uint32_t row = slice * kSlice + l;
    const uint32_t lower = b.diagSlot[row] - b.rowStart[row];
    K = max(K, b.rowLen[row] - lower);
    KL = max(KL, lower);
  }
  b.sellK[slice] = K;
  b.sellSize[slice] = K * 9u * kSlice;
  b.lowK[slice] = KL;
  b.lowSize[slice] = KL * kSlice;
}

// Every slice to the largest slice's upper and lower slot counts (when neither exceeds kUniformSlots).
__global__ void kSellUniform(BuildView b, uint32_t K, uint32_t KL) {
  const uint32_t slice = blockIdx.x * blockDim.x + threadIdx.x;
  if (slice >= b.nSlices) return;
  b.sellK[slice] = K;
  b.sellSize[slice] = K * 9u * kSlice;
  b.lowK[slice] = KL;
  b.lowSize[slice] = KL * kSlice;
}

// Upper slot k of a row: column colIndex[diag + k]; lower slot k: column colIndex[rowStart + k] and the block of
// (that column, this row) in the column's upper slots. Pinned columns point at the dummy column nRows; padding
// slots at the dummy column and the zero block.
__global__ void kSellCol(BuildView b, const uint32_t* sellBase, const uint32_t* lowBase, uint32_t zeroBlock) {
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= b.nRows) return;
  const uint32_t slice = row / kSlice, lane = row % kSlice;
  const uint32_t K = b.sellK[slice], KL = b.lowK[slice];
  const uint32_t base = sellBase[slice] / 9u, lbase = lowBase[slice];
  const uint32_t rs = b.rowStart[row], re = b.rowStart[row + 1], kd = b.diagSlot[row];
  for (uint32_t k = 0; k < K; ++k) {
    uint32_t c = b.nRows;
    if (kd + k < re) {
      c = b.colIndex[kd + k];
      if (b.rowFlags[c] & kRowPinned) c = b.nRows;
    }
    b.sellCol[base + k * kSlice + lane] = c;
  }
  for (uint32_t k = 0; k < KL; ++k) {
    uint32_t c = b.nRows, blk = zeroBlock;
    if (rs + k < kd) {
      const uint32_t j = b.colIndex[rs + k];
      const uint32_t u = lowerBound(b.colIndex, b.rowStart[j], b.rowStart[j + 1], row) - b.diagSlot[j];
      blk = sellBase[j / kSlice] / 9u + u * kSlice + j % kSlice;
      c = (b.rowFlags[j] & kRowPinned) ? b.nRows : j;
    }
    b.lowCol[lbase + k * kSlice + lane] = c;
    b.lowBlk[lbase + k * kSlice + lane] = blk;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Launchers.

void launchSubstepBegin(const View& v, bool learned, hipStream_t st) {
  kSubstepBegin<<<xcdGrid(v.nTiles), 256, 0, st>>>(v, learned ? 1u : 0u);
}
void launchHoldContacts(const View& v, hipStream_t st) { kHoldContacts<<<xcdGrid(v.nTiles), 256, 0, st>>>(v); }
void launchAssemble(const View& v, uint32_t maxK, bool, hipStream_t st) {
  kAssemble<<<xcdGrid(v.nSlices), 64, maxK * 9u * kSlice * sizeof(float), st>>>(v);
}
void launchNewtonBegin(const View& v, hipStream_t st) { kNewtonBegin<<<v.nSegments, 256, 0, st>>>(v); }
void launchPcgInit(const View& v, hipStream_t st) { kPcgInit<<<xcdGrid(v.nTiles), 256, 0, st>>>(v); }
void launchPcgSpmv(const View& v, uint32_t k, hipStream_t st, hipEvent_t start, hipEvent_t stop) {
  F3* pOld = (k & 1u) ? v.pB : v.pA;
  F3* pNew = (k & 1u) ? v.pA : v.pB;
  hipExtLaunchKernelGGL(kPcgSpmv, dim3(xcdGrid(v.nTiles)), dim3(256), 0, st, start, stop, 0, v, k,
                        static_cast<const F3*>(pOld), pNew);
}
void launchPcgUpdate(const View& v, uint32_t k, hipStream_t st) {
  kPcgUpdate<<<xcdGrid(v.nTiles), 256, 0, st>>>(v, k, (k & 1u) ? v.pA : v.pB);
}
void launchPcgIteration(const View& v, uint32_t k, hipStream_t st) {
  launchPcgSpmv(v, k, st);
  launchPcgUpdate(v, k, st);
}
void launchPcgStepInit(const View& v, hipStream_t st) { kPcgStepInit<<<xcdGrid(v.nTiles), 256, 0, st>>>(v); }
void launchPcgStep(const View& v, uint32_t parity, PcgShape shape, hipStream_t st, hipEvent_t start, hipEvent_t stop) {
  const dim3 grid(xcdGrid(v.nTiles));
  const uint32_t q = parity & 1u;
  if (shape.halo && !v.tileHalo) throw std::invalid_argument("PCG halo shape: a tile's columns exceed the window");
  if (shape.waves == 1 && shape.slots == 8 && !shape.halo && v.uniformK && v.nSegments == 1)
    hipExtLaunchKernelGGL((kPcgStep<1, 8, false, true>), grid, dim3(256), 0, st, start, stop, 0, v, q);
  else if (shape.waves == 1 && shape.slots == 8 && !shape.halo)
    hipExtLaunchKernelGGL((kPcgStep<1, 8, false>), grid, dim3(256), 0, st, start, stop, 0, v, q);
  else if (shape.waves == 1 && shape.slots == 4 && !shape.halo)
    hipExtLaunchKernelGGL((kPcgStep<1, 4, false>), grid, dim3(256), 0, st, start, stop, 0, v, q);
  else if (shape.waves == 2 && shape.slots == 8 && !shape.halo)
    hipExtLaunchKernelGGL((kPcgStep<2, 8, false>), grid, dim3(512), 0, st, start, stop, 0, v, q);
  else if (shape.waves == 4 && shape.slots == 8 && !shape.halo)
    hipExtLaunchKernelGGL((kPcgStep<4, 8, false>), grid, dim3(1024), 0, st, start, stop, 0, v, q);
  else if (shape.waves == 1 && shape.slots == 8 && shape.halo)
    hipExtLaunchKernelGGL((kPcgStep<1, 8, true>), grid, dim3(256), 0, st, start, stop, 0, v, q);
  else if (shape.waves == 1 && shape.slots == 4 && shape.halo)
    hipExtLaunchKernelGGL((kPcgStep<1, 4, true>), grid, dim3(256), 0, st, start, stop, 0, v, q);
  else
    throw std::invalid_argument("PCG launch shape must be 1x8, 1x4, 2x8 or 4x8, or 1x8 / 1x4 with the halo");
}
uint32_t pcgSolveCapacity(bool extra) {
  int device = 0, fits = 0;
  hipDeviceProp_t prop{};
  if (hipGetDevice(&device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) return 0u;
  const void* fn = extra ? reinterpret_cast<const void*>(kPcgSolve<true>) : reinterpret_cast<const void*>(kPcgSolve<false>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fits, fn, 256, 0) != hipSuccess) return 0u;
  // one workgroup per CU (the kernel's registers are sized for exactly that)
  return fits >= 1 ? uint32_t(prop.multiProcessorCount) : 0u;
}

uint32_t pcgSolvePairCapacity(bool extra) {
  int device = 0, fits = 0;
  hipDeviceProp_t prop{};
  if (hipGetDevice(&device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) return 0u;
  const void* fn = extra ? reinterpret_cast<const void*>(kPcgSolvePair<true>)
                         : reinterpret_cast<const void*>(kPcgSolvePair<false>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fits, fn, 512, 0) != hipSuccess) return 0u;
  return fits >= 1 ? uint32_t(prop.multiProcessorCount) : 0u;
}

hipError_t launchPcgSolvePairs(const View& v, const SolveCtl& c, uint32_t counterLines, bool extra, bool clearAbort,
                               hipStream_t st) {
  launchFill(c.ctr, 0u, uint64_t(counterLines) * 64u + (clearAbort ? 1u : 0u), st);
  const dim3 grid(c.order ? v.nPairs : xcdGrid(v.nPairs));
  if (extra)
    hipExtLaunchKernelGGL((kPcgSolvePair<true>), grid, dim3(512), 0, st, nullptr, nullptr, 0, v, c);
  else
    hipExtLaunchKernelGGL((kPcgSolvePair<false>), grid, dim3(512), 0, st, nullptr, nullptr, 0, v, c);
  return hipGetLastError();
}

hipError_t launchPcgSolve(const View& v, const SolveCtl& c, uint32_t counterLines, bool extra, bool clearAbort,
                          hipStream_t st) {
  // An ordinary dispatch (capturable in a graph, and traceable by the profiler, which does not survive cooperative
  // dispatches): the host has checked that the grid fits the device at once (pcgSolveCapacity), the solver's stream
  // runs nothing beside it, and a workgroup that found no room would only make its segment's waits give up.
  launchFill(c.ctr, 0u, uint64_t(counterLines) * 64u + (clearAbort ? 1u : 0u), st);
  const dim3 grid(c.order ? v.nTiles : xcdGrid(v.nTiles));
  if (extra)
    hipExtLaunchKernelGGL((kPcgSolve<true>), grid, dim3(256), 0, st, nullptr, nullptr, 0, v, c);
  else
    hipExtLaunchKernelGGL((kPcgSolve<false>), grid, dim3(256), 0, st, nullptr, nullptr, 0, v, c);
  return hipGetLastError();
}
void launchPcgRefine(const View& v, uint32_t maxRestarts, const uint32_t* abortFlag, hipStream_t st) {
  kTrueResidual<<<xcdGrid(v.nTiles), 256, 0, st>>>(v, abortFlag);
  kRefineDecide<<<v.nSegments, 256, 0, st>>>(v, maxRestarts, abortFlag);
}
void launchContactForceSlope(const View& v, hipStream_t st) {
  kContactForceSlope<<<xcdGrid(v.nTiles), 256, 0, st>>>(v);
}
void launchNewtonSlope(const View& v, hipStream_t st) { kNewtonSlope<<<v.nSegments, 256, 0, st>>>(v); }
void launchEnergyTrial(const View& v, hipStream_t st) { kEnergyTrial<<<xcdGrid(v.nTiles), 256, 0, st>>>(v); }
void launchLineSearchDecide(const View& v, hipStream_t st) { kLineSearchDecide<<<v.nSegments, 256, 0, st>>>(v); }
void launchApply(const View& v, hipStream_t st) { kApply<<<xcdGrid(v.nTiles), 256, 0, st>>>(v); }
void launchNewtonEnd(const View& v, hipStream_t st) { kNewtonEnd<<<v.nSegments, 256, 0, st>>>(v); }
void launchFinish(const View& v, hipStream_t st) { kFinish<<<xcdGrid(v.nTiles), 256, 0, st>>>(v); }
void launchStats(const View& v, const double* massF64, double* tilePart, double* segOut, hipStream_t st) {
  kStatsTile<<<xcdGrid(v.nTiles), 256, 0, st>>>(v, massF64, tilePart);
  kStatsSegment<<<v.nSegments, 256, 0, st>>>(v, tilePart, segOut);
}
void launchWarmFeatures(const View& v, hipStream_t st) { kWarmFeatures<<<xcdGrid(v.nTiles), 256, 0, st>>>(v); }
void launchWarmCheck(const View& v, hipStream_t st) {
  kWarmPrepare<<<xcdGrid(v.nTiles), 256, 0, st>>>(v);
  kEnergyTrial<<<xcdGrid(v.nTiles), 256, 0, st>>>(v);
  kWarmLearnedEnergy<<<v.nSegments, 256, 0, st>>>(v);
  kEnergyTrial<<<xcdGrid(v.nTiles), 256, 0, st>>>(v);
  kWarmDecide<<<v.nSegments, 256, 0, st>>>(v);
  kWarmApply<<<xcdGrid(v.nTiles), 256, 0, st>>>(v);
}
void launchFill(uint32_t* p, uint32_t value, uint64_t count, hipStream_t st) {
  if (count == 0) return;
  kFill<<<uint32_t((count + 255) / 256), 256, 0, st>>>(p, value, count);
}
void launchSpmvPlain(const View& v, uint32_t maxK, const F3* in, F3* out, hipStream_t st) {
  if (!v.lowK && maxK <= 16u)
    kSpmv16<<<xcdGrid(v.nTiles), 256, 0, st>>>(v, in, out);
  else
    kSpmvPlain<<<xcdGrid(v.nTiles), 256, 0, st>>>(v, in, out);
}
void launchSellToSlots(const View& v, float* out, hipStream_t st) {
  kSellToSlots<<<(v.nRows + 255) / 256, 256, 0, st>>>(v, out);
}
void launchEvalTerms(int kind, uint32_t n, const float* x, const float* params, const Fabric32* fabrics,
                     float limitScale, float* energy, float* grad, float* hess, hipStream_t st) {
  if (n == 0) return;
  kEvalTerms<<<(n + 63) / 64, 64, 0, st>>>(kind, n, x, params, fabrics, limitScale, energy, grad, hess);
}
void launchFieldQuery(const float* values, Field32 field, uint32_t n, const float* p, float* phi, float* normal,
                      hipStream_t st) {
  if (n == 0) return;
  kFieldQuery<<<(n + 255) / 256, 256, 0, st>>>(values, field, n, p, phi, normal);
}

void launchBuildStructure(const BuildView& bIn, BuiltStructure& out, hipStream_t st) {
  BuildView b = bIn;
  const uint32_t items = 3u * b.nStretch + 4u * b.nHinge + 2u * b.nSeam;
  const uint32_t nRows = b.nRows;
  hipCheck(hipMemsetAsync(b.incCount, 0, sizeof(uint32_t) * nRows, st), "memset");
  hipCheck(hipMemsetAsync(b.error, 0, sizeof(uint32_t), st), "memset");
  if (items) kIncCount<<<(items + 255) / 256, 256, 0, st>>>(b, items);
  kScan<<<1, 1024, 0, st>>>(b.incCount, b.incStart, nRows);
  hipCheck(hipMemsetAsync(b.incCount, 0, sizeof(uint32_t) * nRows, st), "memset");
  if (items) kIncFill<<<(items + 255) / 256, 256, 0, st>>>(b, items, b.incCount);
  kRowSortCount<<<(nRows + 127) / 128, 128, 0, st>>>(b);
  kScan<<<1, 1024, 0, st>>>(b.rowLen, b.rowStart, nRows);
  uint32_t hostTotal = 0, hostErr = 0;
  hipCheck(hipMemcpyAsync(&hostTotal, b.rowStart + nRows, sizeof(uint32_t), hipMemcpyDeviceToHost, st), "copy");
  hipCheck(hipMemcpyAsync(&hostErr, b.error, sizeof(uint32_t), hipMemcpyDeviceToHost, st), "copy");
  hipCheck(hipStreamSynchronize(st), "sync");
  if (hostErr) {
    throw std::invalid_argument(hostErr == 1u ? "hip: a particle has more than 96 incident terms"
                                              : "hip: a particle couples to more than 64 particles");
  }
  out.nBlocks = hostTotal;
  hipCheck(hipMalloc(&b.colIndex, sizeof(uint32_t) * std::max<uint32_t>(out.nBlocks, 1)), "malloc colIndex");
  kPatternFill<<<(nRows + 127) / 128, 128, 0, st>>>(b);
  const uint32_t terms = b.nStretch + b.nHinge + b.nSeam;
  if (terms) kSlots<<<(terms + 255) / 256, 256, 0, st>>>(b);
  kSellK<<<(b.nSlices + 255) / 256, 256, 0, st>>>(b);
  // The largest slot counts: copied back (small) to size the assembly kernel's LDS and pick the layout.
  {
    std::vector<uint32_t> hk(b.nSlices), hl(b.nSlices);
    hipCheck(hipMemcpyAsync(hk.data(), b.sellK, sizeof(uint32_t) * b.nSlices, hipMemcpyDeviceToHost, st), "copy");
    hipCheck(hipMemcpyAsync(hl.data(), b.lowK, sizeof(uint32_t) * b.nSlices, hipMemcpyDeviceToHost, st), "copy");
    hipCheck(hipStreamSynchronize(st), "sync");
    out.maxK = 0;
    out.maxLowK = 0;
    for (uint32_t i = 0; i < b.nSlices; ++i) {
      out.maxK = std::max(out.maxK, hk[i]);
      out.maxLowK = std::max(out.maxLowK, hl[i]);
    }
  }
  out.uniform = out.maxK <= kUniformSlots && out.maxLowK <= kUniformSlots && out.maxK > 0;
  if (out.uniform) kSellUniform<<<(b.nSlices + 255) / 256, 256, 0, st>>>(b, out.maxK, out.maxLowK);
  kScan<<<1, 1024, 0, st>>>(b.sellSize, b.sellSize, b.nSlices);
  kScan<<<1, 1024, 0, st>>>(b.lowSize, b.lowSize, b.nSlices);
  uint32_t hostSell = 0, hostLow = 0;
  hipCheck(hipMemcpyAsync(&hostSell, b.sellSize + b.nSlices, sizeof(uint32_t), hipMemcpyDeviceToHost, st), "copy");
  hipCheck(hipMemcpyAsync(&hostLow, b.lowSize + b.nSlices, sizeof(uint32_t), hipMemcpyDeviceToHost, st), "copy");
  hipCheck(hipStreamSynchronize(st), "sync");
  out.sellFloats = hostSell;
  out.lowEntries = hostLow;
  const uint32_t zeroBlock = hostSell / 9u;
  hipCheck(hipMalloc(&b.sellCol, sizeof(uint32_t) * (uint64_t(hostSell / 9u) + kSellColSlack)), "malloc sellCol");
  hipCheck(hipMalloc(&b.lowCol, sizeof(uint32_t) * (uint64_t(hostLow) + kSellColSlack)), "malloc lowCol");
  hipCheck(hipMalloc(&b.lowBlk, sizeof(uint32_t) * (uint64_t(hostLow) + kSellColSlack)), "malloc lowBlk");
  kSellCol<<<(nRows + 255) / 256, 256, 0, st>>>(b, b.sellSize, b.lowSize, zeroBlock);
  launchFill(b.sellCol + hostSell / 9u, nRows, kSellColSlack, st);
  launchFill(b.lowCol + hostLow, nRows, kSellColSlack, st);
  launchFill(b.lowBlk + hostLow, zeroBlock, kSellColSlack, st);
  out.colIndex = b.colIndex;
  out.sellCol = b.sellCol;
  out.lowCol = b.lowCol;
  out.lowBlk = b.lowBlk;
}

}  // namespace dk
