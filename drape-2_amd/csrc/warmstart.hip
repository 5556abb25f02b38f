// warmstart.hip — kernels and host side of the learned Newton warm start (network described in warmstart.hpp).
//
// Layout: every activation is row-major [rows x columns]; rows are the graph's rows (a multiple of 256). BF16
// values are stored as their 16-bit patterns (uint16_t) and converted with round-to-nearest-even.
//
// MFMA on one wavefront, gfx950's v_mfma_f32_16x16x32_bf16 wherever the reduction depth is a multiple of 32 (every
// layer but the 16-wide ones, which use v_mfma_f32_16x16x16_bf16): lane l holds A[l % 16][K (l / 16) + j] and
// B[K (l / 16) + j][l % 16] (j < K, K = 8 or 4 BF16 values) and receives D[4 (l / 16) + j][l % 16] (four FP32
// values). The GEMM tests check this layout against a host product.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>

#include "warmstart.hpp"

namespace ws {

namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("hip: ") + what + ": " + hipGetErrorString(e));
}

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kChunk = 256;  // rows per workgroup in the reduction kernels

}  // namespace

__device__ inline uint16_t toBf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);  // NaN stays a (quiet) NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                          // round to nearest even
  return uint16_t(u >> 16);
}

__device__ inline float fromBf16(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

// ---------------------------------------------------------------------------------------------------------------
// GEMMs.

// One MFMA step over a reduction slice: 32 deep (8 BF16 values per lane) or 16 deep (4 values).
__device__ inline f32x4 mfma32(u16x8 a, u16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                 0, 0, 0);
}
__device__ inline f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// Y[N x Mo] (FP32, row stride ldy) = X[N x Ki] (BF16, row stride ldx) . Bt^T with Bt [Mo x Ki] BF16 row-major.
// One wavefront per 16 rows, four per workgroup; N a multiple of 64.
template <int Ki, int Mo>
__global__ void __launch_bounds__(256) kGemm(const uint16_t* __restrict__ X, int ldx, const uint16_t* __restrict__ Bt,
                                             float* __restrict__ Y, int ldy, uint32_t N) {
  constexpr bool K32 = Ki % 32 == 0;
  constexpr int KL = K32 ? 8 : 4;  // reduction values per lane per step
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t row0 = (blockIdx.x * 4u + wave) * 16u;
  if (row0 >= N) return;
  const uint32_t r = lane & 15u, kq = (lane >> 4) * KL, oq = (lane >> 4) * 4u;
  f32x4 acc[Mo / 16];
#pragma unroll
  for (int t = 0; t < Mo / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k0 = 0; k0 < Ki; k0 += 4 * KL) {
    if constexpr (K32) {
      const u16x8 a = *reinterpret_cast<const u16x8*>(X + size_t(row0 + r) * ldx + k0 + kq);
#pragma unroll
      for (int t = 0; t < Mo / 16; ++t)
        acc[t] = mfma32(a, *reinterpret_cast<const u16x8*>(Bt + size_t(t * 16 + r) * Ki + k0 + kq), acc[t]);
    } else {
      const s16x4 a = *reinterpret_cast<const s16x4*>(X + size_t(row0 + r) * ldx + k0 + kq);
#pragma unroll
      for (int t = 0; t < Mo / 16; ++t)
        acc[t] = mfma16(a, *reinterpret_cast<const s16x4*>(Bt + size_t(t * 16 + r) * Ki + k0 + kq), acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < Mo / 16; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) Y[size_t(row0 + oq + j) * ldy + t * 16 + r] = acc[t][j];
}

// partial[chunk][Ki x Mo] = X[chunk rows]^T . D[chunk rows] (X BF16 [N x Ki], D BF16 [N x Mo]); 256 rows per
// workgroup, N a multiple of 256, 32 rows per MFMA step. The 16x16 output tiles are dealt to the four wavefronts
// in turn.
template <int Ki, int Mo>
__global__ void __launch_bounds__(256) kGemmTN(const uint16_t* __restrict__ X, int ldx, const uint16_t* __restrict__ D,
                                               int ldd, float* __restrict__ partial) {
  constexpr int TN = Mo / 16, T = (Ki / 16) * TN, PER = (T + 3) / 4;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t rBase = blockIdx.x * kChunk;
  const uint32_t r = lane & 15u, kq = (lane >> 4) * 8u, oq = (lane >> 4) * 4u;
  f32x4 acc[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) acc[p] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (uint32_t rs = rBase; rs < rBase + kChunk; rs += 32u) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int t = int(wave) + 4 * p;
      if (t < T) {
        const int ti = t / TN, tn = t % TN;
        u16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          a[j] = X[size_t(rs + kq + j) * ldx + ti * 16 + r];
          b[j] = D[size_t(rs + kq + j) * ldd + tn * 16 + r];
        }
        acc[p] = mfma32(a, b, acc[p]);
      }
    }
  }
  float* out = partial + size_t(blockIdx.x) * Ki * Mo;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int t = int(wave) + 4 * p;
    if (t < T) {
      const int ti = t / TN, tn = t % TN;
#pragma unroll
      for (int j = 0; j < 4; ++j) out[(ti * 16 + oq + j) * Mo + tn * 16 + r] = acc[p][j];
    }
  }
}

// out[k] (+)= sum over chunks c of partial[c][k], in chunk order.
__global__ void kReduceChunks(const float* __restrict__ partial, uint32_t chunks, uint32_t size, float* out, int accumulate) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= size) return;
  float s = 0.0f;
  for (uint32_t c = 0; c < chunks; ++c) s += partial[size_t(c) * size + k];
  out[k] = accumulate ? out[k] + s : s;
}

// partial[chunk][c] = sum over the chunk's rows of Y[row][c] (c < cols), rows in order.
__global__ void __launch_bounds__(256) kColSum(const float* __restrict__ Y, int ldy, int cols, float* partial) {
  const uint32_t c = threadIdx.x;
  if (c >= uint32_t(cols)) return;
  const uint32_t rBase = blockIdx.x * kChunk;
  float s = 0.0f;
  for (uint32_t r = rBase; r < rBase + kChunk; ++r) s += Y[size_t(r) * ldy + c];
  partial[size_t(blockIdx.x) * cols + c] = s;
}

// ---------------------------------------------------------------------------------------------------------------
// Forward stages.

__global__ void __launch_bounds__(256) kFeatures(GraphView g, FrameView f, uint16_t* __restrict__ feat) {
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= g.nRows) return;
  float x[kIn];
#pragma unroll
  for (int i = 0; i < kIn; ++i) x[i] = 0.0f;
  if (g.rowMask[row]) {
    const dk::F3 v = f.v[row], d = f.dvPrev[row];
    const float4 pn = f.phiNormal[row];
    x[0] = v.x / kVelScale;
    x[1] = v.y / kVelScale;
    x[2] = v.z / kVelScale;
    x[3] = d.x / kDvScale;
    x[4] = d.y / kDvScale;
    x[5] = d.z / kDvScale;
    x[6] = pn.w / kPhiScale;
    x[7] = pn.x;
    x[8] = pn.y;
    x[9] = pn.z;
  }
#pragma unroll
  for (int i = 0; i < kIn; ++i) feat[size_t(row) * kIn + i] = toBf16(x[i]);
}

// Y = relu(Y + b) in place (FP32 [N x 64]); BF16 copy into dst (row stride ldd).
__global__ void __launch_bounds__(256) kBiasRelu(float* __restrict__ Y, const float* __restrict__ b,
                                                 uint16_t* __restrict__ dst, int ldd, float* __restrict__ keep,
                                                 uint32_t N) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * kHidden) return;
  const uint32_t row = i / kHidden, c = i % kHidden;
  const float y = fmaxf(Y[i] + b[c], 0.0f);
  Y[i] = y;
  if (keep) keep[i] = y;
  dst[size_t(row) * ldd + c] = toBf16(y);
}

// Edge features of the edge i -> j.
__device__ inline void edgeFeatures(const GraphView& g, const FrameView& f, uint32_t i, uint32_t j, float e[kEdge]) {
  const dk::F3 xi = f.x[i], xj = f.x[j], vi = f.v[i], vj = f.v[j];
  e[0] = (g.rest2[2 * j] - g.rest2[2 * i]) / kLenScale;
  e[1] = (g.rest2[2 * j + 1] - g.rest2[2 * i + 1]) / kLenScale;
  e[2] = (xj.x - xi.x) / kLenScale;
  e[3] = (xj.y - xi.y) / kLenScale;
  e[4] = (xj.z - xi.z) / kLenScale;
  e[5] = (vj.x - vi.x) / kVelScale;
  e[6] = (vj.y - vi.y) / kVelScale;
  e[7] = (vj.z - vi.z) / kVelScale;
}

__device__ inline float edgePre(const float* pq, uint32_t i, uint32_t j, uint32_t c, const float e[kEdge],
                                const float wc[kEdge], float bmc) {
  float z = pq[size_t(i) * 2 * kHidden + c] + pq[size_t(j) * 2 * kHidden + kHidden + c] + bmc;
#pragma unroll
  for (int k = 0; k < kEdge; ++k) z += e[k] * wc[k];
  return z;
}

// m_i = mean over neighbours j of relu(P_i + Q_j + e_ij Wc + bm); one wavefront per row, one lane per channel.
// Writes BF16 m into columns 64..127 of hm (row stride 128).
__global__ void __launch_bounds__(256) kMessage(GraphView g, FrameView f, const float* __restrict__ pq,
                                                const float* __restrict__ Wc, const float* __restrict__ bm,
                                                uint16_t* __restrict__ hm) {
  const uint32_t row = blockIdx.x * 4u + (threadIdx.x >> 6), c = threadIdx.x & 63u;
  if (row >= g.nRows) return;
  float wc[kEdge];
#pragma unroll
  for (int k = 0; k < kEdge; ++k) wc[k] = Wc[k * kHidden + c];
  const float bmc = bm[c];
  float acc = 0.0f;
  const uint32_t e0 = g.rowStart[row], e1 = g.rowStart[row + 1];
  for (uint32_t e = e0; e < e1; ++e) {
    const uint32_t j = g.nbr[e];
    float ef[kEdge];
    edgeFeatures(g, f, row, j, ef);
    acc += fmaxf(edgePre(pq, row, j, c, ef, wc, bmc), 0.0f);
  }
  const float m = e1 > e0 ? acc / float(e1 - e0) : 0.0f;
  hm[size_t(row) * 2 * kHidden + kHidden + c] = toBf16(m);
}

// U = relu(acc + bu) (kept for the backward), h += U; BF16 h into dst (row stride ldd).
__global__ void __launch_bounds__(256) kUpdate(const float* __restrict__ acc, const float* __restrict__ bu,
                                               float* __restrict__ U, float* __restrict__ h, uint16_t* __restrict__ dst,
                                               int ldd, uint32_t N) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * kHidden) return;
  const uint32_t row = i / kHidden, c = i % kHidden;
  const float u = fmaxf(acc[size_t(row) * kHidden + c] + bu[c], 0.0f);
  U[i] = u;
  const float hn = h[i] + u;
  h[i] = hn;
  dst[size_t(row) * ldd + c] = toBf16(hn);
}

// out += bd; dv (m/s) for masked rows.
__global__ void __launch_bounds__(256) kDecode(GraphView g, float* __restrict__ out, const float* __restrict__ bd,
                                               dk::F3* dv) {
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= g.nRows) return;
  float o[kOutUsed];
#pragma unroll
  for (int c = 0; c < kOut; ++c) {
    const float y = out[size_t(row) * kOut + c] + bd[c];
    out[size_t(row) * kOut + c] = y;
    if (c < kOutUsed) o[c] = y;
  }
  if (dv) dv[row] = g.rowMask[row] ? dk::F3{o[0] * kDvScale, o[1] * kDvScale, o[2] * kDvScale} : dk::F3{0.f, 0.f, 0.f};
}

// ---------------------------------------------------------------------------------------------------------------
// Backward stages.

// Per-chunk loss partials and dOut = 2 (out - t / kDvScale) / count on masked rows (columns >= 3 zero).
__global__ void __launch_bounds__(256) kLoss(GraphView g, const float* __restrict__ out, const dk::F3* __restrict__ t,
                                             float invCount, float* __restrict__ dOut, uint16_t* __restrict__ dOutBf,
                                             double* __restrict__ lossPart) {
  __shared__ double lds[4];
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (row < g.nRows) {
    const dk::F3 tt = t[row];
    const float tv[3] = {tt.x / kDvScale, tt.y / kDvScale, tt.z / kDvScale};
    const bool on = g.rowMask[row] != 0u;
#pragma unroll
    for (int c = 0; c < kOut; ++c) {
      float d = 0.0f;
      if (on && c < kOutUsed) {
        const float diff = out[size_t(row) * kOut + c] - tv[c];
        l += double(diff) * diff;
        d = 2.0f * diff * invCount;
      }
      dOut[size_t(row) * kOut + c] = d;
      dOutBf[size_t(row) * kOut + c] = toBf16(d);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
  if ((threadIdx.x & 63u) == 0) lds[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) lossPart[blockIdx.x] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

// dpre = dY * (Y > 0), FP32 and BF16 (all [N x 64]).
__global__ void __launch_bounds__(256) kReluBack(const float* __restrict__ dY, const float* __restrict__ Y,
                                                 float* __restrict__ dpre, uint16_t* __restrict__ dpreBf, uint32_t N) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * kHidden) return;
  const float d = Y[i] > 0.0f ? dY[i] : 0.0f;
  dpre[i] = d;
  dpreBf[i] = toBf16(d);
}

// Message backward, gather form: for row i, dP_i = sum_j g_ij and dQ_i = sum_{j in N(i)} g_ji, with
// g_ij = dm_i / deg_i * [z_ij > 0]; the row's dWc = sum_j e_ij g_ij^T and dbm = sum_j g_ij are summed per
// workgroup (256 rows, each wavefront taking every fourth row) into partial[block][kEdge * 64 + 64].
__global__ void __launch_bounds__(256) kMessageBack(GraphView g, FrameView f, const float* __restrict__ pq,
                                                    const float* __restrict__ Wc, const float* __restrict__ bm,
                                                    const float* __restrict__ dHm, float* __restrict__ dPq,
                                                    uint16_t* __restrict__ dPqBf, float* __restrict__ partial) {
  __shared__ float red[4][kEdge + 1][kHidden];
  const uint32_t wave = threadIdx.x >> 6, c = threadIdx.x & 63u;
  float wc[kEdge];
#pragma unroll
  for (int k = 0; k < kEdge; ++k) wc[k] = Wc[k * kHidden + c];
  const float bmc = bm[c];
  float dwc[kEdge + 1];
#pragma unroll
  for (int k = 0; k <= kEdge; ++k) dwc[k] = 0.0f;
  const uint32_t rBase = blockIdx.x * kChunk;
  for (uint32_t row = rBase + wave; row < rBase + kChunk; row += 4u) {
    const uint32_t e0 = g.rowStart[row], e1 = g.rowStart[row + 1];
    const float dmi = e1 > e0 ? dHm[size_t(row) * 2 * kHidden + kHidden + c] / float(e1 - e0) : 0.0f;
    float dP = 0.0f, dQ = 0.0f;
    for (uint32_t e = e0; e < e1; ++e) {
      const uint32_t j = g.nbr[e];
      float ef[kEdge];
      edgeFeatures(g, f, row, j, ef);
      if (edgePre(pq, row, j, c, ef, wc, bmc) > 0.0f) {
        dP += dmi;
#pragma unroll
        for (int k = 0; k < kEdge; ++k) dwc[k] += ef[k] * dmi;
        dwc[kEdge] += dmi;
      }
      // Edge j -> i: row i is the neighbour whose Q enters z_ji.
      const uint32_t dj = g.rowStart[j + 1] - g.rowStart[j];
      float eb[kEdge];
      edgeFeatures(g, f, j, row, eb);
      if (edgePre(pq, j, row, c, eb, wc, bmc) > 0.0f) dQ += dHm[size_t(j) * 2 * kHidden + kHidden + c] / float(dj);
    }
    dPq[size_t(row) * 2 * kHidden + c] = dP;
    dPq[size_t(row) * 2 * kHidden + kHidden + c] = dQ;
    dPqBf[size_t(row) * 2 * kHidden + c] = toBf16(dP);
    dPqBf[size_t(row) * 2 * kHidden + kHidden + c] = toBf16(dQ);
  }
#pragma unroll
  for (int k = 0; k <= kEdge; ++k) red[wave][k][c] = dwc[k];
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int k = 0; k <= kEdge; ++k) {
      const float s = ((red[0][k][c] + red[1][k][c]) + red[2][k][c]) + red[3][k][c];
      partial[size_t(blockIdx.x) * (kEdge + 1) * kHidden + k * kHidden + c] = s;
    }
  }
}

// dh = dh + dHm[:, 0:64] + dhPq.
__global__ void __launch_bounds__(256) kAccumulateDh(float* __restrict__ dh, const float* __restrict__ dHm,
                                                     const float* __restrict__ dhPq, uint32_t N) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * kHidden) return;
  const uint32_t row = i / kHidden, c = i % kHidden;
  dh[i] = (dh[i] + dHm[size_t(row) * 2 * kHidden + c]) + dhPq[i];
}

// Adam (FP32 master weights).
__global__ void kAdam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m, float* __restrict__ v,
                      uint32_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = gr[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of a K x M weight: W (row-major [K][M]) and Wt ([M][K]).
__global__ void kPack(const float* __restrict__ W, uint32_t K, uint32_t M, uint16_t* __restrict__ w,
                      uint16_t* __restrict__ wt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * M) return;
  const uint32_t k = i / M, m = i % M;
  const uint16_t b = toBf16(W[i]);
  if (w) w[i] = b;
  if (wt) wt[size_t(m) * K + k] = b;
}

// ---------------------------------------------------------------------------------------------------------------
// Host side.

namespace {

inline uint32_t blocks(uint64_t n, uint32_t t = 256) { return uint32_t((n + t - 1) / t); }

// Offsets (in BF16 elements) of the packed weights inside wbf_.
struct PackLayout {
  size_t WeT, WdT, Wd, WabT[kRounds], Wab[kRounds], WuT[kRounds], Wu[kRounds], total;
  PackLayout() {
    size_t o = 0;
    auto take = [&](size_t n) {
      const size_t r = o;
      o += (n + 7) & ~size_t(7);  // keep 16-byte alignment
      return r;
    };
    WeT = take(kIn * kHidden);
    WdT = take(kHidden * kOut);
    Wd = take(kHidden * kOut);
    for (int r = 0; r < kRounds; ++r) {
      WabT[r] = take(kHidden * 2 * kHidden);
      Wab[r] = take(kHidden * 2 * kHidden);
      WuT[r] = take(2 * kHidden * kHidden);
      Wu[r] = take(2 * kHidden * kHidden);
    }
    total = o;
  }
};
const PackLayout kPackLayout;

}  // namespace

ParamLayout::ParamLayout() {
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t r = o;
    o += n;
    return r;
  };
  We = take(kIn * kHidden);
  be = take(kHidden);
  for (int r = 0; r < kRounds; ++r) {
    Wab[r] = take(kHidden * 2 * kHidden);
    Wc[r] = take(kEdge * kHidden);
    bm[r] = take(kHidden);
    Wu[r] = take(2 * kHidden * kHidden);
    bu[r] = take(kHidden);
  }
  Wd = take(kHidden * kOut);
  bd = take(kOut);
  total = o;
}

Graph::Graph(uint32_t n, const uint32_t* rowStart, const uint32_t* nbr, const double* rest2, const uint32_t* rowMask)
    : n_(n) {
  nRows_ = std::max<uint32_t>(256u, (n + 255u) / 256u * 256u);
  std::vector<uint32_t> rs(nRows_ + 1u, rowStart[n]), mask(nRows_, 0u);
  for (uint32_t i = 0; i <= n; ++i) rs[i] = rowStart[i];
  for (uint32_t i = 0; i < n; ++i) mask[i] = rowMask ? (rowMask[i] ? 1u : 0u) : 1u;
  const uint32_t ne = rowStart[n];
  for (uint32_t e = 0; e < ne; ++e)
    if (nbr[e] >= n) throw std::invalid_argument("graph neighbour index out of range");
  std::vector<float> r2(2 * size_t(nRows_), 0.0f);
  for (size_t i = 0; i < 2 * size_t(n); ++i) r2[i] = float(rest2[i]);
  rowStart_.allocate(sizeof(uint32_t) * rs.size());
  nbr_.allocate(sizeof(uint32_t) * std::max<uint32_t>(ne, 1u));
  rest2_.allocate(sizeof(float) * r2.size());
  mask_.allocate(sizeof(uint32_t) * mask.size());
  check(hipMemcpy(rowStart_.as<void>(), rs.data(), sizeof(uint32_t) * rs.size(), hipMemcpyHostToDevice), "copy");
  if (ne) check(hipMemcpy(nbr_.as<void>(), nbr, sizeof(uint32_t) * ne, hipMemcpyHostToDevice), "copy");
  check(hipMemcpy(rest2_.as<void>(), r2.data(), sizeof(float) * r2.size(), hipMemcpyHostToDevice), "copy");
  check(hipMemcpy(mask_.as<void>(), mask.data(), sizeof(uint32_t) * mask.size(), hipMemcpyHostToDevice), "copy");
  check(hipStreamSynchronize(nullptr), "sync");  // null-stream work done before any other stream reads it
}

GraphView Graph::view() const {
  return GraphView{nRows_, rowStart_.as<uint32_t>(), nbr_.as<uint32_t>(), rest2_.as<float>(), mask_.as<uint32_t>()};
}

Net::Net(uint64_t seed) {
  const size_t n = layout_.total;
  std::vector<float> p(n, 0.0f);
  std::mt19937_64 rng(seed);
  auto init = [&](size_t off, int fanIn, int rows, int cols, int usedRows, int usedCols) {
    // He-style uniform initialisation; unused rows / columns stay zero.
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    const float a = std::sqrt(3.0f / float(fanIn));
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) p[off + size_t(i) * cols + j] = (i < usedRows && j < usedCols) ? a * U(rng) : 0.0f;
  };
  init(layout_.We, 10, kIn, kHidden, 10, kHidden);
  for (int r = 0; r < kRounds; ++r) {
    init(layout_.Wab[r], kHidden, kHidden, 2 * kHidden, kHidden, 2 * kHidden);
    init(layout_.Wc[r], kEdge, kEdge, kHidden, kEdge, kHidden);
    init(layout_.Wu[r], 2 * kHidden, 2 * kHidden, kHidden, 2 * kHidden, kHidden);
    for (int c = 0; c < kHidden; ++c) p[layout_.Wu[r] + size_t(c) * kHidden + c] *= 0.1f;  // small residual branch
  }
  init(layout_.Wd, kHidden, kHidden, kOut, kHidden, kOutUsed);
  for (size_t i = layout_.Wd; i < layout_.Wd + size_t(kHidden) * kOut; ++i) p[i] *= 0.1f;
  check(hipStreamCreateWithFlags(&packStream_, hipStreamNonBlocking), "stream");
  params_.allocate(sizeof(float) * n);
  grad_.allocate(sizeof(float) * n);
  adamM_.allocate(sizeof(float) * n);
  adamV_.allocate(sizeof(float) * n);
  wbf_.allocate(sizeof(uint16_t) * kPackLayout.total);
  check(hipMemcpy(params_.as<void>(), p.data(), sizeof(float) * n, hipMemcpyHostToDevice), "copy");
  check(hipStreamSynchronize(nullptr), "sync");  // the copy is on the null stream, the packing on its own
  packWeights(packStream_);
  check(hipStreamSynchronize(packStream_), "sync");
}

Net::~Net() {
  if (packStream_) {
    (void)hipStreamSynchronize(packStream_);
    (void)hipStreamDestroy(packStream_);
  }
}

void Net::getParams(float* out) const {
  check(hipMemcpy(out, params_.as<void>(), sizeof(float) * layout_.total, hipMemcpyDeviceToHost), "copy");
}

void Net::setParams(const float* in) {
  check(hipMemcpy(params_.as<void>(), in, sizeof(float) * layout_.total, hipMemcpyHostToDevice), "copy");
  check(hipStreamSynchronize(nullptr), "sync");  // null-stream work done before any other stream reads it
  packWeights(packStream_);
  check(hipStreamSynchronize(packStream_), "sync");
}

void Net::resetOptimizer() {
  check(hipMemset(adamM_.as<void>(), 0, sizeof(float) * layout_.total), "memset");
  check(hipMemset(adamV_.as<void>(), 0, sizeof(float) * layout_.total), "memset");
  check(hipStreamSynchronize(nullptr), "sync");  // null-stream work done before any other stream reads it
  adamStep_ = 0;
}

void Net::getGradient(float* out) const {
  check(hipMemcpy(out, grad_.as<void>(), sizeof(float) * layout_.total, hipMemcpyDeviceToHost), "copy");
}

void Net::packWeights(hipStream_t st) {
  const float* P = params_.as<float>();
  uint16_t* B = wbf_.as<uint16_t>();
  const PackLayout& L = kPackLayout;
  kPack<<<blocks(kIn * kHidden), 256, 0, st>>>(P + layout_.We, kIn, kHidden, nullptr, B + L.WeT);
  kPack<<<blocks(kHidden * kOut), 256, 0, st>>>(P + layout_.Wd, kHidden, kOut, B + L.Wd, B + L.WdT);
  for (int r = 0; r < kRounds; ++r) {
    kPack<<<blocks(kHidden * 2 * kHidden), 256, 0, st>>>(P + layout_.Wab[r], kHidden, 2 * kHidden, B + L.Wab[r],
                                                         B + L.WabT[r]);
    kPack<<<blocks(2 * kHidden * kHidden), 256, 0, st>>>(P + layout_.Wu[r], 2 * kHidden, kHidden, B + L.Wu[r],
                                                         B + L.WuT[r]);
  }
  check(hipGetLastError(), "pack launch");
}

void Net::ensureCapacity(uint32_t nRows) {
  if (nRows <= capRows_) return;
  const size_t n = nRows;
  feat_.allocate(2 * n * kIn);
  hEnc_.allocate(4 * n * kHidden);
  hF32_.allocate(4 * n * kHidden);
  for (int r = 0; r < kRounds; ++r) {
    hm_[r].allocate(2 * n * 2 * kHidden);
    pq_[r].allocate(4 * n * 2 * kHidden);
    u_[r].allocate(4 * n * kHidden);
  }
  h3_.allocate(2 * n * kHidden);
  out_.allocate(4 * n * kOut);
  acc_.allocate(4 * n * 2 * kHidden);
  dOut_.allocate(4 * n * kOut);
  dOutBf_.allocate(2 * n * kOut);
  dh_.allocate(4 * n * kHidden);
  dpre_.allocate(4 * n * kHidden);
  dpreBf_.allocate(2 * n * kHidden);
  dHm_.allocate(4 * n * 2 * kHidden);
  dPq_.allocate(4 * n * 2 * kHidden);
  dPqBf_.allocate(2 * n * 2 * kHidden);
  dhPq_.allocate(4 * n * kHidden);
  const size_t chunks = n / kChunk;
  partial_.allocate(4 * chunks * size_t(2 * kHidden * kHidden));  // largest partial: 128 x 64 per chunk
  lossPart_.allocate(8 * blocks(n));
  capRows_ = nRows;
}

void Net::forward(const GraphView& g, const FrameView& f, hipStream_t st) {
  const uint32_t N = g.nRows;
  if (N % 256u != 0u) throw std::invalid_argument("graph rows must be a multiple of 256");
  ensureCapacity(N);
  const float* P = params_.as<float>();
  const uint16_t* B = wbf_.as<uint16_t>();
  const PackLayout& L = kPackLayout;
  const uint32_t gemmBlocks = N / 64u;
  kFeatures<<<blocks(N), 256, 0, st>>>(g, f, feat_.as<uint16_t>());
  kGemm<kIn, kHidden><<<gemmBlocks, 256, 0, st>>>(feat_.as<uint16_t>(), kIn, B + L.WeT, hF32_.as<float>(), kHidden, N);
  kBiasRelu<<<blocks(size_t(N) * kHidden), 256, 0, st>>>(hF32_.as<float>(), P + layout_.be, hm_[0].as<uint16_t>(),
                                                           2 * kHidden, hEnc_.as<float>(), N);
  for (int r = 0; r < kRounds; ++r) {
    uint16_t* hm = hm_[r].as<uint16_t>();
    kGemm<kHidden, 2 * kHidden><<<gemmBlocks, 256, 0, st>>>(hm, 2 * kHidden, B + L.WabT[r], pq_[r].as<float>(),
                                                            2 * kHidden, N);
    kMessage<<<N / 4u, 256, 0, st>>>(g, f, pq_[r].as<float>(), P + layout_.Wc[r], P + layout_.bm[r], hm);
    kGemm<2 * kHidden, kHidden><<<gemmBlocks, 256, 0, st>>>(hm, 2 * kHidden, B + L.WuT[r], acc_.as<float>(), kHidden, N);
    uint16_t* next = r + 1 < kRounds ? hm_[r + 1].as<uint16_t>() : h3_.as<uint16_t>();
    const int ldNext = r + 1 < kRounds ? 2 * kHidden : kHidden;
    kUpdate<<<blocks(size_t(N) * kHidden), 256, 0, st>>>(acc_.as<float>(), P + layout_.bu[r], u_[r].as<float>(),
                                                          hF32_.as<float>(), next, ldNext, N);
  }
  kGemm<kHidden, kOut><<<gemmBlocks, 256, 0, st>>>(h3_.as<uint16_t>(), kHidden, B + L.WdT, out_.as<float>(), kOut, N);
  check(hipGetLastError(), "forward launch");
}

void Net::predict(const GraphView& g, const FrameView& f, dk::F3* dv, hipStream_t st) {
  forward(g, f, st);
  kDecode<<<blocks(g.nRows), 256, 0, st>>>(g, out_.as<float>(), params_.as<float>() + layout_.bd, dv);
  check(hipGetLastError(), "decode launch");
}

double Net::lossAndGradient(const GraphView& g, const FrameView& f, const dk::F3* target, float lr, hipStream_t st) {
  forward(g, f, st);
  const uint32_t N = g.nRows;
  const float* P = params_.as<float>();
  float* G = grad_.as<float>();
  const uint16_t* B = wbf_.as<uint16_t>();
  const PackLayout& L = kPackLayout;
  const uint32_t gemmBlocks = N / 64u, chunks = N / kChunk;
  float* part = partial_.as<float>();
  kDecode<<<blocks(N), 256, 0, st>>>(g, out_.as<float>(), P + layout_.bd, nullptr);
  // Number of rows in the loss.
  std::vector<uint32_t> mask(N);
  check(hipMemcpyAsync(mask.data(), g.rowMask, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, st), "copy");
  check(hipStreamSynchronize(st), "sync");
  uint32_t count = 0;
  for (uint32_t m : mask) count += m ? 1u : 0u;
  const float invCount = 1.0f / float(std::max<uint32_t>(count, 1u));
  const uint32_t lossBlocks = blocks(N);
  kLoss<<<lossBlocks, 256, 0, st>>>(g, out_.as<float>(), target, invCount, dOut_.as<float>(), dOutBf_.as<uint16_t>(),
                                    lossPart_.as<double>());
  // Decoder: dWd = h3^T dOut, dbd = colsum(dOut), dh = dOut Wd^T.
  kGemmTN<kHidden, kOut><<<chunks, 256, 0, st>>>(h3_.as<uint16_t>(), kHidden, dOutBf_.as<uint16_t>(), kOut, part);
  kReduceChunks<<<blocks(kHidden * kOut), 256, 0, st>>>(part, chunks, kHidden * kOut, G + layout_.Wd, 0);
  kColSum<<<chunks, 256, 0, st>>>(dOut_.as<float>(), kOut, kOut, part);
  kReduceChunks<<<1, 256, 0, st>>>(part, chunks, kOut, G + layout_.bd, 0);
  kGemm<kOut, kHidden><<<gemmBlocks, 256, 0, st>>>(dOutBf_.as<uint16_t>(), kOut, B + L.Wd, dh_.as<float>(), kHidden, N);
  for (int r = kRounds - 1; r >= 0; --r) {
    const uint16_t* hm = hm_[r].as<uint16_t>();
    // Update: dpre = dh * (U > 0); dWu = [h | m]^T dpre; dbu = colsum(dpre); dHm = dpre Wu^T.
    kReluBack<<<blocks(size_t(N) * kHidden), 256, 0, st>>>(dh_.as<float>(), u_[r].as<float>(), dpre_.as<float>(),
                                                            dpreBf_.as<uint16_t>(), N);
    kGemmTN<2 * kHidden, kHidden><<<chunks, 256, 0, st>>>(hm, 2 * kHidden, dpreBf_.as<uint16_t>(), kHidden, part);
    kReduceChunks<<<blocks(2 * kHidden * kHidden), 256, 0, st>>>(part, chunks, 2 * kHidden * kHidden,
                                                                 G + layout_.Wu[r], 0);
    kColSum<<<chunks, 256, 0, st>>>(dpre_.as<float>(), kHidden, kHidden, part);
    kReduceChunks<<<1, 256, 0, st>>>(part, chunks, kHidden, G + layout_.bu[r], 0);
    kGemm<kHidden, 2 * kHidden><<<gemmBlocks, 256, 0, st>>>(dpreBf_.as<uint16_t>(), kHidden, B + L.Wu[r],
                                                            dHm_.as<float>(), 2 * kHidden, N);
    // Message: dPq, and dWc followed by dbm (contiguous in the parameter vector).
    kMessageBack<<<chunks, 256, 0, st>>>(g, f, pq_[r].as<float>(), P + layout_.Wc[r], P + layout_.bm[r],
                                         dHm_.as<float>(), dPq_.as<float>(), dPqBf_.as<uint16_t>(), part);
    kReduceChunks<<<blocks((kEdge + 1) * kHidden), 256, 0, st>>>(part, chunks, (kEdge + 1) * kHidden,
                                                                 G + layout_.Wc[r], 0);
    // Wab: dWab = h^T dPq; dhPq = dPq Wab^T; dh += dHm[:, :64] + dhPq.
    kGemmTN<kHidden, 2 * kHidden><<<chunks, 256, 0, st>>>(hm, 2 * kHidden, dPqBf_.as<uint16_t>(), 2 * kHidden, part);
    kReduceChunks<<<blocks(kHidden * 2 * kHidden), 256, 0, st>>>(part, chunks, kHidden * 2 * kHidden,
                                                                 G + layout_.Wab[r], 0);
    kGemm<2 * kHidden, kHidden><<<gemmBlocks, 256, 0, st>>>(dPqBf_.as<uint16_t>(), 2 * kHidden, B + L.Wab[r],
                                                            dhPq_.as<float>(), kHidden, N);
    kAccumulateDh<<<blocks(size_t(N) * kHidden), 256, 0, st>>>(dh_.as<float>(), dHm_.as<float>(), dhPq_.as<float>(), N);
  }
  // Encoder: dpre = dh * (h0 > 0); dWe = f^T dpre; dbe = colsum(dpre).
  kReluBack<<<blocks(size_t(N) * kHidden), 256, 0, st>>>(dh_.as<float>(), hEnc_.as<float>(), dpre_.as<float>(),
                                                          dpreBf_.as<uint16_t>(), N);
  kGemmTN<kIn, kHidden><<<chunks, 256, 0, st>>>(feat_.as<uint16_t>(), kIn, dpreBf_.as<uint16_t>(), kHidden, part);
  kReduceChunks<<<blocks(kIn * kHidden), 256, 0, st>>>(part, chunks, kIn * kHidden, G + layout_.We, 0);
  kColSum<<<chunks, 256, 0, st>>>(dpre_.as<float>(), kHidden, kHidden, part);
  kReduceChunks<<<1, 256, 0, st>>>(part, chunks, kHidden, G + layout_.be, 0);
  check(hipGetLastError(), "backward launch");
  std::vector<double> lp(lossBlocks);
  check(hipMemcpyAsync(lp.data(), lossPart_.as<void>(), sizeof(double) * lossBlocks, hipMemcpyDeviceToHost, st), "copy");
  check(hipStreamSynchronize(st), "sync");
  double loss = 0.0;
  for (double x : lp) loss += x;
  loss *= double(invCount);
  if (lr > 0.0f) {
    ++adamStep_;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, float(adamStep_)), c2 = 1.0f - std::pow(b2, float(adamStep_));
    kAdam<<<blocks(layout_.total), 256, 0, st>>>(params_.as<float>(), G, adamM_.as<float>(), adamV_.as<float>(),
                                                  uint32_t(layout_.total), lr, b1, b2, 1e-8f, c1, c2);
    packWeights(st);
    check(hipStreamSynchronize(st), "sync");
  }
  return loss;
}

}  // namespace ws
