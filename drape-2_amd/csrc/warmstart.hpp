// warmstart.hpp — learned warm start of the Newton solve: a per-vertex message-passing network over the mesh edge
// graph that predicts the velocity change dv of the coming step, so that the Newton guess u = h (v + dv) replaces
// the reference's previous-velocity guess u = h v (cpu_solver.cpp:244; kernels.h:34-51).
//
// Network (all layers per vertex; hidden width 64):
//   node features f (16, 10 used): v / 1 m/s, previous dv / 0.5 m/s, SDF distance / 2 cm, SDF normal
//   edge features e_ij (8): rest edge / 1 cm (pattern space, 2), current edge / 1 cm (3), relative velocity (3)
//   encoder     h = relu(f We + be)
//   3 rounds    [P | Q] = h Wab;  m_i = mean_j relu(P_i + Q_j + e_ij Wc + bm);  h += relu([h | m] Wu + bu)
//   decoder     dv / 0.5 m/s = h Wd + bd   (3 of 16 columns used)
// The dense layers run as BF16 MFMA GEMMs with FP32 accumulation (weights and activations in BF16, FP32 master
// weights, Adam); the message step and every elementwise stage run in FP32. The backward pass mirrors the forward
// layer by layer; all reductions (weight and bias gradients, loss) are per-workgroup partials summed in a fixed
// order, so training is deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_buffer.hpp"
#include "layout.hpp"

namespace ws {

constexpr int kIn = 16;      // node feature columns (10 used)
constexpr int kEdge = 8;     // edge features
constexpr int kHidden = 64;  // hidden width
constexpr int kRounds = 3;   // message-passing rounds
constexpr int kOut = 16;     // decoder columns (3 used)
constexpr int kOutUsed = 3;

constexpr float kVelScale = 1.0f;    // m/s
constexpr float kDvScale = 0.5f;     // m/s, previous dv feature and output
constexpr float kLenScale = 0.01f;   // m, edge lengths
constexpr float kPhiScale = 0.02f;   // m, SDF distance

// Offsets of the parameter tensors in the flat FP32 parameter vector (row-major [in][out]).
struct ParamLayout {
  size_t We, be, Wd, bd;
  size_t Wab[kRounds], Wc[kRounds], bm[kRounds], Wu[kRounds], bu[kRounds];
  size_t total;
  ParamLayout();
};

// Mesh edge graph on the device. Rows are padded to a multiple of 256; rowMask is 1 for a real particle row.
struct GraphView {
  uint32_t nRows;
  const uint32_t* rowStart;  // nRows + 1
  const uint32_t* nbr;
  const float* rest2;        // 2 per row, pattern space (m)
  const uint32_t* rowMask;
};

// Per-row state the features are computed from (device pointers, nRows entries each).
struct FrameView {
  const dk::F3* x;
  const dk::F3* v;
  const dk::F3* dvPrev;
  const float4* phiNormal;  // SDF distance (m) in w, unit normal in xyz (zero when there is no body)
};

// A graph owned on the device, built from host CSR arrays over n rows (padded to a multiple of 256 with empty,
// masked rows). rowMask (may be null: every row real) marks the rows that are particles.
class Graph {
 public:
  Graph(uint32_t n, const uint32_t* rowStart, const uint32_t* nbr, const double* rest2,
        const uint32_t* rowMask = nullptr);
  GraphView view() const;
  uint32_t particles() const { return n_; }
  uint32_t rows() const { return nRows_; }

 private:
  uint32_t n_ = 0, nRows_ = 0;
  dh::DeviceBuffer rowStart_, nbr_, rest2_, mask_;
};

class Net {
 public:
  explicit Net(uint64_t seed);
  ~Net();
  size_t paramCount() const { return layout_.total; }
  void getParams(float* out) const;
  void setParams(const float* in);
  void resetOptimizer();

  // Inference: dv (m/s) per row into `dv` (device, nRows entries; zero on masked rows).
  void predict(const GraphView& g, const FrameView& f, dk::F3* dv, hipStream_t st);
  // Loss (mean squared error of dv / kDvScale over masked rows, 3 components) and parameter gradients; when
  // lr > 0 an Adam step follows. target: dv (m/s) per row, device.
  double lossAndGradient(const GraphView& g, const FrameView& f, const dk::F3* target, float lr, hipStream_t st);
  void getGradient(float* out) const;

 private:
  void ensureCapacity(uint32_t nRows);
  void packWeights(hipStream_t st);
  void forward(const GraphView& g, const FrameView& f, hipStream_t st);

  ParamLayout layout_;
  hipStream_t packStream_ = nullptr;
  uint32_t capRows_ = 0;
  uint64_t adamStep_ = 0;
  dh::DeviceBuffer params_, grad_, adamM_, adamV_;
  dh::DeviceBuffer wbf_;  // BF16 weights in both layouts (see packWeights)
  // activations
  dh::DeviceBuffer feat_, hEnc_, hF32_, hm_[kRounds], pq_[kRounds], u_[kRounds], h3_, out_, acc_;
  // backward
  dh::DeviceBuffer dOut_, dOutBf_, dh_, dpre_, dpreBf_, dHm_, dPq_, dPqBf_, dhPq_, partial_, lossPart_;
};

}  // namespace ws
