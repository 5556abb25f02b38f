// solver.hpp — the HIP implementation of the reference's Solver surface (libs/sim/include/drape/sim/solver.hpp:85-98)
// on one MI355X: device state, structure build and the Newton–PCG substep orchestration.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "device_buffer.hpp"
#include "kernels.hpp"
#include "model.hpp"
#include "normals.hpp"
#include "warmstart.hpp"

namespace dh {

class HipSolver {
 public:
  HipSolver();
  ~HipSolver();

  void build(const drape_scene_desc* scenes, uint32_t nScenes);
  void updateFabric(const std::string& garment, const std::vector<Fabric>& fabrics);
  void setPhase(const std::string& garment, int32_t phase);
  void step(double dt, int32_t substeps, int32_t iterations);
  void readPositions(float* out, size_t n) const;
  void readVelocities(float* out, size_t n) const;
  drape_stats stats() const;
  // Positions and unit vertex normals (every garment's triangles, seam pairs welded) in particle order.
  void readPositionsNormals(float* positions, float* normals, size_t n) const;
  // Stats of segments [first, first + count) (one scene of a batched build each), from one device reduction.
  void segmentStats(uint32_t first, uint32_t count, drape_stats* out) const;
  size_t snapshotSize() const;
  void snapshotSave(void* out, size_t bytes) const;
  void snapshotRestore(const void* in, size_t bytes);
  void range(const std::string& garment, uint32_t& offset, uint32_t& count) const;
  size_t particleCount() const { return model_.x.size(); }
  uint32_t segmentCount() const { return static_cast<uint32_t>(model_.segments.size()); }
  void segmentRange(uint32_t s, uint32_t& offset, uint32_t& count) const;
  drape_counters counters() const;
  drape_pcg_report pcgReport() const;
  void resetCounters();
  void setGraph(bool enabled) { useGraph_ = enabled; }
  // 1: one kernel and one reduction per PCG iteration (default); 0: two kernels per iteration.
  void setPcgVariant(int variant);
  // FP64 true-residual refinement of the one-kernel PCG: at most `restarts` continuations of a solve from its true
  // residual (0: off).
  void setRefinement(int restarts) { refineRestarts_ = static_cast<uint32_t>(std::max(0, restarts)); }
  int refinement() const { return static_cast<int>(refineRestarts_); }
  // Wavefronts per 64-row slice in the one-kernel PCG iteration: 1, 2 or 4, or 0 to time all three during the
  // first solves and keep the fastest (results are bit-identical for every choice).
  // Launch shape of the one-kernel PCG iteration (wavefronts per slice x slots per batch; 0x0 = tune); halo 2 = the
  // persistent solve (one launch per solve, 1x8 with the halo).
  void setPcgShape(int waves, int slots, int halo);
  void pcgShape(int& waves, int& slots, int& halo) const;  // 0x0 while still tuning
  // Polls of a persistent solve's barrier before a workgroup gives up (the solve is then re-run with one launch per
  // iteration, and the persistent form is not used again by this solver).
  void setSolveSpinCap(uint32_t cap) { solveSpinCap_ = cap; }
  // The persistent solve fits this build (at most 16 + 16 slots per slice; one tile per CU: each tile's columns
  // within kSolveWindows row ranges of kHaloRows rows in all and at most kSolveExtraCap slots past the eighth; tile
  // pairs where the tiles outnumber the CUs: each pair's within kPairWindows ranges of kPairRows rows and twice the
  // slots; each segment's tiles resident together) and has not given up.
  bool persistentAvailable() const { return persistentOk_; }
  uint64_t persistentAborts() const { return persistentAborts_; }
  // Every `every`-th PCG SpMV launch of the step is bracketed by HIP events (0 = off).
  void setSpmvSampling(int every);
  void spmvTiming(uint64_t& launches, double& meanUs, double& bytesPerLaunch) const;

  drape_structure_sizes structureSizes() const;
  void readStructure(uint32_t* rowStart, uint32_t* colIndex, uint32_t* diagSlot, uint32_t* stretchSlots,
                     uint32_t* hingeSlots, uint32_t* seamSlots) const;
  void readTerms(uint32_t* stretchV, uint32_t* hingeV, double* hingeK, uint32_t* seamAB) const;
  // The assembly's summation order as gather lists in the reference's contribution numbering.
  void gatherLists(std::vector<uint32_t>& gradStart, std::vector<uint32_t>& gradIdx, std::vector<uint32_t>& blockStart,
                   std::vector<uint32_t>& blockIdx) const;
  void debugAssemble(double h, const float* u, double* energy, float* grad, float* blocks);
  void debugSpmv(const float* x, float* y);
  void debugSystem(float* blocks, float* grad, float* precond, uint32_t* contactFlags, float* contactNormals);

  // Learned warm start: the network predicts each step's velocity residual beyond free flight, which seeds Newton
  // at the inertial target plus h times it; null restores the previous-velocity guess. The solver does not own
  // the network.
  void setWarmStart(ws::Net* net);
  // Network inputs of the current state in particle order: x, v, the last substep's realised dv (3 each) and the
  // SDF distance and normal (4: normal, distance). Any output may be null.
  void frameInputs(float* x, float* v, float* dvPrev, float* phiNormal);
  // The mesh edge graph the network runs on (rows = internal rows), as host CSR in particle order.
  void meshGraph(std::vector<uint32_t>& rowStart, std::vector<uint32_t>& nbr) const;

 private:
  void requireBuilt() const;
  void uploadSegParams(double h);
  void uploadGarments();
  void uploadFabricsAndMass();
  void uploadHinges();
  void uploadSeams();
  void syncStateFromDevice() const;   // x, v, contact force -> host model
  void syncStateToDevice();           // host model -> x, v, contact force
  void pollStates() const;            // device SegState -> hostStates_
  std::vector<double> segmentPartials() const;  // launchStats -> host, kStatsPerSegment per segment
  void toDevice(void* dst, const void* src, size_t bytes) const;
  void fromDevice(void* dst, const void* src, size_t bytes) const;
  template <class T>
  void upload(DeviceBuffer& b, const std::vector<T>& data, size_t count = 0);
  bool allSegments(bool (*pred)(const dk::SegState&)) const;
  void substep(double h, int iterations);
  void runPcg();
  void runPcgStandard();
  void runPcgSingleKernel();
  void releaseGraph();
  void timedPcgStep(uint32_t parity, dk::PcgShape shape, bool solveStarted);
  int shapeIndex() const { return pcgShape_ >= 0 ? pcgShape_ : tunedShape_; }
  std::pair<hipEvent_t, hipEvent_t> takeEvents();
  void runPcgTuning(uint32_t maxCg);
  // Runs the rest of the current solve (and its refinement continuations when `refine`) with launchPcgSolve; false
  // if a barrier gave up or the launch failed (the persistent form is then switched off).
  bool runPcgPersistent(bool refine, bool timeIt);
  void redoSolveWithLaunches(uint32_t maxCg);
  void globalStructure(std::vector<uint32_t>& rowStart, std::vector<uint32_t>& colIndex,
                       std::vector<uint32_t>& diagSlot, std::vector<uint32_t>& internalToGlobalSlot) const;

  hipStream_t stream_ = nullptr;
  Model model_;
  bool built_ = false;
  bool useGraph_ = true;
  double lastH_ = -1.0;
  uint32_t nRows_ = 0, nTiles_ = 0, nSlices_ = 0, nBlocks_ = 0, maxK_ = 0;
  uint32_t maxLowK_ = 0;
  bool uniformLayout_ = false;  // every slice has maxK_ upper and maxLowK_ lower slots
  uint64_t sellFloats_ = 0, lowEntries_ = 0;
  uint32_t lastPcgIterations_ = 0;
  std::vector<uint32_t> rowOf_;  // global particle -> internal row
  dk::View view_{};

  // Device buffers.
  DeviceBuffer rowFlags_, mass_, rowFabric_, rowGarment_, fExt_, pinPos_;
  DeviceBuffer x_, v_, xStart_, yRel_, u_, grad_, step_, r_, z_, pA_, pB_, ap_, warm_, precond_;
  DeviceBuffer cnormal_, cflags_, cforce_, clambda_;
  DeviceBuffer dvLast_, phiN_;
  DeviceBuffer massF64_, statsTile_, statsSeg_;  // device stats: FP64 masses, per-tile and per-segment partials
  std::unique_ptr<ws::Graph> meshGraph_;
  mutable std::unique_ptr<nrm::Mesh> normalMesh_;  // built on the first normals request after a build
  mutable DeviceBuffer normalsOut_;
  ws::Net* net_ = nullptr;
  DeviceBuffer tileRange_, tileHalo_, tileWindows_;
  DeviceBuffer tileSeg_, segParams_, segState_, partA_, partB_, partC_, partE_, partMax_;
  DeviceBuffer stretch_, hinges_, seams_, fabrics_, garments_;
  DeviceBuffer incCount_, incStart_, inc_, rowLen_, rowStart_, colIndex_, diagSlot_;
  DeviceBuffer stretchSlots_, hingeSlots_, seamSlots_, sellK_, sellBase_, sellCol_, sellVal_, buildError_;
  DeviceBuffer lowK_, lowBase_, lowCol_, lowBlk_;
  DeviceBuffer fieldValues_, scratchF3_, scratchF3b_;
  DeviceBuffer cgBuf_[10], cgPart0_, cgPart1_, tileDone_;
  int pcgVariant_ = 1;
  uint32_t refineRestarts_ = 4;
  static constexpr int kShapeCount = dk::kPcgShapeCount;
  int pcgShape_ = -1;     // requested index into dk::kPcgShapes (-1 = tune)
  int tunedShape_ = -1;   // chosen by the tuner (-1 until it has measured every shape)
  double tuneMs_[kShapeCount] = {};
  bool tuneWarm_[kShapeCount] = {};  // the shape's first timed sample was taken (and not counted: it loads code)
  uint32_t tuneLaunches_[kShapeCount] = {};
  bool shapeAvailable(int i) const {
    return dk::kPcgShapes[i].persistent ? persistentOk_ : (!dk::kPcgShapes[i].halo || view_.tileHalo != nullptr);
  }
  // Persistent solve: counters (per segment dk::solveCounterLines lines of 64 words, then the abort flag), each
  // segment's first line, the abort flag's pinned host mirror.
  DeviceBuffer solveCtr_, solveSegCtr_;
  uint32_t solveCtrLines_ = 0;
  uint32_t* solveAbortHost_ = nullptr;
  bool persistentOk_ = false;
  bool solveQueued_ = false;                       // more tiles than resident workgroups: segments queue
  bool solveExtra_ = false;                        // some slice has more than 8 upper or lower slots
  uint32_t tileExtrasMax_ = 0;                     // most slots past the eighth in one tile's rows
  bool solvePairs_ = false;                        // the persistent solve runs by tile pairs (kPcgSolvePair)
  bool pairsFit_ = false;                          // every pair's window and extra slots fit kPcgSolvePair
  uint32_t nPairs_ = 0;
  DeviceBuffer pairTiles_, pairWindows_;
  std::vector<std::pair<uint32_t, uint32_t>> segPairs_;  // per segment: its pair workgroups [first, end)
  DeviceBuffer solveOrder_;                        // the queue's tile order (device)
  std::vector<uint32_t> solveOrderHost_;           // the order last uploaded
  std::vector<std::pair<uint32_t, uint32_t>> segTiles_;  // per segment: first tile, end tile
  std::vector<uint32_t> segLastIters_;             // per segment: PCG iterations of its last persistent solve
  void queueSegments();
  uint64_t persistentAborts_ = 0;
  uint32_t solveSpinCap_ = 1u << 20;
  // PCG iterations per captured graph, i.e. between two looks at the convergence flags after a solve's first
  // batch (even: the parity pattern repeats). Measured on the 65k sheet over 8 ... 64: a look at the flags costs
  // more than the launches that find the solve already ended.
  static constexpr uint32_t kGraphChunk = 64;
  // The tuner times chunks of 16 launches (a solve must outlast a chunk for it to count: with the reference's
  // default cap of 50 iterations a 64-launch chunk never would) and gives up waiting for full samples after
  // kTuneSolves solves.
  static constexpr uint32_t kTuneChunk = 16;
  static constexpr uint32_t kTuneSolves = 64;
  uint32_t tuneSolves_ = 0;
  int fastestSampledShape() const;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graphExec_ = nullptr;
  // One Newton iteration's launches before the PCG solve (activate, assemble, e0) and after it (contact force,
  // slope, kGraphTrials line-search trials, apply, end), each captured once and replayed; the device-side flags
  // make every launch a no-op for a segment that has no work left in it.
  struct Captured {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void release();
  };
  Captured newtonPre_, newtonPost_;
  // The first launches of a continuation from the true residual, replayed behind the refinement decision.
  static constexpr uint32_t kRefineChunk = 16;
  Captured refineExec_;
  Captured tuneExec_[kShapeCount];  // the tuner's chunk of each launch shape
  int refineShape_ = -1;
  static constexpr int kGraphTrials = 2;
  template <class F>
  void launchCaptured(Captured& c, F&& launches);
  int graphShape_ = -1;  // shape index the captured graph launches
  dk::SegState* hostStates_ = nullptr;  // pinned host copy
  int spmvEvery_ = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> eventPool_, eventPending_;
  mutable uint64_t spmvTimed_ = 0;
  mutable double spmvTotalMs_ = 0.0, spmvTotalBytes_ = 0.0;
  std::vector<double> eventBytes_;        // algorithmic bytes of each pending timed launch
  mutable bool mirrorFresh_ = false;      // the host mirror was read during the current substep
  std::vector<double> segBlocks_, segParticles_;  // per segment: stored blocks of its particle rows, particles
  double launchBytes(double blocks, double particles) const;
  // Bytes of a PCG launch over the segments still solving at the last poll (at a solve's start: not finished).
  double activeLaunchBytes(bool solveStarted) const;
  void pushTimed(std::pair<hipEvent_t, hipEvent_t> pr, bool solveStarted);
  void harvestEvents() const;
};

}  // namespace dh
