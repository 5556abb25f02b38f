// kernels.hpp — device view of the solver state and the host-side launchers of the HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.hpp"

namespace dk {

// Pointers and sizes every kernel reads. Vectors over rows hold nRows + 1 entries: entry nRows is the zero
// "dummy column" that padding and pinned-column SELL entries point at.
struct View {
  uint32_t nRows, nTiles, nSlices, nSegments;
  uint32_t nPairs;  // workgroups of the persistent solve by tile pairs (kPcgSolvePair)
  uint32_t nStretch, nHinge, nSeam;
  uint32_t zeroBlock;  // index (in blocks) of an all-zero block after the last slice's values
  uint32_t uniformK, uniformKL;  // > 0: every slice has this many upper / lower slots (0: see sellK, lowK)
  // per row
  const uint32_t* rowFlags;
  const float* mass;
  const uint32_t* rowFabric;
  const uint32_t* rowGarment;
  const F3* fExt;
  const F3* pinPos;
  F3* x;
  F3* v;
  F3* xStart;
  F3* yRel;
  F3* u;
  F3* grad;
  F3* step;
  F3* r;
  F3* z;
  F3* pA;        // PCG search direction, double-buffered by iteration parity
  F3* pB;
  F3* ap;
  F3* warm;      // learned warm-start velocity change per row (zero when unused)
  F3* dvLast;    // velocity change the last substep realised: u / h - v at its start (zero for pinned rows)
  float4* phiN;  // SDF distance (w) and unit normal (xyz) at x, for the warm-start features
  // One-reduction PCG (kPcgStep): r, ap = S A p, p, zr = Mf r and t = Mf ap, each double-buffered by kernel
  // parity (neighbours read the previous kernel's copy while the owner writes the new one); FP64 partials per
  // tile (7 per parity).
  F3* cgR[2];
  F3* cgAp[2];
  F3* cgP[2];
  F3* cgZr[2];
  F3* cgT[2];
  double* cgPart[2];  // per tile: the PCG's FP64 partial sums, kPartStride doubles apart
  uint32_t* tileDone[2];  // per tile, by kernel parity: 1 once the tile's PCG solve has ended (kPcgStep)
  float* precond;  // kPrecondFloats per row: the symmetric block-Jacobi inverse's upper triangle (00 01 02 11 12 22)
  float4* cnormal; // xyz normal, w = offset r - phi(xStart)
  uint32_t* cflags;
  float* cforce;
  float* clambda;
  // per tile / slice / segment
  const uint32_t* tileSeg;
  const uint32_t* tileRange;  // per tile: its segment's first tile and end tile
  const uint32_t* tileHalo;   // per tile: first row and row count of the window holding all its columns
  // per tile: kSolveWindows (first row, row count) ranges holding all its columns (count 0: unused), for kPcgSolve
  const uint32_t* tileWindows;
  // per pair workgroup of kPcgSolvePair: its first tile | (tiles - 1) << 31, and kPairWindows (first row, row count)
  // ranges holding every column of its tiles' rows (count 0: unused)
  const uint32_t* pairTiles;
  const uint32_t* pairWindows;
  const SegParams* seg;
  SegState* state;
  double* partA;  // per tile
  double* partB;  // per tile
  double* partC;  // per tile (also the true residual's squared product terms, kTrueResidual -> kRefineDecide)
  double* partE;  // per slice
  float* partMax; // per tile
  // terms
  const Stretch32* stretch;
  const Hinge32* hinges;
  const Seam32* seams;
  const Fabric32* fabrics;
  const Garment32* garments;
  // structure
  const uint32_t* incStart;
  const uint32_t* inc;
  const uint32_t* rowStart;
  const uint32_t* colIndex;
  const uint32_t* diagSlot;
  const uint32_t* stretchSlots;
  const uint32_t* hingeSlots;
  const uint32_t* seamSlots;
  // Symmetric SELL-64: the upper part (diagonal first, then the columns above it) holds block values; the lower part
  // holds, per slot, the column and the index of the transposed block in that column's upper part. lowK == nullptr
  // means a general matrix stored whole in the upper arrays.
  const uint32_t* sellK;     // per slice: upper slots
  const uint32_t* sellBase;  // per slice, in floats (column base = sellBase / 9)
  const uint32_t* sellCol;
  float* sellVal;
  const uint32_t* lowK;      // per slice: lower slots
  const uint32_t* lowBase;   // per slice, in entries
  const uint32_t* lowCol;
  const uint32_t* lowBlk;    // block index (float offset / 9) of block (column, row) in sellVal
  const float* fieldValues;
};

// Structure builder inputs/outputs (device pointers).
struct BuildView {
  uint32_t nRows, nStretch, nHinge, nSeam, nSlices;
  const Stretch32* stretch;
  const Hinge32* hinges;
  const Seam32* seams;
  const uint32_t* rowFlags;
  uint32_t* incCount;   // nRows
  uint32_t* incStart;   // nRows + 1
  uint32_t* inc;
  uint32_t* rowLen;     // nRows
  uint32_t* rowStart;   // nRows + 1
  uint32_t* colIndex;
  uint32_t* diagSlot;
  uint32_t* stretchSlots;
  uint32_t* hingeSlots;
  uint32_t* seamSlots;
  uint32_t* sellK;      // nSlices: upper slots
  uint32_t* sellSize;   // nSlices + 1 (scan of K * 576)
  uint32_t* sellCol;
  uint32_t* lowK;       // nSlices: lower slots
  uint32_t* lowSize;    // nSlices + 1 (scan of lowK * 64)
  uint32_t* lowCol;
  uint32_t* lowBlk;
  uint32_t* error;      // nonzero: a row exceeded kMaxIncidence or kMaxRow
};

// ---- launchers (kernels.hip); all asynchronous on `st` ----
// learned: the Newton guess is the inertial target plus h times the network's residual in View::warm (else the
// reference's u = h v).
void launchSubstepBegin(const View& v, bool learned, hipStream_t st);
void launchHoldContacts(const View& v, hipStream_t st);
void launchAssemble(const View& v, uint32_t maxK, bool contact, hipStream_t st);
void launchNewtonBegin(const View& v, hipStream_t st);
void launchPcgInit(const View& v, hipStream_t st);
void launchPcgIteration(const View& v, uint32_t k, hipStream_t st);
// start/stop (optional): events the launch records at kernel start and end (hipExtLaunchKernel).
void launchPcgSpmv(const View& v, uint32_t k, hipStream_t st, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
void launchPcgUpdate(const View& v, uint32_t k, hipStream_t st);
void launchPcgStepInit(const View& v, hipStream_t st);
// SimModel::stats per segment: segOut[6 s + i] = kinetic energy, sum of |v|^2 and count of unpinned rows, then
// the non-finite flag (1.0), the maximum body penetration and the maximum seam gap. tilePart: 6 per tile.
constexpr uint32_t kStatsPerSegment = 6;
void launchStats(const View& v, const double* massF64, double* tilePart, double* segOut, hipStream_t st);
// Launch shape of the one-kernel PCG iteration: `waves` wavefronts share each 64-row slice, each batch covers
// `slots` lower and `slots` upper slots, and `halo` stages the tile's column window in LDS. Supported: 1x8, 1x4,
// 2x8, 4x8, and 1x8 / 1x4 with the halo; the last entry is the persistent solve (launchPcgSolve), which iterates
// as 1x8 with the halo. Every shape gives bit-identical results.
struct PcgShape {
  uint32_t waves = 1, slots = 8;
  bool halo = false;        // gathers from an LDS window of the tile's columns (every tile's window <= kHaloRows rows)
  bool persistent = false;  // one launch per solve (kPcgSolve, kPcgSolvePair) instead of one launch per iteration
  bool operator==(const PcgShape& o) const {
    return waves == o.waves && slots == o.slots && halo == o.halo && persistent == o.persistent;
  }
};
constexpr int kPcgShapeCount = 7;
constexpr int kPcgPersistentShape = 6;
constexpr PcgShape kPcgShapes[kPcgShapeCount] = {{1, 8, false, false}, {1, 4, false, false}, {2, 8, false, false},
                                                 {4, 8, false, false}, {1, 8, true, false},  {1, 4, true, false},
                                                 {1, 8, true, true}};
// One PCG iteration of the one-reduction form.
void launchPcgStep(const View& v, uint32_t parity, PcgShape shape, hipStream_t st, hipEvent_t start = nullptr,
                   hipEvent_t stop = nullptr);
// Persistent solve (kPcgSolve): one launch runs the one-kernel PCG until every segment's solve has ended, with the
// matrix and the rows' private data held in registers and a barrier per segment; bit-identical to launchPcgStep
// with the 1x8 halo shape. It needs every tile's columns within kSolveWindows row ranges of kHaloRows rows in all
// (View::tileWindows), at most 16 + 16 slots per slice (8 + 8 held in registers), and every segment's tiles resident
// at once (pcgSolveCapacity): all tiles of the launch, or, with SolveCtl::order, each segment's (the segments then
// pass through the resident workgroups in that order).
// Distance in doubles between consecutive tiles' PCG partial sums (View::cgPart).
constexpr uint32_t kPartStride = 7;
constexpr uint32_t kSolveWindows = 4;  // row ranges of a tile's gathered window in the persistent solve
// Per tile, the persistent solve's LDS list of matrix slots past the eighth (a row's upper or lower slots 9..16).
constexpr uint32_t kSolveExtraCap = 64;
// The persistent solve by tile pairs: row ranges and rows of a pair's window.
constexpr uint32_t kPairWindows = 8;
constexpr uint32_t kPairRows = 3072;
struct SolveCtl {
  uint32_t* ctr;            // per segment 1 + 2 ceil(tiles / 16) lines of 64 words: top, group arrivals, generations
  const uint32_t* segCtr;   // per segment: its first counter, in units of 64 words
  uint32_t* abortFlag;      // set by a workgroup that gave up waiting (or by the host): every workgroup leaves
  uint32_t spinCap;         // polls of a barrier's generation counter before a workgroup gives up
  // Queued launch (null: one launch of resident tiles): the tile of each workgroup, every segment's tiles
  // consecutive, so that segments run through the device's resident workgroups in this order.
  const uint32_t* order;
};
// Tiles (or tile pairs) per group of the persistent solve's two-level segment barrier.
constexpr uint32_t kBarrierGroup = 16u;
// Counter lines a segment of `tiles` tiles needs: the top counter, and per group its arrivals and its generation.
inline uint32_t solveCounterLines(uint32_t tiles) {
  return 1u + 2u * ((tiles + kBarrierGroup - 1u) / kBarrierGroup);
}
// Workgroups of kPcgSolve the device holds at once, one per CU (0 if one cannot be resident); extra: the instance
// for slices with more than 8 upper or lower slots.
uint32_t pcgSolveCapacity(bool extra);
// Clears the counters (counterLines * 64 words at c.ctr) and, with clearAbort, the abort flag after them, and
// launches the solve (an ordinary dispatch of xcdGrid(nTiles) workgroups; the caller checks the grid against
// pcgSolveCapacity). A flag left set makes the launch's waits give up at once. Returns the launch's error code.
hipError_t launchPcgSolve(const View& v, const SolveCtl& c, uint32_t counterLines, bool extra, bool clearAbort,
                          hipStream_t st);
// Persistent solve by tile pairs (kPcgSolvePair): for launches whose tiles outnumber the CUs, one 512-thread
// workgroup per CU takes two consecutive tiles of a segment, with one window (the union of the tiles' columns, at
// most kPairWindows ranges of kPairRows rows in all; View::pairWindows); bit-identical to kPcgSolve.
// Pair workgroups of kPcgSolvePair the device holds at once (one per CU; 0 if one cannot be resident).
uint32_t pcgSolvePairCapacity(bool extra);
// As launchPcgSolve, by tile pairs: xcdGrid(nPairs) workgroups of 512 threads (with c.order: nPairs, order[i] a pair).
hipError_t launchPcgSolvePairs(const View& v, const SolveCtl& c, uint32_t counterLines, bool extra, bool clearAbort,
                               hipStream_t st);
// FP64 true residual of the finished solve and, per segment, a continuation of the solve from it when it is above
// tol r0 (at most maxRestarts per solve; see kTrueResidual). abortFlag (or null): when it is set, the persistent
// solve before gave up, and no segment continues.
void launchPcgRefine(const View& v, uint32_t maxRestarts, const uint32_t* abortFlag, hipStream_t st);
void launchContactForceSlope(const View& v, hipStream_t st);
void launchNewtonSlope(const View& v, hipStream_t st);
void launchEnergyTrial(const View& v, hipStream_t st);
void launchLineSearchDecide(const View& v, hipStream_t st);
void launchApply(const View& v, hipStream_t st);
void launchNewtonEnd(const View& v, hipStream_t st);
void launchFinish(const View& v, hipStream_t st);
// Learned warm start: SDF features at x; the energy check of the learned guess against the previous-velocity
// guess (two energy-only evaluations, then per segment keep the lower; cpu_solver.cpp:264-274 energy).
void launchWarmFeatures(const View& v, hipStream_t st);
void launchWarmCheck(const View& v, hipStream_t st);
void launchFill(uint32_t* p, uint32_t value, uint64_t count, hipStream_t st);

// Debug / standalone paths.
// y = A x (pinned rows and columns excluded); maxK = the most upper slots of a slice (a general matrix, lowK null,
// with maxK <= 16 takes the one-batch kernel; the value buffer must then carry kSellSlack zero floats past its end).
void launchSpmvPlain(const View& v, uint32_t maxK, const F3* in, F3* out, hipStream_t st);
void launchSellToSlots(const View& v, float* out /* 9 per slot */, hipStream_t st);
void launchEvalTerms(int kind, uint32_t n, const float* x, const float* params, const Fabric32* fabrics,
                     float limitScale, float* energy, float* grad, float* hess, hipStream_t st);
void launchFieldQuery(const float* values, Field32 field, uint32_t n, const float* p, float* phi, float* normal,
                      hipStream_t st);

// Structure builder (all on the device).
// Device arrays it allocates are handed to the caller, which frees them.
struct BuiltStructure {
  uint32_t nBlocks = 0;
  uint32_t maxK = 0;          // upper slots of the largest slice
  uint32_t maxLowK = 0;       // lower slots of the largest slice
  bool uniform = false;       // every slice has maxK upper and maxLowK lower slots (see kSellUniform)
  uint64_t sellFloats = 0;    // upper block values (the value buffer adds kSellSlack zero floats)
  uint64_t lowEntries = 0;    // lower slots times 64
  uint32_t* colIndex = nullptr;  // nBlocks
  uint32_t* sellCol = nullptr;   // sellFloats / 9 + kSellColSlack
  uint32_t* lowCol = nullptr;    // lowEntries + kSellColSlack
  uint32_t* lowBlk = nullptr;    // lowEntries + kSellColSlack
};
void launchBuildStructure(const BuildView& b, BuiltStructure& out, hipStream_t st);

}  // namespace dk
