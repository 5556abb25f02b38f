// layout.hpp — device data layout of the MI355X cloth step, shared by the host orchestration and the kernels.
//
// Rows (particles) are grouped in slices of 64 (one wavefront, one row per lane) and tiles of 256 (four slices,
// one 256-thread workgroup). Every segment (an independent scene of a batched build) starts on a tile boundary, so
// a tile never mixes segments; the gap rows are inert (flag kRowPad, treated as pinned). For a single-segment
// build the internal row of a particle is its global index in the reference's particle order
// (libs/sim/include/drape/sim/solver.hpp:84); only padding rows follow the last particle.
//
// The Hessian is block-CSR over 3x3 blocks with the reference's pattern and slot numbering
// (libs/sim/src/sim_model.cpp:41-74): rowStart/colIndex/diagSlot. It is symmetric, and the device stores each
// off-diagonal block once, in a sliced-ELL layout (SELL-64) that one lane per row streams coalesced:
//  - upper part: a row's diagonal block and the blocks of the columns above it (upper slot 0 is the diagonal).
//    Slice t holds K_t = the most such blocks of its 64 rows; value (slot k of lane l, component c) is at
//    sellBase[t] + (k * 64 + l) * 9 + c — a lane's block is 9 consecutive floats and the 64 lanes' blocks of one
//    slot are consecutive, so a wavefront reads 2,304 contiguous bytes per slot — and the column index at
//    sellBase[t] / 9 + k * 64 + l. Components are row-major (c = 3 * row + col).
//  - lower part: for each column below the diagonal, the column (lowCol) and the block index (lowBlk, in units
//    of 9 floats) of block (column, row) in that column's upper part, whose transpose is block (row, column);
//    slice t holds KL_t slots at lowBase[t] + k * 64 + l.
// Padding entries carry the dummy column `nRows` (a zero entry of every gathered vector), zero values and, in the
// lower part, the index of a zero block after the last slice; pinned columns also carry the dummy column, which
// excludes them from the product (cpu_solver.cpp:147-159).
#pragma once

#include <cstdint>

#ifndef __HIPCC__
#define __host__
#define __device__
#endif

namespace dk {

constexpr uint32_t kSlice = 64;   // rows per slice = lanes per wavefront
constexpr uint32_t kTile = 256;   // rows per tile = threads per workgroup
constexpr uint32_t kSlicesPerTile = kTile / kSlice;
constexpr uint32_t kMaxIncidence = 96;  // term incidences per particle the structure builder accepts
constexpr uint32_t kMaxRow = 64;        // blocks per row the structure builder accepts
// Zero floats after the last slice's block values (the first zero block is the lower part's padding block): a
// kernel may read a whole batch of up to 16 slots of any slice without testing each slot against K.
constexpr uint32_t kSellSlack = 16u * 9u * 64u;
// Entries after the last slice's upper columns, lower columns and lower block indices: a kernel may load the
// indices of 16 slots past any slice's K (and discards them).
constexpr uint32_t kSellColSlack = 16u * 64u;
// When no slice needs more than this many upper or lower slots, every slice gets the largest slice's counts (the
// extra slots of a smaller slice are padding: the dummy column and the zero block, as past a shorter row's last
// slot). A slice's first slot then sits at slice times that count, so the iteration kernel can address its
// matrix without reading the slice headers first.
constexpr uint32_t kUniformSlots = 8u;
// Largest window of rows whose (zr, t, p) a tile stages in LDS for the halo form of the PCG iteration (36 B a row:
// 54 KB; 6 rows per thread of a 256-thread workgroup).
constexpr uint32_t kHaloRows = 1536u;
// Floats per row of the block-Jacobi preconditioner: the upper triangle of the symmetric inverse diagonal block.
constexpr uint32_t kPrecondFloats = 6u;

// Row flags.
constexpr uint32_t kRowPinned = 1u;
constexpr uint32_t kRowPad = 2u;

// Contact flags (GpuContact nearBody / active, gpu_layout.h:74-78).
constexpr uint32_t kContactNear = 1u;
constexpr uint32_t kContactActive = 2u;

// Incidence entry: kind (2 bits) | local vertex a (2 bits) | term index (28 bits); ascending order = (kind, term, a),
// the reference's gather order (gpu_solver_core.cpp:184-224).
constexpr uint32_t kKindStretch = 0, kKindHinge = 1, kKindSeam = 2;
__host__ __device__ inline uint32_t incEncode(uint32_t kind, uint32_t term, uint32_t a) {
  return (kind << 30) | (term << 2) | a;
}
__host__ __device__ inline uint32_t incKind(uint32_t e) { return e >> 30; }
__host__ __device__ inline uint32_t incTerm(uint32_t e) { return (e >> 2) & 0x0FFFFFFFu; }
__host__ __device__ inline uint32_t incLocal(uint32_t e) { return e & 3u; }

struct F3 {
  float x, y, z;
};

// FabricPhysical fields the kernels read, as float (GpuFabric, gpu_layout.h:26-31).
struct Fabric32 {
  float stretchWarp, stretchWeft, stretchBias;
  float strainLimitWarp, strainLimitWeft;
  float weight, thickness, friction, damping;
  float pad[3];
};

struct Stretch32 {
  uint32_t v[3];
  uint32_t fabric;
  float dmInv[4];  // (0,0), (1,0), (0,1), (1,1)
  float area;
  float pad[3];
};

struct Hinge32 {
  uint32_t v[4];  // opposite A, opposite B, edge start, edge end
  float K, restAngle;
  float pad[2];
};

struct Seam32 {
  uint32_t a, b, garment;
  float startGap;
};

struct Garment32 {
  uint32_t phase;  // 0 Assembling, 1 Settling
  float assemblyTime;
  uint32_t pad[2];
};

// Body field of one segment: samples at origin + (i,j,k) cell, x fastest (CollisionField, collision_field.hpp:11-20).
struct Field32 {
  float ox, oy, oz, cell;
  uint32_t dx, dy, dz;
  float band;
  uint64_t offset;  // first value in the concatenated field buffer
  uint32_t present;
  uint32_t pad;
};

// Per-segment constants (StepParams, gpu_layout.h:57-62) and layout.
struct SegParams {
  float gx, gy, gz, h;
  float invH2, limitScale, seamStiffness, closingSpeed;
  float frictionSlip, collisionOffset, dampingScale, assemblyExtraDamping;
  float contactReach, contactSnap, pad0, pad1;
  double cgTolerance;
  uint32_t maxCg;
  uint32_t tileBegin, tileEnd;  // tiles [tileBegin, tileEnd)
  uint32_t rowBegin, rowEnd;    // internal rows [rowBegin, rowEnd) (rowEnd padded to a tile)
  uint32_t pad2;
  Field32 field;
  double fieldOrigin[3], fieldCell, fieldBand;  // the body grid's FP64 geometry (the FP64 sample of the stats)
};

// Per-segment solver state on the device. Written by the first workgroup of the segment (or by the per-segment
// control kernels); every other workgroup recomputes the same values from the same partial sums in the same order.
struct SegState {
  double rzPar[2];                      // PCG r.z, by iteration parity (iteration k reads [(k+1)&1], writes [k&1])
  double pAp, alpha, beta, r0, rr;      // PCG (CgScalars, gpu_layout.h:54-57, in FP64)
  double e0, eTrial, slope, lsAlpha;    // Newton energy, trial energy, grad . step, line-search step
  float maxMove;
  uint32_t pcgConverged;  // 1 once |r| <= tol |r0| or pAp <= 0 (read by the host; kPcgStep reads per-tile flags)
  uint32_t pcgIter;
  uint32_t newtonDone;    // 1 once slope >= 0 or max move < 1e-9 (no more Newton iterations this substep)
  uint32_t lsDone;        // 1 once the line search decided this Newton iteration, 2 once the step is applied
  uint32_t lsTrial;
  uint32_t active;        // 1 while this segment runs Newton iterations this substep
  uint32_t warmRejected;  // 1 when the learned warm start was rejected this substep
  // One-kernel PCG: iteration index by kernel parity. A kernel of parity q reads kIter[q] (written by the previous
  // kernel) and writes slot 1-q, so no workgroup ever reads a value another workgroup of the same launch writes
  // (the per-tile solve-ended flags, View::tileDone, follow the same rule).
  uint32_t kIter[2];
  uint64_t cntNewton, cntPcg, cntLs, cntSubsteps, cntWarmRejected, cntWarmAccepted;
  // Linear-solve report: why each PCG solve ended (kPcgExit*), counted per reason, and the FP64 true relative
  // residual |S(-grad - A step)| / |S(-grad)| at each solve's end (S the contact filter).
  uint32_t pcgExit;
  // FP64 true-residual refinement: restartActive = 1 while a solve continues from its true residual (it keeps the
  // solve's r0); kBase = the updates done before the continuation (the iteration cap counts the whole solve).
  uint32_t restartActive, kBase, restartsThisSolve, pad3;
  uint64_t cntSolves, cntExitTolerance, cntExitCap, cntExitCurvature, cntExitZero, cntRestarts;
  double trueRelMax, trueRelSum, trueRelLast, trueRelMaxConverged;  // the last: over solves ended on the tolerance test
  // Over solves ended on the tolerance test: the largest true residual over max(tol |r0|, the part of it that
  // rounding step to FP32 leaves, estimated as in kRefineDecide).
  double trueExcessMax;
};

// Why a PCG solve ended (cpu_solver.cpp:188-204 stopping rules).
constexpr uint32_t kPcgExitTolerance = 1;  // |r| <= tol |r0| after an update
constexpr uint32_t kPcgExitCap = 2;        // maxCg updates done
constexpr uint32_t kPcgExitCurvature = 3;  // p.Ap <= 0 before an update
constexpr uint32_t kPcgExitZero = 4;       // r0 = 0: nothing to solve

}  // namespace dk
