// solver.cpp — HipSolver: build, substep orchestration and state access on one MI355X.
//
// Substep (cpu_solver.cpp:230-314 semantics, every segment independent):
//   predict + contact setup
//   Newton iteration (up to `iterations`, stopping a segment once its slope is not negative or its largest move
//   is below 1e-9 m):
//     activate contacts; assemble energy / gradient / block-CSR Hessian; e0
//     block-Jacobi PCG (one kernel per iteration, replayed from hipGraph chunks, or the two-kernel form; the host
//     polls the per-segment converged flags after each batch of chunks)
//     contact forces from the unfiltered residual; slope
//     Armijo line search (one trial kernel + one decision kernel per trial)
//     apply; largest move
//   finish (push-out, pins, damped velocity)
// All decisions are taken on the device from FP64 sums; the host only reads flags to know when to stop
// enqueueing work.
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace dh {

namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("hip: ") + what + ": " + hipGetErrorString(e));
}

dk::F3 f3(const Vec3f& v) { return {v.x, v.y, v.z}; }
dk::F3 f3(const Vec3& v) { return {static_cast<float>(v.x), static_cast<float>(v.y), static_cast<float>(v.z)}; }

dk::Fabric32 fabric32(const Fabric& f) {
  dk::Fabric32 g{};
  g.stretchWarp = static_cast<float>(f.stretchWarp);
  g.stretchWeft = static_cast<float>(f.stretchWeft);
  g.stretchBias = static_cast<float>(f.stretchBias);
  g.strainLimitWarp = static_cast<float>(f.strainLimitWarp);
  g.strainLimitWeft = static_cast<float>(f.strainLimitWeft);
  g.weight = static_cast<float>(f.weight);
  g.thickness = static_cast<float>(f.thickness);
  g.friction = static_cast<float>(f.friction);
  g.damping = static_cast<float>(f.damping);
  return g;
}

constexpr double kContactReach = 0.01;   // sim_model.hpp:14
constexpr double kContactSnap = 0.0005;  // sim_model.hpp:15

}  // namespace

// Work on the null stream is not ordered with a solver's stream (created non-blocking): the fill must be complete
// before this returns, or a kernel on that stream may write the buffer first.
void DeviceBuffer::allocate(size_t bytes) {
  release();
  const size_t b = std::max<size_t>(bytes, 16);
  check(hipMalloc(&ptr_, b), "hipMalloc");
  check(hipMemsetAsync(ptr_, 0, b, nullptr), "hipMemset");
  check(hipStreamSynchronize(nullptr), "hipMemset");
  bytes_ = b;
}

void DeviceBuffer::release() {
  if (ptr_) (void)hipFree(ptr_);
  ptr_ = nullptr;
  bytes_ = 0;
}

HipSolver::HipSolver() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw std::runtime_error("hip: no GPU device available");
  hipDeviceProp_t prop;
  check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    throw std::runtime_error(std::string("hip: device is ") + prop.gcnArchName + ", this build targets gfx950");
  check(hipSetDevice(0), "hipSetDevice");
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
}

HipSolver::~HipSolver() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  releaseGraph();
  for (auto& pr : eventPool_) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  for (auto& pr : eventPending_) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  if (hostStates_) (void)hipHostFree(hostStates_);
  if (solveAbortHost_) (void)hipHostFree(solveAbortHost_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void HipSolver::requireBuilt() const {
  if (!built_) throw std::logic_error("hip: solver not built");
}

// Copies on the solver's stream, complete when they return (the stream is non-blocking, so copies on the null
// stream would not be ordered with its kernels).
void HipSolver::toDevice(void* dst, const void* src, size_t bytes) const {
  check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), "copy");
  check(hipStreamSynchronize(stream_), "copy");
}

void HipSolver::fromDevice(void* dst, const void* src, size_t bytes) const {
  check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_), "copy");
  check(hipStreamSynchronize(stream_), "copy");
}

template <class T>
void HipSolver::upload(DeviceBuffer& b, const std::vector<T>& data, size_t count) {
  const size_t n = std::max(count, data.size());
  b.allocate(sizeof(T) * std::max<size_t>(n, 1));
  if (!data.empty()) toDevice(b.as<void>(), data.data(), sizeof(T) * data.size());
}

void HipSolver::build(const drape_scene_desc* scenes, uint32_t nScenes) {
  Model m = Model::build(scenes, nScenes);
  check(hipStreamSynchronize(stream_), "sync");
  releaseGraph();
  model_ = std::move(m);
  built_ = false;
  normalMesh_.reset();
  lastH_ = -1.0;
  const Model& M = model_;
  nRows_ = M.rowCount();
  nTiles_ = nRows_ / dk::kTile;
  nSlices_ = nRows_ / dk::kSlice;
  const size_t n = M.x.size();
  rowOf_.resize(n);
  for (uint32_t i = 0; i < n; ++i) rowOf_[i] = M.internalRow(i);

  // Per-row constants (padding rows are pinned and inert).
  std::vector<uint32_t> flags(nRows_, dk::kRowPinned | dk::kRowPad), fab(nRows_, 0), gar(nRows_, 0);
  std::vector<float> mass(nRows_, 1.0f);
  std::vector<dk::F3> fext(nRows_ + 1, dk::F3{0, 0, 0}), pins(nRows_ + 1, dk::F3{0, 0, 0});
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t r = rowOf_[i];
    flags[r] = M.pinned[i] ? dk::kRowPinned : 0u;
    fab[r] = M.pFabric[i];
    gar[r] = M.pGarment[i];
    mass[r] = static_cast<float>(M.mass[i]);
    fext[r] = f3(M.fExt[i]);
  }
  for (const auto& [i, pos] : M.pins) pins[rowOf_[i]] = f3(pos);
  upload(rowFlags_, flags);
  upload(mass_, mass);
  std::vector<double> massF64(nRows_, 0.0);
  for (uint32_t i = 0; i < n; ++i) massF64[rowOf_[i]] = M.mass[i];
  upload(massF64_, massF64);
  upload(rowFabric_, fab);
  upload(rowGarment_, gar);
  upload(fExt_, fext);
  upload(pinPos_, pins);
  const size_t vecBytes = sizeof(dk::F3) * (nRows_ + 1);
  for (DeviceBuffer* b : {&x_, &v_, &xStart_, &yRel_, &u_, &grad_, &step_, &r_, &z_, &pA_, &pB_, &ap_, &warm_,
                          &scratchF3_, &scratchF3b_, &dvLast_})
    b->allocate(vecBytes);
  phiN_.allocate(sizeof(float4) * (nRows_ + 1));
  {
    // Mesh edge graph over internal rows (triangle edges, sorted unique neighbours) for the learned warm start.
    std::vector<std::vector<uint32_t>> adj(nRows_);
    for (const StretchTerm& t : M.stretch)
      for (int a = 0; a < 3; ++a) {
        const uint32_t i = rowOf_[t.v[a]], j = rowOf_[t.v[(a + 1) % 3]];
        adj[i].push_back(j);
        adj[j].push_back(i);
      }
    std::vector<uint32_t> rs(nRows_ + 1, 0), nb, mask(nRows_, 0);
    std::vector<double> rest(2 * size_t(nRows_), 0.0);
    for (uint32_t r = 0; r < nRows_; ++r) {
      std::sort(adj[r].begin(), adj[r].end());
      adj[r].erase(std::unique(adj[r].begin(), adj[r].end()), adj[r].end());
      nb.insert(nb.end(), adj[r].begin(), adj[r].end());
      rs[r + 1] = static_cast<uint32_t>(nb.size());
    }
    for (uint32_t i = 0; i < n; ++i) {
      mask[rowOf_[i]] = 1u;
      rest[2 * size_t(rowOf_[i])] = M.rest[2 * size_t(i)];
      rest[2 * size_t(rowOf_[i]) + 1] = M.rest[2 * size_t(i) + 1];
    }
    meshGraph_ = std::make_unique<ws::Graph>(nRows_, rs.data(), nb.data(), rest.data(), mask.data());
  }
  // Row arrays a gathered column may index hold the zero dummy entry nRows as well; the PCG vectors carry 16 more
  // zero bytes, the most the persistent solve's 16-byte window chunks read past the dummy entry.
  for (DeviceBuffer& b : cgBuf_) b.allocate(vecBytes + 16u);
  precond_.allocate(sizeof(float) * dk::kPrecondFloats * (nRows_ + 1));
  cnormal_.allocate(sizeof(float4) * (nRows_ + 1));
  cflags_.allocate(sizeof(uint32_t) * (nRows_ + 1));
  cforce_.allocate(sizeof(float) * nRows_);
  clambda_.allocate(sizeof(float) * nRows_);

  // Segments.
  std::vector<uint32_t> tileSeg(nTiles_);
  std::vector<float> fieldValues;
  std::vector<dk::SegParams> params(M.segments.size());
  for (uint32_t s = 0; s < M.segments.size(); ++s) {
    const Segment& sg = M.segments[s];
    for (uint32_t t = sg.rowBegin / dk::kTile; t < (sg.rowBegin + sg.rowCount) / dk::kTile; ++t) tileSeg[t] = s;
    dk::SegParams& P = params[s];
    P = dk::SegParams{};
    P.gx = static_cast<float>(sg.settings.gravity[0]);
    P.gy = static_cast<float>(sg.settings.gravity[1]);
    P.gz = static_cast<float>(sg.settings.gravity[2]);
    P.limitScale = static_cast<float>(sg.settings.limitScale);
    P.seamStiffness = static_cast<float>(sg.settings.seamStiffness);
    P.closingSpeed = static_cast<float>(sg.settings.closingSpeed);
    P.frictionSlip = static_cast<float>(sg.settings.frictionSlip);
    P.collisionOffset = static_cast<float>(sg.settings.collisionOffset);
    P.dampingScale = static_cast<float>(sg.settings.dampingScale);
    P.assemblyExtraDamping = static_cast<float>(sg.settings.extraDamping);
    P.contactReach = static_cast<float>(kContactReach);
    P.contactSnap = static_cast<float>(kContactSnap);
    P.cgTolerance = sg.settings.cgTolerance;
    P.maxCg = static_cast<uint32_t>(std::max(0, sg.settings.maxCg));
    P.tileBegin = sg.rowBegin / dk::kTile;
    P.tileEnd = (sg.rowBegin + sg.rowCount) / dk::kTile;
    P.rowBegin = sg.rowBegin;
    P.rowEnd = sg.rowBegin + sg.rowCount;
    if (sg.body.present()) {
      P.field.ox = static_cast<float>(sg.body.origin[0]);
      P.field.oy = static_cast<float>(sg.body.origin[1]);
      P.field.oz = static_cast<float>(sg.body.origin[2]);
      P.field.cell = static_cast<float>(sg.body.cell);
      P.field.dx = static_cast<uint32_t>(sg.body.dims[0]);
      P.field.dy = static_cast<uint32_t>(sg.body.dims[1]);
      P.field.dz = static_cast<uint32_t>(sg.body.dims[2]);
      P.field.offset = fieldValues.size();
      P.field.present = 1u;
      fieldValues.insert(fieldValues.end(), sg.body.values.begin(), sg.body.values.end());
    }
    P.field.band = static_cast<float>(sg.body.band);
    for (int a = 0; a < 3; ++a) P.fieldOrigin[a] = sg.body.origin[a];
    P.fieldCell = sg.body.cell;
    P.fieldBand = sg.body.band;
  }
  upload(tileSeg_, tileSeg);
  std::vector<uint32_t> tileRange(2 * size_t(nTiles_), 0u);
  for (uint32_t t = 0; t < nTiles_; ++t) {
    tileRange[2 * t] = params[tileSeg[t]].tileBegin;
    tileRange[2 * t + 1] = params[tileSeg[t]].tileEnd;
  }
  upload(tileRange_, tileRange);
  upload(segParams_, params);
  segState_.allocate(sizeof(dk::SegState) * std::max<size_t>(M.segments.size(), 1));
  if (hostStates_) (void)hipHostFree(hostStates_);
  hostStates_ = nullptr;
  check(hipHostMalloc(reinterpret_cast<void**>(&hostStates_), sizeof(dk::SegState) * std::max<size_t>(M.segments.size(), 1),
                      hipHostMallocDefault),
        "hipHostMalloc");
  std::memset(hostStates_, 0, sizeof(dk::SegState) * std::max<size_t>(M.segments.size(), 1));
  upload(fieldValues_, fieldValues);
  partA_.allocate(sizeof(double) * nTiles_);
  partB_.allocate(sizeof(double) * nTiles_);
  partC_.allocate(sizeof(double) * nTiles_);
  partE_.allocate(sizeof(double) * nSlices_);
  partMax_.allocate(sizeof(float) * nTiles_);
  statsTile_.allocate(sizeof(double) * dk::kStatsPerSegment * nTiles_);
  statsSeg_.allocate(sizeof(double) * dk::kStatsPerSegment * std::max<size_t>(M.segments.size(), 1));
  tileDone_.allocate(sizeof(uint32_t) * 2 * nTiles_);
  cgPart0_.allocate(sizeof(double) * dk::kPartStride * nTiles_);
  cgPart1_.allocate(sizeof(double) * dk::kPartStride * nTiles_);

  // Terms in internal row indices.
  std::vector<dk::Stretch32> st(M.stretch.size());
  for (size_t k = 0; k < M.stretch.size(); ++k) {
    const StretchTerm& t = M.stretch[k];
    dk::Stretch32 g{};
    for (int a = 0; a < 3; ++a) g.v[a] = rowOf_[t.v[a]];
    for (int a = 0; a < 4; ++a) g.dmInv[a] = static_cast<float>(t.dmInv[a]);
    g.area = static_cast<float>(t.area);
    g.fabric = t.fabric;
    st[k] = g;
  }
  upload(stretch_, st);
  std::vector<dk::Seam32> sm(M.seams.size());
  for (size_t k = 0; k < M.seams.size(); ++k)
    sm[k] = dk::Seam32{rowOf_[M.seams[k].a], rowOf_[M.seams[k].b], M.seams[k].garment, M.seams[k].startGap};
  upload(seams_, sm);
  hinges_.allocate(sizeof(dk::Hinge32) * std::max<size_t>(M.hinges.size(), 1));
  uploadHinges();
  fabrics_.allocate(sizeof(dk::Fabric32) * std::max<size_t>(M.fabrics.size(), 1));
  garments_.allocate(sizeof(dk::Garment32) * std::max<size_t>(M.garments.size(), 1));
  uploadFabricsAndMass();
  uploadGarments();

  // Structure on the device.
  const uint32_t S = static_cast<uint32_t>(M.stretch.size()), H = static_cast<uint32_t>(M.hinges.size()),
                 K = static_cast<uint32_t>(M.seams.size());
  incCount_.allocate(sizeof(uint32_t) * nRows_);
  incStart_.allocate(sizeof(uint32_t) * (nRows_ + 1));
  inc_.allocate(sizeof(uint32_t) * std::max<size_t>(3ull * S + 4ull * H + 2ull * K, 1));
  rowLen_.allocate(sizeof(uint32_t) * nRows_);
  rowStart_.allocate(sizeof(uint32_t) * (nRows_ + 1));
  diagSlot_.allocate(sizeof(uint32_t) * nRows_);
  stretchSlots_.allocate(sizeof(uint32_t) * 9 * std::max<uint32_t>(S, 1));
  hingeSlots_.allocate(sizeof(uint32_t) * 16 * std::max<uint32_t>(H, 1));
  seamSlots_.allocate(sizeof(uint32_t) * 4 * std::max<uint32_t>(K, 1));
  sellK_.allocate(sizeof(uint32_t) * nSlices_);
  sellBase_.allocate(sizeof(uint32_t) * (nSlices_ + 1));
  lowK_.allocate(sizeof(uint32_t) * nSlices_);
  lowBase_.allocate(sizeof(uint32_t) * (nSlices_ + 1));
  buildError_.allocate(sizeof(uint32_t));
  dk::BuildView b{};
  b.nRows = nRows_;
  b.nStretch = S;
  b.nHinge = H;
  b.nSeam = K;
  b.nSlices = nSlices_;
  b.stretch = stretch_.as<dk::Stretch32>();
  b.hinges = hinges_.as<dk::Hinge32>();
  b.seams = seams_.as<dk::Seam32>();
  b.rowFlags = rowFlags_.as<uint32_t>();
  b.incCount = incCount_.as<uint32_t>();
  b.incStart = incStart_.as<uint32_t>();
  b.inc = inc_.as<uint32_t>();
  b.rowLen = rowLen_.as<uint32_t>();
  b.rowStart = rowStart_.as<uint32_t>();
  b.diagSlot = diagSlot_.as<uint32_t>();
  b.stretchSlots = stretchSlots_.as<uint32_t>();
  b.hingeSlots = hingeSlots_.as<uint32_t>();
  b.seamSlots = seamSlots_.as<uint32_t>();
  b.sellK = sellK_.as<uint32_t>();
  b.sellSize = sellBase_.as<uint32_t>();
  b.lowK = lowK_.as<uint32_t>();
  b.lowSize = lowBase_.as<uint32_t>();
  b.error = buildError_.as<uint32_t>();
  dk::BuiltStructure bs;
  dk::launchBuildStructure(b, bs, stream_);
  colIndex_.adopt(bs.colIndex, sizeof(uint32_t) * std::max<uint32_t>(bs.nBlocks, 1));
  sellCol_.adopt(bs.sellCol, sizeof(uint32_t) * (bs.sellFloats / 9 + dk::kSellColSlack));
  lowCol_.adopt(bs.lowCol, sizeof(uint32_t) * (bs.lowEntries + dk::kSellColSlack));
  lowBlk_.adopt(bs.lowBlk, sizeof(uint32_t) * (bs.lowEntries + dk::kSellColSlack));
  nBlocks_ = bs.nBlocks;
  sellFloats_ = bs.sellFloats;
  lowEntries_ = bs.lowEntries;
  maxK_ = bs.maxK;
  maxLowK_ = bs.maxLowK;
  uniformLayout_ = bs.uniform;
  if (maxK_ * 9ull * dk::kSlice * sizeof(float) > 64ull * 1024)
    throw std::invalid_argument("hip: a 64-row slice has more than 28 blocks on or above the diagonal in one row");
  // Per tile, the window of rows [first, first + count) that holds every column of its rows and the rows
  // themselves; the halo launch shapes run only when every window fits kHaloRows (a banded matrix).
  bool haloFits = true, windowsFit = true;
  {
    std::vector<uint32_t> rs(nRows_ + 1), ci(std::max<uint32_t>(nBlocks_, 1));
    fromDevice(rs.data(), rowStart_.as<void>(), sizeof(uint32_t) * (nRows_ + 1));
    fromDevice(ci.data(), colIndex_.as<void>(), sizeof(uint32_t) * nBlocks_);
    segBlocks_.assign(M.segments.size(), 0.0);
    segParticles_.assign(M.segments.size(), 0.0);
    for (size_t s = 0; s < M.segments.size(); ++s) {
      const Segment& sg = M.segments[s];
      for (uint32_t r = sg.rowBegin; r < sg.rowBegin + sg.rowCount; ++r) {
        if (flags[r] & dk::kRowPad) continue;
        segBlocks_[s] += double(rs[r + 1] - rs[r]);
        segParticles_[s] += 1.0;
      }
    }
    std::vector<uint32_t> halo(2 * size_t(nTiles_), 0u);
    for (uint32_t t = 0; t < nTiles_; ++t) {
      uint32_t lo = t * dk::kTile, hi = (t + 1) * dk::kTile - 1;
      for (uint32_t r = t * dk::kTile; r < (t + 1) * dk::kTile; ++r)
        for (uint32_t k = rs[r]; k < rs[r + 1]; ++k) {
          lo = std::min(lo, ci[k]);
          hi = std::max(hi, ci[k]);
        }
      halo[2 * t] = lo;
      halo[2 * t + 1] = hi - lo + 1;
      haloFits = haloFits && hi - lo + 1 <= dk::kHaloRows;
    }
    upload(tileHalo_, halo);
    // The persistent solve's windows: per tile at most dk::kSolveWindows row ranges holding every column of its rows
    // (and the rows themselves), kHaloRows rows in all. The tile's distinct columns, split where consecutive ones lie
    // more than 64 rows apart, the closest ranges merged until at most kSolveWindows remain.
    std::vector<uint32_t> win(2 * size_t(dk::kSolveWindows) * nTiles_, 0u);
    windowsFit = true;
    std::vector<uint32_t> cols;
    for (uint32_t t = 0; t < nTiles_ && windowsFit; ++t) {
      cols.clear();
      for (uint32_t r = t * dk::kTile; r < (t + 1) * dk::kTile; ++r) {
        cols.push_back(r);
        for (uint32_t k = rs[r]; k < rs[r + 1]; ++k) cols.push_back(ci[k]);
      }
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      std::vector<std::pair<uint32_t, uint32_t>> runs{{cols[0], cols[0]}};
      for (size_t k = 1; k < cols.size(); ++k) {
        if (cols[k] - runs.back().second <= 64u)
          runs.back().second = cols[k];
        else
          runs.push_back({cols[k], cols[k]});
      }
      while (runs.size() > dk::kSolveWindows) {
        size_t best = 0;
        for (size_t k = 1; k + 1 < runs.size(); ++k)
          if (runs[k + 1].first - runs[k].second < runs[best + 1].first - runs[best].second) best = k;
        runs[best].second = runs[best + 1].second;
        runs.erase(runs.begin() + long(best) + 1);
      }
      uint32_t total = 0;
      for (size_t k = 0; k < runs.size(); ++k) {
        win[2 * (dk::kSolveWindows * size_t(t) + k)] = runs[k].first;
        win[2 * (dk::kSolveWindows * size_t(t) + k) + 1] = runs[k].second - runs[k].first + 1;
        total += runs[k].second - runs[k].first + 1;
      }
      windowsFit = total <= dk::kHaloRows;
    }
    upload(tileWindows_, win);
    // Per tile, the slots past the eighth of its rows (upper: the diagonal and the columns after it; lower: those
    // before), which the persistent solve keeps in an LDS list of kSolveExtraCap entries.
    tileExtrasMax_ = 0;
    for (uint32_t t = 0; t < nTiles_; ++t) {
      uint32_t extras = 0;
      for (uint32_t r = t * dk::kTile; r < (t + 1) * dk::kTile; ++r) {
        const uint32_t lower = uint32_t(std::lower_bound(ci.begin() + rs[r], ci.begin() + rs[r + 1], r) -
                                        (ci.begin() + rs[r]));
        const uint32_t upper = rs[r + 1] - rs[r] - lower;
        extras += (lower > 8u ? lower - 8u : 0u) + (upper > 8u ? upper - 8u : 0u);
      }
      tileExtrasMax_ = std::max(tileExtrasMax_, extras);
    }
    // The persistent solve by tile pairs: per segment, consecutive tiles two to a workgroup (the last alone when the
    // count is odd); each pair's window the union of its tiles' columns, in at most kPairWindows ranges of kPairRows
    // rows in all, and its LDS list of slots past the eighth at most 2 kSolveExtraCap long.
    std::vector<uint32_t> pairTiles, pairWin;
    segPairs_.assign(M.segments.size(), {0u, 0u});
    pairsFit_ = true;
    for (size_t sg = 0; sg < M.segments.size(); ++sg) {
      segPairs_[sg].first = uint32_t(pairTiles.size());
      for (uint32_t t = params[sg].tileBegin; t < params[sg].tileEnd; t += 2u) {
        const uint32_t n = std::min(2u, params[sg].tileEnd - t);
        pairTiles.push_back(t | (n - 1u) << 31);
        cols.clear();
        uint32_t extras = 0;
        for (uint32_t r = t * dk::kTile; r < (t + n) * dk::kTile; ++r) {
          cols.push_back(r);
          for (uint32_t k = rs[r]; k < rs[r + 1]; ++k) cols.push_back(ci[k]);
          const uint32_t lower = uint32_t(std::lower_bound(ci.begin() + rs[r], ci.begin() + rs[r + 1], r) -
                                          (ci.begin() + rs[r]));
          const uint32_t upper = rs[r + 1] - rs[r] - lower;
          extras += (lower > 8u ? lower - 8u : 0u) + (upper > 8u ? upper - 8u : 0u);
        }
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        std::vector<std::pair<uint32_t, uint32_t>> runs{{cols[0], cols[0]}};
        for (size_t k = 1; k < cols.size(); ++k) {
          if (cols[k] - runs.back().second <= 64u)
            runs.back().second = cols[k];
          else
            runs.push_back({cols[k], cols[k]});
        }
        while (runs.size() > dk::kPairWindows) {
          size_t best = 0;
          for (size_t k = 1; k + 1 < runs.size(); ++k)
            if (runs[k + 1].first - runs[k].second < runs[best + 1].first - runs[best].second) best = k;
          runs[best].second = runs[best + 1].second;
          runs.erase(runs.begin() + long(best) + 1);
        }
        uint32_t total = 0;
        for (uint32_t k = 0; k < dk::kPairWindows; ++k) {
          pairWin.push_back(k < runs.size() ? runs[k].first : 0u);
          pairWin.push_back(k < runs.size() ? runs[k].second - runs[k].first + 1 : 0u);
          if (k < runs.size()) total += runs[k].second - runs[k].first + 1;
        }
        pairsFit_ = pairsFit_ && total <= dk::kPairRows && extras <= 2u * dk::kSolveExtraCap;
      }
      segPairs_[sg].second = uint32_t(pairTiles.size());
    }
    nPairs_ = uint32_t(pairTiles.size());
    upload(pairTiles_, pairTiles);
    upload(pairWindows_, pairWin);
  }
  tunedShape_ = -1;
  tuneSolves_ = 0;
  std::fill(tuneMs_, tuneMs_ + kShapeCount, 0.0);
  std::fill(tuneLaunches_, tuneLaunches_ + kShapeCount, 0u);
  std::fill(tuneWarm_, tuneWarm_ + kShapeCount, false);
  sellVal_.allocate(sizeof(float) * (sellFloats_ + dk::kSellSlack));  // zero slack: see kSellSlack

  dk::View& v = view_;
  v = dk::View{};
  v.nRows = nRows_;
  v.nTiles = nTiles_;
  v.nSlices = nSlices_;
  v.nSegments = static_cast<uint32_t>(M.segments.size());
  v.nStretch = S;
  v.nHinge = H;
  v.nSeam = K;
  v.zeroBlock = static_cast<uint32_t>(sellFloats_ / 9);
  v.uniformK = uniformLayout_ ? maxK_ : 0u;
  v.uniformKL = uniformLayout_ ? maxLowK_ : 0u;
  v.rowFlags = rowFlags_.as<uint32_t>();
  v.mass = mass_.as<float>();
  v.rowFabric = rowFabric_.as<uint32_t>();
  v.rowGarment = rowGarment_.as<uint32_t>();
  v.fExt = fExt_.as<dk::F3>();
  v.pinPos = pinPos_.as<dk::F3>();
  v.x = x_.as<dk::F3>();
  v.v = v_.as<dk::F3>();
  v.xStart = xStart_.as<dk::F3>();
  v.yRel = yRel_.as<dk::F3>();
  v.u = u_.as<dk::F3>();
  v.grad = grad_.as<dk::F3>();
  v.step = step_.as<dk::F3>();
  v.r = r_.as<dk::F3>();
  v.z = z_.as<dk::F3>();
  v.pA = pA_.as<dk::F3>();
  v.pB = pB_.as<dk::F3>();
  v.ap = ap_.as<dk::F3>();
  v.warm = warm_.as<dk::F3>();
  v.dvLast = dvLast_.as<dk::F3>();
  v.phiN = phiN_.as<float4>();
  for (int b = 0; b < 2; ++b) {
    v.cgR[b] = cgBuf_[0 + b].as<dk::F3>();
    v.cgAp[b] = cgBuf_[2 + b].as<dk::F3>();
    v.cgP[b] = cgBuf_[4 + b].as<dk::F3>();
    v.cgZr[b] = cgBuf_[6 + b].as<dk::F3>();
    v.cgT[b] = cgBuf_[8 + b].as<dk::F3>();
  }
  v.tileDone[0] = tileDone_.as<uint32_t>();
  v.tileDone[1] = tileDone_.as<uint32_t>() + nTiles_;
  v.cgPart[0] = cgPart0_.as<double>();
  v.cgPart[1] = cgPart1_.as<double>();
  v.precond = precond_.as<float>();
  v.cnormal = cnormal_.as<float4>();
  v.cflags = cflags_.as<uint32_t>();
  v.cforce = cforce_.as<float>();
  v.clambda = clambda_.as<float>();
  v.tileSeg = tileSeg_.as<uint32_t>();
  v.tileRange = tileRange_.as<uint32_t>();
  v.tileHalo = haloFits ? tileHalo_.as<uint32_t>() : nullptr;
  v.tileWindows = windowsFit ? tileWindows_.as<uint32_t>() : nullptr;
  v.nPairs = nPairs_;
  v.pairTiles = pairTiles_.as<uint32_t>();
  v.pairWindows = pairWindows_.as<uint32_t>();
  v.seg = segParams_.as<dk::SegParams>();
  v.state = segState_.as<dk::SegState>();
  v.partA = partA_.as<double>();
  v.partB = partB_.as<double>();
  v.partC = partC_.as<double>();
  v.partE = partE_.as<double>();
  v.partMax = partMax_.as<float>();
  v.stretch = stretch_.as<dk::Stretch32>();
  v.hinges = hinges_.as<dk::Hinge32>();
  v.seams = seams_.as<dk::Seam32>();
  v.fabrics = fabrics_.as<dk::Fabric32>();
  v.garments = garments_.as<dk::Garment32>();
  v.incStart = incStart_.as<uint32_t>();
  v.inc = inc_.as<uint32_t>();
  v.rowStart = rowStart_.as<uint32_t>();
  v.colIndex = colIndex_.as<uint32_t>();
  v.diagSlot = diagSlot_.as<uint32_t>();
  v.stretchSlots = stretchSlots_.as<uint32_t>();
  v.hingeSlots = hingeSlots_.as<uint32_t>();
  v.seamSlots = seamSlots_.as<uint32_t>();
  v.sellK = sellK_.as<uint32_t>();
  v.sellBase = sellBase_.as<uint32_t>();
  v.sellCol = sellCol_.as<uint32_t>();
  v.sellVal = sellVal_.as<float>();
  v.lowK = lowK_.as<uint32_t>();
  v.lowBase = lowBase_.as<uint32_t>();
  v.lowCol = lowCol_.as<uint32_t>();
  v.lowBlk = lowBlk_.as<uint32_t>();
  v.fieldValues = fieldValues_.as<float>();
  // Persistent solve: every tile's columns in at most four row ranges (kHaloRows rows in all), at most 16 + 16 slots
  // per slice, and each segment's tiles resident together; counters per segment for its barrier.
  {
    std::vector<uint32_t> segCtr(std::max<size_t>(M.segments.size(), 1), 0u);
    uint32_t lines = 0;
    for (size_t sg = 0; sg < M.segments.size(); ++sg) {
      segCtr[sg] = lines;
      lines += dk::solveCounterLines(params[sg].tileEnd - params[sg].tileBegin);
    }
    solveCtrLines_ = lines;
    solveCtr_.allocate(sizeof(uint32_t) * (size_t(lines) * 64u + 1u));
    upload(solveSegCtr_, segCtr);
    if (!solveAbortHost_)
      check(hipHostMalloc(reinterpret_cast<void**>(&solveAbortHost_), sizeof(uint32_t), hipHostMallocDefault),
            "hipHostMalloc");
    *solveAbortHost_ = 0u;
    // Every segment's tiles must be resident together: all tiles in one launch when they fit the device, else a
    // queue of segments (solveQueued_) in which each segment fits it. One tile per workgroup and per CU when every
    // tile fits that way (each wavefront then has a SIMD to itself), else tile pairs, one pair per CU.
    const uint32_t grid = 8u * ((nTiles_ + 7u) / 8u);
    solveExtra_ = maxK_ > 8u || maxLowK_ > 8u;
    const uint32_t capacity = dk::pcgSolveCapacity(solveExtra_);
    segTiles_.assign(M.segments.size(), {0u, 0u});
    uint32_t largestPairs = 0;
    for (size_t sg = 0; sg < M.segments.size(); ++sg) {
      segTiles_[sg] = {params[sg].tileBegin, params[sg].tileEnd};
      largestPairs = std::max(largestPairs, segPairs_[sg].second - segPairs_[sg].first);
    }
    solvePairs_ = grid > capacity;
    if (solvePairs_) {
      const uint32_t pairCapacity = dk::pcgSolvePairCapacity(solveExtra_);
      solveQueued_ = 8u * ((nPairs_ + 7u) / 8u) > pairCapacity;
      persistentOk_ = pairsFit_ && maxK_ <= 16u && maxLowK_ <= 16u && pairCapacity > 0 && largestPairs <= pairCapacity;
    } else {
      solveQueued_ = false;
      persistentOk_ = windowsFit && maxK_ <= 16u && maxLowK_ <= 16u && tileExtrasMax_ <= dk::kSolveExtraCap &&
                      capacity > 0;
    }
    solveOrderHost_.clear();
    segLastIters_.assign(M.segments.size(), 0u);
    if (solveQueued_) solveOrder_.allocate(sizeof(uint32_t) * nTiles_);
  }
  built_ = true;
  syncStateToDevice();
  check(hipStreamSynchronize(stream_), "build");
}

void HipSolver::uploadSegParams(double h) {
  if (h == lastH_) return;
  std::vector<dk::SegParams> params(model_.segments.size());
  check(hipMemcpyAsync(params.data(), segParams_.as<void>(), sizeof(dk::SegParams) * params.size(),
                       hipMemcpyDeviceToHost, stream_),
        "copy");
  check(hipStreamSynchronize(stream_), "sync");
  for (auto& P : params) {
    P.h = static_cast<float>(h);
    P.invH2 = static_cast<float>(1.0 / (h * h));
  }
  check(hipMemcpyAsync(segParams_.as<void>(), params.data(), sizeof(dk::SegParams) * params.size(),
                       hipMemcpyHostToDevice, stream_),
        "copy");
  check(hipStreamSynchronize(stream_), "sync");
  lastH_ = h;
}

void HipSolver::uploadGarments() {
  std::vector<dk::Garment32> g(model_.garments.size());
  for (size_t k = 0; k < g.size(); ++k)
    g[k] = dk::Garment32{model_.garments[k].phase == DRAPE_PHASE_SETTLING ? 1u : 0u,
                         static_cast<float>(model_.garments[k].assemblyTime), {0, 0}};
  if (!g.empty())
    check(hipMemcpyAsync(garments_.as<void>(), g.data(), sizeof(dk::Garment32) * g.size(), hipMemcpyHostToDevice, stream_),
          "copy");
  check(hipStreamSynchronize(stream_), "sync");
}

void HipSolver::uploadFabricsAndMass() {
  std::vector<dk::Fabric32> f(model_.fabrics.size());
  for (size_t k = 0; k < f.size(); ++k) f[k] = fabric32(model_.fabrics[k]);
  if (!f.empty())
    toDevice(fabrics_.as<void>(), f.data(), sizeof(dk::Fabric32) * f.size());
  std::vector<float> mass(nRows_, 1.0f);
  std::vector<double> massF64(nRows_, 0.0);
  for (uint32_t i = 0; i < model_.x.size(); ++i) {
    mass[rowOf_[i]] = static_cast<float>(model_.mass[i]);
    massF64[rowOf_[i]] = model_.mass[i];
  }
  toDevice(mass_.as<void>(), mass.data(), sizeof(float) * nRows_);
  toDevice(massF64_.as<void>(), massF64.data(), sizeof(double) * nRows_);
}

void HipSolver::uploadHinges() {
  std::vector<dk::Hinge32> hg(model_.hinges.size());
  for (size_t k = 0; k < hg.size(); ++k) {
    const HingeTerm& h = model_.hinges[k];
    dk::Hinge32 g{};
    for (int a = 0; a < 4; ++a) g.v[a] = rowOf_[h.v[a]];
    g.K = static_cast<float>(h.K);
    g.restAngle = 0.0f;
    hg[k] = g;
  }
  if (!hg.empty())
    toDevice(hinges_.as<void>(), hg.data(), sizeof(dk::Hinge32) * hg.size());
}

void HipSolver::uploadSeams() {
  std::vector<dk::Seam32> sm(model_.seams.size());
  for (size_t k = 0; k < sm.size(); ++k)
    sm[k] = dk::Seam32{rowOf_[model_.seams[k].a], rowOf_[model_.seams[k].b], model_.seams[k].garment,
                       model_.seams[k].startGap};
  if (!sm.empty())
    toDevice(seams_.as<void>(), sm.data(), sizeof(dk::Seam32) * sm.size());
}

void HipSolver::syncStateFromDevice() const {
  const size_t n = model_.x.size();
  std::vector<dk::F3> xs(nRows_ + 1), vs(nRows_ + 1);
  std::vector<float> cf(nRows_);
  check(hipMemcpyAsync(xs.data(), x_.as<void>(), sizeof(dk::F3) * (nRows_ + 1), hipMemcpyDeviceToHost, stream_), "copy");
  check(hipMemcpyAsync(vs.data(), v_.as<void>(), sizeof(dk::F3) * (nRows_ + 1), hipMemcpyDeviceToHost, stream_), "copy");
  check(hipMemcpyAsync(cf.data(), cforce_.as<void>(), sizeof(float) * nRows_, hipMemcpyDeviceToHost, stream_), "copy");
  check(hipStreamSynchronize(stream_), "sync");
  Model& m = const_cast<Model&>(model_);
  for (size_t i = 0; i < n; ++i) {
    const dk::F3 a = xs[rowOf_[i]], b = vs[rowOf_[i]];
    m.x[i] = {a.x, a.y, a.z};
    m.v[i] = {b.x, b.y, b.z};
    m.contactForce[i] = cf[rowOf_[i]];
  }
}

void HipSolver::syncStateToDevice() {
  std::vector<dk::F3> xs(nRows_ + 1, dk::F3{0, 0, 0}), vs(nRows_ + 1, dk::F3{0, 0, 0});
  std::vector<float> cf(nRows_, 0.0f);
  for (size_t i = 0; i < model_.x.size(); ++i) {
    xs[rowOf_[i]] = f3(model_.x[i]);
    vs[rowOf_[i]] = f3(model_.v[i]);
    cf[rowOf_[i]] = static_cast<float>(model_.contactForce[i]);
  }
  check(hipMemcpyAsync(x_.as<void>(), xs.data(), sizeof(dk::F3) * (nRows_ + 1), hipMemcpyHostToDevice, stream_), "copy");
  check(hipMemcpyAsync(v_.as<void>(), vs.data(), sizeof(dk::F3) * (nRows_ + 1), hipMemcpyHostToDevice, stream_), "copy");
  check(hipMemcpyAsync(cforce_.as<void>(), cf.data(), sizeof(float) * nRows_, hipMemcpyHostToDevice, stream_), "copy");
  check(hipStreamSynchronize(stream_), "sync");
}

void HipSolver::pollStates() const {
  check(hipGetLastError(), "kernel launch");
  check(hipMemcpyAsync(hostStates_, segState_.as<void>(), sizeof(dk::SegState) * model_.segments.size(),
                       hipMemcpyDeviceToHost, stream_),
        "copy");
  check(hipStreamSynchronize(stream_), "step");
  mirrorFresh_ = true;
  harvestEvents();
}

void HipSolver::setSpmvSampling(int every) {
  spmvEvery_ = std::max(0, every);
  spmvTimed_ = 0;
  spmvTotalMs_ = 0.0;
  spmvTotalBytes_ = 0.0;
}

// Algorithmic bytes of one PCG matrix launch over `blocks` stored blocks of `particles` rows. The matrix is stored
// symmetric: every block on or above the diagonal once with its column (40 B; (blocks + n) / 2 of them), every
// slot below it as a column and the index of its transposed block (8 B; (blocks - n) / 2). Per particle
// - two-kernel form (kPcgSpmv): row and contact flags (8 B), z and p_old gathered once (24 B), p_new and ap
//   written (24 B): 56 B;
// - one-kernel form (kPcgStep): zr, t, p gathered once (36 B), own r and ap (24 B), step (12 B), the
//   preconditioner (24 B: the symmetric inverse's upper triangle), contact flags and normal (20 B), row flags
//   (4 B) read; r, ap, p, zr, t and step written (72 B): 192 B.
double HipSolver::launchBytes(double blocks, double n) const {
  return 40.0 * 0.5 * (blocks + n) + 8.0 * 0.5 * (blocks - n) + (pcgVariant_ == 1 ? 192.0 : 56.0) * n;
}

double HipSolver::activeLaunchBytes(bool solveStarted) const {
  double b = 0.0;
  for (size_t s = 0; s < model_.segments.size(); ++s) {
    const dk::SegState& st = hostStates_[s];
    if (mirrorFresh_ && (st.newtonDone || (solveStarted && st.pcgConverged))) continue;
    b += launchBytes(segBlocks_[s], segParticles_[s]);
  }
  return b;
}

void HipSolver::pushTimed(std::pair<hipEvent_t, hipEvent_t> pr, bool solveStarted) {
  eventPending_.push_back(pr);
  eventBytes_.push_back(activeLaunchBytes(solveStarted));
}

// Reads the elapsed times of completed event pairs (the stream was synchronised) and recycles them.
void HipSolver::harvestEvents() const {
  auto* self = const_cast<HipSolver*>(this);
  for (size_t i = 0; i < self->eventPending_.size(); ++i) {
    const auto pr = self->eventPending_[i];
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, pr.first, pr.second), "event");
    spmvTotalMs_ += ms;
    spmvTotalBytes_ += self->eventBytes_[i];
    ++spmvTimed_;
    self->eventPool_.push_back(pr);
  }
  self->eventPending_.clear();
  self->eventBytes_.clear();
}

// Mean duration and mean algorithmic bytes of the timed launches: each counts the segments that were still solving
// at the host's last look (a segment that converges between that poll and the launch is counted in full). With
// no timed launch, the bytes of a launch over every segment.
void HipSolver::spmvTiming(uint64_t& launches, double& meanUs, double& bytesPerLaunch) const {
  launches = spmvTimed_;
  meanUs = spmvTimed_ ? 1000.0 * spmvTotalMs_ / double(spmvTimed_) : 0.0;
  if (spmvTimed_) {
    bytesPerLaunch = spmvTotalBytes_ / double(spmvTimed_);
    return;
  }
  const double n = double(model_.x.size());
  bytesPerLaunch = launchBytes(double(nBlocks_) - (double(nRows_) - n), n);
}

bool HipSolver::allSegments(bool (*pred)(const dk::SegState&)) const {
  for (size_t s = 0; s < model_.segments.size(); ++s)
    if (!pred(hostStates_[s])) return false;
  return true;
}

void HipSolver::Captured::release() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
}

template <class F>
void HipSolver::launchCaptured(Captured& c, F&& launches) {
  if (!c.exec) {
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "capture");
    launches();
    check(hipStreamEndCapture(stream_, &c.graph), "capture");
    check(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0), "graph instantiate");
  }
  check(hipGraphLaunch(c.exec, stream_), "graph launch");
}

void HipSolver::releaseGraph() {
  if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graphExec_ = nullptr;
  graph_ = nullptr;
  newtonPre_.release();
  newtonPost_.release();
  refineExec_.release();
  for (auto& c : tuneExec_) c.release();
}

void HipSolver::setPcgVariant(int variant) {
  if (variant != 0 && variant != 1) throw std::invalid_argument("pcg variant must be 0 or 1");
  pcgVariant_ = variant;
}

void HipSolver::setPcgShape(int waves, int slots, int halo) {
  if (waves == 0 && slots == 0 && halo == 0) {
    pcgShape_ = -1;
    return;
  }
  for (int i = 0; i < kShapeCount; ++i) {
    const dk::PcgShape& sh = dk::kPcgShapes[i];
    const int code = sh.persistent ? 2 : sh.halo ? 1 : 0;
    if (int(sh.waves) == waves && int(sh.slots) == slots && code == halo) {
      pcgShape_ = i;
      return;
    }
  }
  throw std::invalid_argument(
      "PCG launch shape must be 0x0 (tune), 1x8, 1x4, 2x8 or 4x8 (wavefronts x slots), 1x8 / 1x4 with the halo, or the "
      "persistent 1x8 solve (halo 2)");
}

void HipSolver::pcgShape(int& waves, int& slots, int& halo) const {
  const int i = shapeIndex();
  waves = i < 0 ? 0 : int(dk::kPcgShapes[i].waves);
  slots = i < 0 ? 0 : int(dk::kPcgShapes[i].slots);
  halo = i < 0 ? 0 : dk::kPcgShapes[i].persistent ? 2 : dk::kPcgShapes[i].halo ? 1 : 0;
}

void HipSolver::runPcg() {
  if (pcgVariant_ == 1)
    runPcgSingleKernel();
  else
    runPcgStandard();
}

static uint32_t maxCgOf(const Model& m) {
  uint32_t maxCg = 0;
  for (const auto& sg : m.segments) maxCg = std::max<uint32_t>(maxCg, static_cast<uint32_t>(std::max(0, sg.settings.maxCg)));
  return maxCg;
}

void HipSolver::runPcgStandard() {
  dk::launchPcgInit(view_, stream_);
  const uint32_t maxCg = maxCgOf(model_);
  // First chunk sized from the previous solve; later chunks short. Iterations past convergence are no-ops.
  uint32_t k = 0;
  uint32_t chunk = std::max<uint32_t>(8, lastPcgIterations_ * 3 / 4);
  bool polled = false;  // the host mirror holds this solve's convergence flags
  while (k < maxCg) {
    const uint32_t end = std::min(maxCg, k + chunk);
    for (; k < end; ++k) {
      if (spmvEvery_ > 0 && k % uint32_t(spmvEvery_) == 0) {
        const auto pr = takeEvents();
        dk::launchPcgSpmv(view_, k, stream_, pr.first, pr.second);
        pushTimed(pr, polled);
        dk::launchPcgUpdate(view_, k, stream_);
      } else {
        dk::launchPcgIteration(view_, k, stream_);
      }
    }
    pollStates();
    polled = true;
    if (allSegments([](const dk::SegState& s) { return s.newtonDone != 0u || s.pcgConverged != 0u; })) break;
    chunk = 16;
  }
  uint32_t total = 0;
  for (size_t s = 0; s < model_.segments.size(); ++s) total = std::max(total, hostStates_[s].pcgIter);
  lastPcgIterations_ = total;
}

std::pair<hipEvent_t, hipEvent_t> HipSolver::takeEvents() {
  if (eventPool_.empty()) {
    hipEvent_t a, b;
    check(hipEventCreate(&a), "event");
    check(hipEventCreate(&b), "event");
    eventPool_.push_back({a, b});
  }
  const auto pr = eventPool_.back();
  eventPool_.pop_back();
  return pr;
}

void HipSolver::timedPcgStep(uint32_t parity, dk::PcgShape shape, bool solveStarted) {
  const auto pr = takeEvents();
  dk::launchPcgStep(view_, parity, shape, stream_, pr.first, pr.second);
  pushTimed(pr, solveStarted);
}

// Solves run before a launch shape is chosen: chunks of kTuneChunk launches rotate over the shapes of
// dk::kPcgShapes that the matrix admits (the halo shapes need every tile's window to fit; the shape with the fewest
// samples goes next), each bracketed by HIP events and followed by a poll. A chunk counts only if the solve was
// still running at its end, so every timed launch did a full iteration; each shape's first chunk is not counted (it
// loads the kernel's code and captures the chunk's graph). The persistent solve takes the rest of a solve as its
// sample and counts its iterations. Once every shape has four chunks' worth of counted iterations the fastest mean
// is kept; after kTuneSolves solves without that (solves too short to fill a chunk), the fastest shape sampled so far
// is kept, else 1x8. The shape never changes the results.
void HipSolver::runPcgTuning(uint32_t maxCg) {
  uint32_t launched = 0;
  while (launched < maxCg + 1) {
    int o = 0;
    for (int i = 1; i < kShapeCount; ++i)
      if (shapeAvailable(i) && tuneLaunches_[i] < tuneLaunches_[o]) o = i;
    const int si = tunedShape_ >= 0 ? tunedShape_ : o;
    const dk::PcgShape shape = dk::kPcgShapes[si];
    if (shape.persistent) {
      // The persistent candidate runs the rest of the solve in one launch and counts as a sample of that many
      // iterations when it ran at least a chunk's worth.
      pollStates();
      uint32_t before = 0;
      for (size_t g = 0; g < model_.segments.size(); ++g) before = std::max(before, hostStates_[g].pcgIter);
      const auto pr = takeEvents();
      check(hipEventRecord(pr.first, stream_), "event");
      const bool ok = runPcgPersistent(false, false);
      check(hipEventRecord(pr.second, stream_), "event");
      if (!ok) {
        eventPool_.push_back(pr);
        redoSolveWithLaunches(maxCg);
        return;
      }
      pollStates();
      uint32_t after = 0;
      for (size_t g = 0; g < model_.segments.size(); ++g) after = std::max(after, hostStates_[g].pcgIter);
      if (tunedShape_ < 0 && after - before >= kTuneChunk && !tuneWarm_[si]) {
        tuneWarm_[si] = true;  // the first persistent solve loads the kernel's code: not counted
      } else if (tunedShape_ < 0 && after - before >= kTuneChunk) {
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, pr.first, pr.second), "event time");
        tuneMs_[si] += ms;
        tuneLaunches_[si] += after - before;
        bool sampled = true;
        for (int i = 0; i < kShapeCount; ++i) sampled = sampled && (!shapeAvailable(i) || tuneLaunches_[i] >= 4u * kTuneChunk);
        if (sampled) tunedShape_ = fastestSampledShape();
      }
      eventPool_.push_back(pr);
      break;
    }
    const auto pr = takeEvents();
    check(hipEventRecord(pr.first, stream_), "event");
    // Each shape's chunk is a captured graph, as the solves after tuning replay them: the events then time the
    // launches themselves, not how fast the host can submit them.
    if (useGraph_)
      launchCaptured(tuneExec_[si], [&] {
        for (uint32_t i = 0; i < kTuneChunk; ++i) dk::launchPcgStep(view_, i & 1u, shape, stream_);
      });
    else
      for (uint32_t i = 0; i < kTuneChunk; ++i) dk::launchPcgStep(view_, i & 1u, shape, stream_);
    check(hipEventRecord(pr.second, stream_), "event");
    launched += kTuneChunk;
    pollStates();
    const bool done = allSegments([](const dk::SegState& s) { return s.newtonDone != 0u || s.pcgConverged != 0u; });
    if (!done && tunedShape_ < 0 && !tuneWarm_[o]) {
      tuneWarm_[o] = true;  // the shape's first chunk loads its kernel and captures its graph: not counted
    } else if (!done && tunedShape_ < 0) {
      float ms = 0.0f;
      check(hipEventElapsedTime(&ms, pr.first, pr.second), "event time");
      tuneMs_[o] += ms;
      tuneLaunches_[o] += kTuneChunk;
      bool sampled = true;
      for (int i = 0; i < kShapeCount; ++i) sampled = sampled && (!shapeAvailable(i) || tuneLaunches_[i] >= 4u * kTuneChunk);
      if (sampled) tunedShape_ = fastestSampledShape();
    }
    eventPool_.push_back(pr);
    if (done) break;
  }
  if (tunedShape_ < 0 && ++tuneSolves_ >= kTuneSolves) tunedShape_ = fastestSampledShape();
  const int refineIdx = tunedShape_ >= 0 ? tunedShape_ : 0;
  if (dk::kPcgShapes[refineIdx].persistent && shapeAvailable(refineIdx)) {
    if (!runPcgPersistent(true, false)) redoSolveWithLaunches(maxCg);
    return;
  }
  const dk::PcgShape shape = dk::kPcgShapes[dk::kPcgShapes[refineIdx].persistent ? 0 : refineIdx];
  for (uint32_t r = 0; r < refineRestarts_; ++r) {
    dk::launchPcgRefine(view_, refineRestarts_, nullptr, stream_);
    pollStates();
    if (allSegments([](const dk::SegState& s) { return s.newtonDone != 0u || s.pcgConverged != 0u; })) break;
    for (uint32_t n = 0; n < maxCg + 1; n += kTuneChunk) {
      for (uint32_t i = 0; i < kTuneChunk; ++i) dk::launchPcgStep(view_, i & 1u, shape, stream_);
      pollStates();
      if (allSegments([](const dk::SegState& s) { return s.newtonDone != 0u || s.pcgConverged != 0u; })) break;
    }
  }
  uint32_t total = 0;
  for (size_t s = 0; s < model_.segments.size(); ++s) total = std::max(total, hostStates_[s].pcgIter);
  lastPcgIterations_ = total;
}

// The shape with the lowest mean chunk time among those sampled (1x8 when none is).
int HipSolver::fastestSampledShape() const {
  int best = -1;
  for (int i = 0; i < kShapeCount; ++i)
    if (shapeAvailable(i) && tuneLaunches_[i] > 0 &&
        (best < 0 || tuneMs_[i] / tuneLaunches_[i] < tuneMs_[best] / tuneLaunches_[best]))
      best = i;
  return best < 0 ? 0 : best;
}

// The rest of the current solve with the persistent kernel: one launch runs every segment's solve to its end, then
// (when `refine`) each FP64 true-residual refinement round is a decision and, for the segments that continue, one
// more launch. The first round goes out right behind the solve, and the host looks at the flags once after both:
// the abort flag is cleared only before the solve, so a solve that gave up makes the round a no-op (no segment
// continues, and the continuation's waits give up at once) and is still seen. With `timeIt` the launches between
// two looks are bracketed by events and counted in the iteration timing by the iterations they ran. A launch that
// fails, or whose barrier gave up (abort flag), switches the persistent form off for this solver and returns false:
// the caller then re-runs the solve.
bool HipSolver::runPcgPersistent(bool refine, bool timeIt) {
  dk::SolveCtl c{};
  c.ctr = solveCtr_.as<uint32_t>();
  c.segCtr = solveSegCtr_.as<uint32_t>();
  c.abortFlag = c.ctr + size_t(solveCtrLines_) * 64u;
  c.spinCap = solveSpinCap_;
  const auto fail = [&] {
    persistentOk_ = false;
    ++persistentAborts_;
    if (pcgShape_ >= 0 && dk::kPcgShapes[pcgShape_].persistent) pcgShape_ = -1;
    if (tunedShape_ >= 0 && dk::kPcgShapes[tunedShape_].persistent) tunedShape_ = fastestSampledShape();
    return false;
  };
  // Iterations done before each look, per segment (the first launch follows kPcgStepInit when timed: none).
  // A segment whose Newton loop had already finished keeps its last solve's count (kPcgStepInit leaves it).
  std::vector<uint32_t> before(model_.segments.size(), 0u);
  for (size_t g = 0; g < model_.segments.size(); ++g)
    if (mirrorFresh_ && hostStates_[g].newtonDone) before[g] = hostStates_[g].pcgIter;
  c.order = solveQueued_ ? solveOrder_.as<uint32_t>() : nullptr;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> timed;  // the launches since the last look
  const auto launch = [&](bool clearAbort) -> bool {
    if (solveQueued_) queueSegments();
    std::pair<hipEvent_t, hipEvent_t> pr{};
    if (timeIt) {
      pr = takeEvents();
      check(hipEventRecord(pr.first, stream_), "event");
    }
    const hipError_t e = solvePairs_ ? dk::launchPcgSolvePairs(view_, c, solveCtrLines_, solveExtra_, clearAbort, stream_)
                                     : dk::launchPcgSolve(view_, c, solveCtrLines_, solveExtra_, clearAbort, stream_);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      if (timeIt) {
        timed.push_back(pr);
        for (const auto& t : timed) eventPool_.push_back(t);
        timed.clear();
      }
      return false;
    }
    if (timeIt) {
      check(hipEventRecord(pr.second, stream_), "event");
      timed.push_back(pr);
    }
    return true;
  };
  const auto look = [&]() -> bool {
    check(hipMemcpyAsync(solveAbortHost_, c.abortFlag, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_), "copy");
    pollStates();
    if (*solveAbortHost_) {
      for (const auto& pr : timed) eventPool_.push_back(pr);
      timed.clear();
      return false;
    }
    if (timeIt) {
      // Counted as many iterations as its longest segment ran, carrying the bytes every segment's own iterations
      // move: bytes over time is then the launches' achieved rate.
      uint32_t most = 0;
      double bytes = 0.0;
      for (size_t g = 0; g < model_.segments.size(); ++g) {
        const uint32_t ran = hostStates_[g].pcgIter - std::min(before[g], hostStates_[g].pcgIter);
        most = std::max(most, ran);
        bytes += double(ran) * launchBytes(segBlocks_[g], segParticles_[g]);
        before[g] = hostStates_[g].pcgIter;
      }
      float ms = 0.0f;
      for (const auto& pr : timed) {
        float part = 0.0f;
        check(hipEventElapsedTime(&part, pr.first, pr.second), "event");
        ms += part;
        eventPool_.push_back(pr);
      }
      timed.clear();
      if (most > 0) {
        spmvTotalMs_ += ms;
        spmvTimed_ += most;
        spmvTotalBytes_ += bytes;
      }
    }
    return true;
  };
  const auto continued = [&] {
    bool any = false;
    for (size_t g = 0; g < model_.segments.size(); ++g) any = any || hostStates_[g].restartActive != 0u;
    return any;
  };
  if (!launch(true)) return fail();
  if (refine && refineRestarts_ > 0) {
    dk::launchPcgRefine(view_, refineRestarts_, c.abortFlag, stream_);
    if (!launch(false)) return fail();
  }
  if (!look()) return fail();
  if (refine)
    for (uint32_t r = 1; r < refineRestarts_ && continued(); ++r) {
      dk::launchPcgRefine(view_, refineRestarts_, c.abortFlag, stream_);
      if (!launch(false) || !look()) return fail();
    }
  uint32_t total = 0;
  for (size_t g = 0; g < model_.segments.size(); ++g) {
    total = std::max(total, hostStates_[g].pcgIter);
    segLastIters_[g] = hostStates_[g].pcgIter;
  }
  lastPcgIterations_ = total;
  return true;
}

// The queue of a persistent launch with more tiles than the device holds at once: segments by decreasing expected
// work (tiles times the iterations of the segment's last solve), each segment's tiles consecutive, so that the
// longest solves start first and the short ones fill the CUs they leave; segments whose Newton loop has finished
// (their tiles leave at once) go last. Uploaded only when it changes.
void HipSolver::queueSegments() {
  std::vector<std::pair<double, uint32_t>> work;
  for (uint32_t sg = 0; sg < segTiles_.size(); ++sg) {
    const double tiles = double(segTiles_[sg].second - segTiles_[sg].first);
    const bool idle = mirrorFresh_ && hostStates_[sg].newtonDone != 0u;
    work.push_back({idle ? -1.0 : tiles * double(std::max<uint32_t>(1u, segLastIters_[sg])), sg});
  }
  std::stable_sort(work.begin(), work.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  std::vector<uint32_t> order;
  order.reserve(nTiles_);
  for (const auto& w : work) {
    const auto& span = solvePairs_ ? segPairs_[w.second] : segTiles_[w.second];  // workgroups: pairs or tiles
    for (uint32_t t = span.first; t < span.second; ++t) order.push_back(t);
  }
  if (order == solveOrderHost_) return;
  solveOrderHost_ = order;
  check(hipMemcpyAsync(solveOrder_.as<void>(), solveOrderHost_.data(), sizeof(uint32_t) * order.size(),
                       hipMemcpyHostToDevice, stream_),
        "copy");
}

// After the persistent form gave up: the whole solve again from its start, one launch per iteration (the results
// are those the persistent form would have given).
void HipSolver::redoSolveWithLaunches(uint32_t maxCg) {
  (void)maxCg;
  runPcgSingleKernel();
}

// Chunks of kGraphChunk one-kernel iterations, replayed from a captured hipGraph (or launched eagerly). The
// iteration index lives on the device, so one graph serves every chunk; iterations after convergence return at
// once. The host reads the converged flags after the first batch (sized from the previous solve) and then after
// every chunk.
void HipSolver::runPcgSingleKernel() {
  const int shapeIdx = shapeIndex();
  if (shapeIdx >= 0 && !shapeAvailable(shapeIdx)) {
    if (dk::kPcgShapes[shapeIdx].persistent)
      throw std::invalid_argument(
          "PCG persistent solve: needs every 256-row tile's columns within four row ranges of 1536 rows in all (or, "
          "with more tiles than CUs, every tile pair's within eight ranges of 3072 rows), at most 16 + 16 slots per "
          "64-row slice, each scene's tiles resident together (512 per scene), and must not have given up before");
    throw std::invalid_argument("PCG halo shape: a 256-row tile of this matrix reaches more than 1536 rows");
  }
  dk::launchPcgStepInit(view_, stream_);
  const uint32_t maxCg = maxCgOf(model_);
  if (shapeIdx < 0) {
    runPcgTuning(maxCg);
    return;
  }
  if (dk::kPcgShapes[shapeIdx].persistent) {
    if (!runPcgPersistent(true, spmvEvery_ > 0)) redoSolveWithLaunches(maxCg);
    return;
  }
  const dk::PcgShape shape = dk::kPcgShapes[shapeIdx];
  const bool sampling = spmvEvery_ > 0;
  if (useGraph_ && (!graphExec_ || graphShape_ != shapeIdx)) {
    releaseGraph();
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "capture");
    for (uint32_t i = 0; i < kGraphChunk; ++i) dk::launchPcgStep(view_, i & 1u, shape, stream_);
    check(hipStreamEndCapture(stream_, &graph_), "capture");
    check(hipGraphInstantiate(&graphExec_, graph_, nullptr, nullptr, 0), "graph instantiate");
    graphShape_ = shapeIdx;
  }
  // Chunks until every segment's solve has ended: a first batch of `chunks`, then one chunk per look at the flags.
  const auto runChunks = [&](uint32_t chunks, uint32_t budget) {
    uint32_t launched = 0;
    bool polled = false;  // the host mirror holds this solve's convergence flags
    std::vector<std::pair<hipEvent_t, hipEvent_t>> chunkEvents;
    while (launched < budget + 1) {
      for (uint32_t c = 0; c < chunks; ++c) {
        if (useGraph_) {
          // With sampling, each replayed chunk is bracketed by events: a chunk's time over its 64 launches is the
          // per-iteration time with the gaps between launches, counted below when every launch did an iteration.
          std::pair<hipEvent_t, hipEvent_t> pr{};
          if (sampling) {
            pr = takeEvents();
            check(hipEventRecord(pr.first, stream_), "event");
          }
          check(hipGraphLaunch(graphExec_, stream_), "graph launch");
          if (sampling) {
            check(hipEventRecord(pr.second, stream_), "event");
            chunkEvents.push_back(pr);
          }
        } else {
          for (uint32_t i = 0; i < kGraphChunk; ++i) {
            if (sampling && (launched + i) % uint32_t(spmvEvery_) == 1)
              timedPcgStep(i & 1u, shape, polled);
            else
              dk::launchPcgStep(view_, i & 1u, shape, stream_);
          }
        }
        launched += kGraphChunk;
      }
      pollStates();
      polled = true;
      const bool ended = allSegments([](const dk::SegState& s) { return s.newtonDone != 0u || s.pcgConverged != 0u; });
      // The batch's chunks count only when a solve was still running at this look (no launch found it ended).
      for (const auto& pr : chunkEvents) {
        if (!ended) {
          float ms = 0.0f;
          check(hipEventElapsedTime(&ms, pr.first, pr.second), "event");
          spmvTotalMs_ += ms;
          spmvTimed_ += kGraphChunk;
          spmvTotalBytes_ += double(kGraphChunk) * activeLaunchBytes(true);
        }
        eventPool_.push_back(pr);
      }
      chunkEvents.clear();
      if (ended) break;
      chunks = 1;
    }
  };
  // First batch: five eighths of the previous solve's iterations, in whole chunks (measured on the 65k sheet with
  // 64-launch chunks against 3/4 and 7/8: a batch that overshoots the solve costs more in launches that find it
  // ended than one more look at the flags costs).
  runChunks(std::max<uint32_t>(1, (lastPcgIterations_ * 5 / 8) / kGraphChunk), maxCg);
  // FP64 true-residual refinement: a segment whose solve stopped on the FP32 recursion's test while its true
  // residual is still above tol r0 continues from the true residual (kTrueResidual, kRefineDecide).
  // The decision and a first short chunk of the continuation go out together (the chunk's launches are no-ops for
  // a segment that does not continue), so a look at the flags follows both.
  for (uint32_t r = 0; r < refineRestarts_; ++r) {
    dk::launchPcgRefine(view_, refineRestarts_, nullptr, stream_);
    if (useGraph_) {
      if (!refineExec_.exec || refineShape_ != shapeIdx) {
        refineExec_.release();
        launchCaptured(refineExec_, [&] {
          for (uint32_t i = 0; i < kRefineChunk; ++i) dk::launchPcgStep(view_, i & 1u, shape, stream_);
        });
        refineShape_ = shapeIdx;
      } else {
        check(hipGraphLaunch(refineExec_.exec, stream_), "graph launch");
      }
    } else {
      for (uint32_t i = 0; i < kRefineChunk; ++i) dk::launchPcgStep(view_, i & 1u, shape, stream_);
    }
    pollStates();
    bool continued = false;
    for (size_t g = 0; g < model_.segments.size(); ++g) continued = continued || hostStates_[g].restartActive != 0u;
    if (!continued) break;
    if (!allSegments([](const dk::SegState& s) { return s.newtonDone != 0u || s.pcgConverged != 0u; }))
      runChunks(1, maxCg);
  }
  uint32_t total = 0;
  for (size_t s = 0; s < model_.segments.size(); ++s) total = std::max(total, hostStates_[s].pcgIter);
  lastPcgIterations_ = total;
}

void HipSolver::substep(double h, int iterations) {
  mirrorFresh_ = false;  // every segment starts the substep's Newton loop
  uploadSegParams(h);
  if (net_) {
    dk::launchWarmFeatures(view_, stream_);
    const ws::FrameView f{view_.x, view_.v, view_.dvLast, view_.phiN};
    net_->predict(meshGraph_->view(), f, warm_.as<dk::F3>(), stream_);
  }
  dk::launchSubstepBegin(view_, net_ != nullptr, stream_);
  if (net_) dk::launchWarmCheck(view_, stream_);
  // Each Newton iteration: the pre-solve launches, the PCG solve (chunks, polled), then the post-solve launches.
  // With graphs, the post-solve part runs the first kGraphTrials line-search trials, the apply and the end test
  // in one replay and the host polls once; a line search that is still undecided then continues eagerly, one poll
  // per trial, and its apply and end test run after it (they act only on a decided, not yet applied segment).
  const auto lsDecided = [](const dk::SegState& s) { return s.newtonDone != 0u || s.lsDone != 0u; };
  for (int it = 0; it < iterations; ++it) {
    const auto pre = [&] {
      dk::launchHoldContacts(view_, stream_);
      dk::launchAssemble(view_, maxK_, true, stream_);
      dk::launchNewtonBegin(view_, stream_);
    };
    if (useGraph_)
      launchCaptured(newtonPre_, pre);
    else
      pre();
    runPcg();
    int trials = 0;
    if (useGraph_) {
      launchCaptured(newtonPost_, [&] {
        dk::launchContactForceSlope(view_, stream_);
        dk::launchNewtonSlope(view_, stream_);
        for (int ls = 0; ls < kGraphTrials; ++ls) {
          dk::launchEnergyTrial(view_, stream_);
          dk::launchLineSearchDecide(view_, stream_);
        }
        dk::launchApply(view_, stream_);
        dk::launchNewtonEnd(view_, stream_);
      });
      trials = kGraphTrials;
      pollStates();
    } else {
      dk::launchContactForceSlope(view_, stream_);
      dk::launchNewtonSlope(view_, stream_);
    }
    if (!useGraph_ || !allSegments(lsDecided)) {
      for (; trials < 20; ++trials) {
        dk::launchEnergyTrial(view_, stream_);
        dk::launchLineSearchDecide(view_, stream_);
        pollStates();
        if (allSegments(lsDecided)) break;
      }
      dk::launchApply(view_, stream_);
      dk::launchNewtonEnd(view_, stream_);
      pollStates();
    }
    if (allSegments([](const dk::SegState& s) { return s.newtonDone != 0u; })) break;
  }
  dk::launchFinish(view_, stream_);
  bool assembling = false;
  for (auto& g : model_.garments)
    if (g.phase == DRAPE_PHASE_ASSEMBLING) {
      g.assemblyTime += h;
      assembling = true;
    }
  if (assembling) uploadGarments();
}

void HipSolver::step(double dt, int32_t substeps, int32_t iterations) {
  requireBuilt();
  if (model_.x.empty()) return;
  const int ns = std::max(1, substeps);
  const double h = dt / ns;
  const int its = std::max(1, iterations);
  for (int s = 0; s < ns; ++s) substep(h, its);
  check(hipStreamSynchronize(stream_), "step");
}

void HipSolver::readPositions(float* out, size_t n) const {
  requireBuilt();
  if (n != model_.x.size()) throw std::invalid_argument("readPositions: span size must equal particleCount()");
  std::vector<dk::F3> xs(nRows_ + 1);
  check(hipMemcpyAsync(xs.data(), x_.as<void>(), sizeof(dk::F3) * (nRows_ + 1), hipMemcpyDeviceToHost, stream_), "copy");
  check(hipStreamSynchronize(stream_), "sync");
  for (size_t i = 0; i < n; ++i) {
    const dk::F3 a = xs[rowOf_[i]];
    out[3 * i] = a.x;
    out[3 * i + 1] = a.y;
    out[3 * i + 2] = a.z;
  }
}

void HipSolver::readVelocities(float* out, size_t n) const {
  requireBuilt();
  if (n != model_.x.size()) throw std::invalid_argument("span size must equal particleCount()");
  std::vector<dk::F3> vs(nRows_ + 1);
  check(hipMemcpyAsync(vs.data(), v_.as<void>(), sizeof(dk::F3) * (nRows_ + 1), hipMemcpyDeviceToHost, stream_), "copy");
  check(hipStreamSynchronize(stream_), "sync");
  for (size_t i = 0; i < n; ++i) {
    const dk::F3 a = vs[rowOf_[i]];
    out[3 * i] = a.x;
    out[3 * i + 1] = a.y;
    out[3 * i + 2] = a.z;
  }
}

// vertexNormals (libs/render/src/normals.cpp:5-24) of the current positions: every garment's triangles (one
// stretch term per mesh triangle, in triangle order), every seam pair welded, in internal rows; copied out with
// the positions in one pass.
void HipSolver::readPositionsNormals(float* positions, float* normals, size_t n) const {
  requireBuilt();
  if (n != model_.x.size()) throw std::invalid_argument("span size must equal particleCount()");
  if (!normalMesh_) {
    std::vector<uint32_t> tri(3 * model_.stretch.size()), weld(2 * model_.seams.size());
    for (size_t t = 0; t < model_.stretch.size(); ++t)
      for (int k = 0; k < 3; ++k) tri[3 * t + k] = rowOf_[model_.stretch[t].v[k]];
    for (size_t k = 0; k < model_.seams.size(); ++k) {
      weld[2 * k] = rowOf_[model_.seams[k].a];
      weld[2 * k + 1] = rowOf_[model_.seams[k].b];
    }
    normalMesh_ = std::make_unique<nrm::Mesh>(nRows_, tri.data(), uint32_t(model_.stretch.size()), weld.data(),
                                              uint32_t(model_.seams.size()));
    normalsOut_.allocate(sizeof(dk::F3) * (nRows_ + 1));
  }
  normalMesh_->compute(x_.as<dk::F3>(), normalsOut_.as<dk::F3>(), stream_);
  std::vector<dk::F3> xs(nRows_), ns(nRows_);
  check(hipMemcpyAsync(xs.data(), x_.as<void>(), sizeof(dk::F3) * nRows_, hipMemcpyDeviceToHost, stream_), "copy");
  check(hipMemcpyAsync(ns.data(), normalsOut_.as<void>(), sizeof(dk::F3) * nRows_, hipMemcpyDeviceToHost, stream_),
        "copy");
  check(hipStreamSynchronize(stream_), "sync");
  for (size_t i = 0; i < n; ++i) {
    const dk::F3 a = xs[rowOf_[i]], b = ns[rowOf_[i]];
    if (positions) {
      positions[3 * i] = a.x;
      positions[3 * i + 1] = a.y;
      positions[3 * i + 2] = a.z;
    }
    normals[3 * i] = b.x;
    normals[3 * i + 1] = b.y;
    normals[3 * i + 2] = b.z;
  }
}

// SimModel::stats (sim_model.cpp:222-241) reduced on the device: per-tile FP64 partials, summed per segment in a
// fixed order (kStatsTile, kStatsSegment); only kStatsPerSegment doubles per segment are copied back.
std::vector<double> HipSolver::segmentPartials() const {
  const dk::View& v = view_;
  dk::launchStats(v, massF64_.as<double>(), statsTile_.as<double>(), statsSeg_.as<double>(), stream_);
  std::vector<double> out(dk::kStatsPerSegment * model_.segments.size());
  check(hipMemcpyAsync(out.data(), statsSeg_.as<void>(), sizeof(double) * out.size(), hipMemcpyDeviceToHost, stream_),
        "copy");
  check(hipStreamSynchronize(stream_), "sync");
  return out;
}

namespace {
drape_stats combineSegments(const double* p, uint32_t count) {
  drape_stats st{};
  double sumV2 = 0.0, moving = 0.0;
  for (uint32_t s = 0; s < count; ++s, p += dk::kStatsPerSegment) {
    st.kinetic_energy += p[0];
    sumV2 += p[1];
    moving += p[2];
    if (p[3] != 0.0) st.has_non_finite = 1;
    st.max_body_penetration = std::max(st.max_body_penetration, p[4]);
    st.max_seam_gap = std::max(st.max_seam_gap, p[5]);
  }
  st.rms_speed = moving > 0.0 ? std::sqrt(sumV2 / moving) : 0.0;
  return st;
}
}  // namespace

drape_stats HipSolver::stats() const {
  requireBuilt();
  const std::vector<double> p = segmentPartials();
  return combineSegments(p.data(), static_cast<uint32_t>(model_.segments.size()));
}

void HipSolver::segmentStats(uint32_t first, uint32_t count, drape_stats* out) const {
  requireBuilt();
  if (uint64_t(first) + count > model_.segments.size()) throw std::out_of_range("segment range out of range");
  const std::vector<double> p = segmentPartials();
  for (uint32_t s = 0; s < count; ++s) out[s] = combineSegments(p.data() + size_t(first + s) * dk::kStatsPerSegment, 1);
}

// Snapshot blob: u64 particles, u64 garments, u64 seams; positions, velocities (3 floats each); phases (int32);
// assemblyTime (double); seamStartGap (float); contactForce (double) — Snapshot (solver.hpp:76-82).
size_t HipSolver::snapshotSize() const {
  requireBuilt();
  return 3 * sizeof(uint64_t) + model_.x.size() * 24 + model_.garments.size() * 12 + model_.seams.size() * 4 +
         model_.x.size() * 8;
}

void HipSolver::snapshotSave(void* out, size_t bytes) const {
  requireBuilt();
  if (bytes != snapshotSize()) throw std::invalid_argument("snapshot buffer size mismatch");
  syncStateFromDevice();
  const Model& m = model_;
  char* p = static_cast<char*>(out);
  const uint64_t head[3] = {m.x.size(), m.garments.size(), m.seams.size()};
  std::memcpy(p, head, sizeof(head));
  p += sizeof(head);
  std::memcpy(p, m.x.data(), m.x.size() * 12);
  p += m.x.size() * 12;
  std::memcpy(p, m.v.data(), m.v.size() * 12);
  p += m.v.size() * 12;
  for (const auto& g : m.garments) {
    std::memcpy(p, &g.phase, 4);
    p += 4;
  }
  for (const auto& g : m.garments) {
    std::memcpy(p, &g.assemblyTime, 8);
    p += 8;
  }
  for (const auto& s : m.seams) {
    std::memcpy(p, &s.startGap, 4);
    p += 4;
  }
  std::memcpy(p, m.contactForce.data(), m.contactForce.size() * 8);
}

void HipSolver::snapshotRestore(const void* in, size_t bytes) {
  requireBuilt();
  Model& m = model_;
  const char* p = static_cast<const char*>(in);
  if (bytes < 3 * sizeof(uint64_t)) throw std::invalid_argument("snapshot does not match this scene");
  uint64_t head[3];
  std::memcpy(head, p, sizeof(head));
  p += sizeof(head);
  const size_t expect = 3 * sizeof(uint64_t) + head[0] * 24 + head[1] * 12 + head[2] * 4 + head[0] * 8;
  if (head[0] != m.x.size() || head[1] != m.garments.size() || bytes != expect)
    throw std::invalid_argument("snapshot does not match this scene");
  std::memcpy(m.x.data(), p, m.x.size() * 12);
  p += m.x.size() * 12;
  std::memcpy(m.v.data(), p, m.v.size() * 12);
  p += m.v.size() * 12;
  for (auto& g : m.garments) {
    std::memcpy(&g.phase, p, 4);
    p += 4;
  }
  for (auto& g : m.garments) {
    std::memcpy(&g.assemblyTime, p, 8);
    p += 8;
  }
  if (head[2] == m.seams.size()) {
    for (auto& s : m.seams) {
      std::memcpy(&s.startGap, p, 4);
      p += 4;
    }
  } else {
    p += head[2] * 4;
  }
  std::memcpy(m.contactForce.data(), p, m.contactForce.size() * 8);
  syncStateToDevice();
  uploadGarments();
  uploadSeams();
  // The realised dv is not part of the reference's snapshot: the learned warm start restarts from dv = 0.
  check(hipMemsetAsync(dvLast_.as<void>(), 0, dvLast_.bytes(), stream_), "memset");
  check(hipStreamSynchronize(stream_), "sync");
}

void HipSolver::updateFabric(const std::string& garment, const std::vector<Fabric>& fabrics) {
  requireBuilt();
  Garment& g = model_.garments[model_.garmentIndex(garment)];
  if (fabrics.size() != g.pieceCount) throw std::invalid_argument("one fabric per pattern piece required");
  std::copy(fabrics.begin(), fabrics.end(), model_.fabrics.begin() + g.fabricOffset);
  for (uint32_t i = g.offset; i < g.offset + g.count; ++i) model_.updateMass(i);
  for (auto& h : model_.hinges) {
    const uint32_t fi = h.side[0].fabric;
    if (fi >= g.fabricOffset && fi < g.fabricOffset + g.pieceCount) h.K = model_.hingeStiffness(h);
  }
  check(hipStreamSynchronize(stream_), "sync");
  uploadFabricsAndMass();
  uploadHinges();
}

void HipSolver::setPhase(const std::string& garment, int32_t phase) {
  requireBuilt();
  const size_t gi = model_.garmentIndex(garment);
  Garment& g = model_.garments[gi];
  g.phase = phase;
  if (phase == DRAPE_PHASE_ASSEMBLING) {
    g.assemblyTime = 0;
    syncStateFromDevice();
    for (auto& sp : model_.seams) {
      if (sp.garment != gi) continue;
      const Vec3f a = model_.x[sp.a], b = model_.x[sp.b];
      const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
      sp.startGap = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    uploadSeams();
  }
  uploadGarments();
}

void HipSolver::range(const std::string& garment, uint32_t& offset, uint32_t& count) const {
  requireBuilt();
  const Garment& g = model_.garments[model_.garmentIndex(garment)];
  offset = g.offset;
  count = g.count;
}

void HipSolver::segmentRange(uint32_t s, uint32_t& offset, uint32_t& count) const {
  requireBuilt();
  if (s >= model_.segments.size()) throw std::out_of_range("segment index out of range");
  offset = model_.segments[s].particleBegin;
  count = model_.segments[s].particleCount;
}

drape_counters HipSolver::counters() const {
  requireBuilt();
  pollStates();
  drape_counters c{};
  for (size_t s = 0; s < model_.segments.size(); ++s) {
    const dk::SegState& st = hostStates_[s];
    c.substeps += st.cntSubsteps;
    c.newton_iterations += st.cntNewton;
    c.pcg_iterations += st.cntPcg;
    c.line_search_trials += st.cntLs;
    c.warm_start_fallbacks += st.cntWarmRejected;
    c.warm_start_accepted += st.cntWarmAccepted;
  }
  return c;
}

drape_pcg_report HipSolver::pcgReport() const {
  requireBuilt();
  pollStates();
  drape_pcg_report r{};
  double sum = 0.0;
  for (size_t s = 0; s < model_.segments.size(); ++s) {
    const dk::SegState& st = hostStates_[s];
    r.solves += st.cntSolves;
    r.exit_tolerance += st.cntExitTolerance;
    r.exit_cap += st.cntExitCap;
    r.exit_curvature += st.cntExitCurvature;
    r.exit_zero += st.cntExitZero;
    r.restarts += st.cntRestarts;
    r.true_rel_max = std::max(r.true_rel_max, st.trueRelMax);
    r.true_rel_last = std::max(r.true_rel_last, st.trueRelLast);
    r.true_rel_max_converged = std::max(r.true_rel_max_converged, st.trueRelMaxConverged);
    r.true_rel_excess_max = std::max(r.true_rel_excess_max, st.trueExcessMax);
    sum += st.trueRelSum;
  }
  r.true_rel_mean = r.solves ? sum / double(r.solves) : 0.0;
  r.persistent_aborts = persistentAborts_;
  return r;
}

void HipSolver::resetCounters() {
  requireBuilt();
  pollStates();
  for (size_t s = 0; s < model_.segments.size(); ++s) {
    dk::SegState& st = hostStates_[s];
    st.cntSubsteps = st.cntNewton = st.cntPcg = st.cntLs = st.cntWarmRejected = st.cntWarmAccepted = 0;
    st.cntSolves = st.cntExitTolerance = st.cntExitCap = st.cntExitCurvature = st.cntExitZero = st.cntRestarts = 0;
    st.trueRelMax = st.trueRelSum = st.trueRelLast = st.trueRelMaxConverged = st.trueExcessMax = 0.0;
  }
  check(hipMemcpyAsync(segState_.as<void>(), hostStates_, sizeof(dk::SegState) * model_.segments.size(),
                       hipMemcpyHostToDevice, stream_),
        "copy");
  check(hipStreamSynchronize(stream_), "sync");
}

drape_structure_sizes HipSolver::structureSizes() const {
  requireBuilt();
  drape_structure_sizes s{};
  s.n_particles = static_cast<uint32_t>(model_.x.size());
  s.n_stretch = static_cast<uint32_t>(model_.stretch.size());
  s.n_hinges = static_cast<uint32_t>(model_.hinges.size());
  s.n_seams = static_cast<uint32_t>(model_.seams.size());
  s.n_blocks = nBlocks_ - (nRows_ - static_cast<uint32_t>(model_.x.size()));  // padding rows hold one block each
  return s;
}

// The device structure re-indexed to global particles (identical to the internal one for a single scene, whose
// padding rows all follow the last particle).
void HipSolver::globalStructure(std::vector<uint32_t>& rowStart, std::vector<uint32_t>& colIndex,
                                std::vector<uint32_t>& diagSlot, std::vector<uint32_t>& internalToGlobalSlot) const {
  std::vector<uint32_t> rs(nRows_ + 1), ci(nBlocks_);
  fromDevice(rs.data(), rowStart_.as<void>(), sizeof(uint32_t) * (nRows_ + 1));
  fromDevice(ci.data(), colIndex_.as<void>(), sizeof(uint32_t) * nBlocks_);
  const size_t n = model_.x.size();
  std::vector<uint32_t> particleOfRow(nRows_, UINT32_MAX);
  for (uint32_t i = 0; i < n; ++i) particleOfRow[rowOf_[i]] = i;
  rowStart.assign(n + 1, 0);
  colIndex.clear();
  diagSlot.assign(n, 0);
  internalToGlobalSlot.assign(nBlocks_, UINT32_MAX);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t r = rowOf_[i];
    for (uint32_t s = rs[r]; s < rs[r + 1]; ++s) {
      const uint32_t c = particleOfRow[ci[s]];
      if (c == i) diagSlot[i] = static_cast<uint32_t>(colIndex.size());
      internalToGlobalSlot[s] = static_cast<uint32_t>(colIndex.size());
      colIndex.push_back(c);
    }
    rowStart[i + 1] = static_cast<uint32_t>(colIndex.size());
  }
}

void HipSolver::readStructure(uint32_t* rowStart, uint32_t* colIndex, uint32_t* diagSlot, uint32_t* stretchSlots,
                              uint32_t* hingeSlots, uint32_t* seamSlots) const {
  requireBuilt();
  check(hipStreamSynchronize(stream_), "sync");
  std::vector<uint32_t> rs, ci, ds, map;
  globalStructure(rs, ci, ds, map);
  if (rowStart) std::copy(rs.begin(), rs.end(), rowStart);
  if (colIndex) std::copy(ci.begin(), ci.end(), colIndex);
  if (diagSlot) std::copy(ds.begin(), ds.end(), diagSlot);
  auto slots = [&](const DeviceBuffer& b, size_t count, uint32_t* out) {
    if (!out || count == 0) return;
    std::vector<uint32_t> s(count);
    fromDevice(s.data(), b.as<void>(), sizeof(uint32_t) * count);
    for (size_t k = 0; k < count; ++k) out[k] = map[s[k]];
  };
  slots(stretchSlots_, 9 * model_.stretch.size(), stretchSlots);
  slots(hingeSlots_, 16 * model_.hinges.size(), hingeSlots);
  slots(seamSlots_, 4 * model_.seams.size(), seamSlots);
}

// Each row adds its own term, then its incidence entries in list order (kind, term, local vertex a), each term's
// block (a, c) going to slot(a, c): so a gradient's summands are [own, entries in order] and a slot's summands are
// [own (diagonal), the entries' (a, c) landing there, in entry order] — the reference's term-ordered gather lists
// when the device lists are sorted as the reference's are.
void HipSolver::gatherLists(std::vector<uint32_t>& gradStart, std::vector<uint32_t>& gradIdx,
                            std::vector<uint32_t>& blockStart, std::vector<uint32_t>& blockIdx) const {
  requireBuilt();
  check(hipStreamSynchronize(stream_), "sync");
  std::vector<uint32_t> rs, ci, ds, map;
  globalStructure(rs, ci, ds, map);
  const uint32_t n = static_cast<uint32_t>(model_.x.size());
  const uint32_t S = static_cast<uint32_t>(model_.stretch.size()), H = static_cast<uint32_t>(model_.hinges.size());
  const uint32_t K = static_cast<uint32_t>(model_.seams.size());
  std::vector<uint32_t> incStart(nRows_ + 1);
  fromDevice(incStart.data(), incStart_.as<void>(), sizeof(uint32_t) * (nRows_ + 1));
  std::vector<uint32_t> inc(incStart[nRows_]);
  if (!inc.empty()) fromDevice(inc.data(), inc_.as<void>(), sizeof(uint32_t) * inc.size());
  auto fetch = [&](const DeviceBuffer& b, size_t count) {
    std::vector<uint32_t> v(count);
    if (count) fromDevice(v.data(), b.as<void>(), sizeof(uint32_t) * count);
    return v;
  };
  const std::vector<uint32_t> sSlots = fetch(stretchSlots_, 9ull * S), hSlots = fetch(hingeSlots_, 16ull * H);
  const std::vector<uint32_t> kSlots = fetch(seamSlots_, 4ull * K);
  const uint32_t nBlocks = static_cast<uint32_t>(ci.size());
  std::vector<std::vector<uint32_t>> perSlot(nBlocks);
  gradStart.assign(1, 0);
  gradIdx.clear();
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t r = rowOf_[i];
    gradIdx.push_back(i);
    perSlot[ds[i]].push_back(i);
    for (uint32_t q = incStart[r]; q < incStart[r + 1]; ++q) {
      const uint32_t e = inc[q], kind = dk::incKind(e), t = dk::incTerm(e), a = dk::incLocal(e);
      uint32_t m, contrib, hessBase;
      const uint32_t* sl;
      if (kind == dk::kKindStretch) {
        m = 3, contrib = n + 3 * t + a, hessBase = n + 9 * t + 3 * a, sl = &sSlots[9 * size_t(t) + 3 * a];
      } else if (kind == dk::kKindHinge) {
        m = 4, contrib = n + 3 * S + 4 * t + a, hessBase = n + 9 * S + 16 * t + 4 * a, sl = &hSlots[16 * size_t(t) + 4 * a];
      } else {
        m = 2, contrib = n + 3 * S + 4 * H + 2 * t + a, hessBase = n + 9 * S + 16 * H + 4 * t + 2 * a,
        sl = &kSlots[4 * size_t(t) + 2 * a];
      }
      gradIdx.push_back(contrib);
      for (uint32_t c = 0; c < m; ++c) perSlot[map[sl[c]]].push_back(hessBase + c);
    }
    gradStart.push_back(static_cast<uint32_t>(gradIdx.size()));
  }
  blockStart.assign(1, 0);
  blockIdx.clear();
  for (const auto& l : perSlot) {
    blockIdx.insert(blockIdx.end(), l.begin(), l.end());
    blockStart.push_back(static_cast<uint32_t>(blockIdx.size()));
  }
}

void HipSolver::readTerms(uint32_t* stretchV, uint32_t* hingeV, double* hingeK, uint32_t* seamAB) const {
  requireBuilt();
  for (size_t k = 0; k < model_.stretch.size(); ++k)
    if (stretchV)
      for (int a = 0; a < 3; ++a) stretchV[3 * k + a] = model_.stretch[k].v[a];
  for (size_t k = 0; k < model_.hinges.size(); ++k) {
    if (hingeV)
      for (int a = 0; a < 4; ++a) hingeV[4 * k + a] = model_.hinges[k].v[a];
    if (hingeK) hingeK[k] = model_.hinges[k].K;
  }
  for (size_t k = 0; k < model_.seams.size(); ++k)
    if (seamAB) {
      seamAB[2 * k] = model_.seams[k].a;
      seamAB[2 * k + 1] = model_.seams[k].b;
    }
}

void HipSolver::debugAssemble(double h, const float* u, double* energy, float* grad, float* blocks) {
  requireBuilt();
  uploadSegParams(h);
  // Keep the contact force the next step's friction reads.
  std::vector<float> savedForce(nRows_);
  fromDevice(savedForce.data(), cforce_.as<void>(), sizeof(float) * nRows_);
  dk::launchSubstepBegin(view_, false, stream_);
  std::vector<dk::F3> us(nRows_ + 1, dk::F3{0, 0, 0});
  for (size_t i = 0; i < model_.x.size(); ++i) us[rowOf_[i]] = dk::F3{u[3 * i], u[3 * i + 1], u[3 * i + 2]};
  check(hipMemcpyAsync(u_.as<void>(), us.data(), sizeof(dk::F3) * (nRows_ + 1), hipMemcpyHostToDevice, stream_), "copy");
  check(hipMemsetAsync(cflags_.as<void>(), 0, sizeof(uint32_t) * nRows_, stream_), "memset");
  check(hipMemsetAsync(clambda_.as<void>(), 0, sizeof(float) * nRows_, stream_), "memset");
  dk::launchAssemble(view_, maxK_, false, stream_);
  dk::launchNewtonBegin(view_, stream_);
  pollStates();
  if (energy) *energy = hostStates_[0].e0;
  for (size_t s = 1; s < model_.segments.size() && energy; ++s) *energy += hostStates_[s].e0;
  if (grad) {
    std::vector<dk::F3> g(nRows_ + 1);
    fromDevice(g.data(), grad_.as<void>(), sizeof(dk::F3) * (nRows_ + 1));
    for (size_t i = 0; i < model_.x.size(); ++i) {
      grad[3 * i] = g[rowOf_[i]].x;
      grad[3 * i + 1] = g[rowOf_[i]].y;
      grad[3 * i + 2] = g[rowOf_[i]].z;
    }
  }
  if (blocks) {
    DeviceBuffer tmp;
    tmp.allocate(sizeof(float) * 9 * nBlocks_);
    dk::launchSellToSlots(view_, tmp.as<float>(), stream_);
    std::vector<float> internal(9 * size_t(nBlocks_));
    check(hipMemcpyAsync(internal.data(), tmp.as<void>(), sizeof(float) * internal.size(), hipMemcpyDeviceToHost, stream_),
          "copy");
    check(hipStreamSynchronize(stream_), "sync");
    std::vector<uint32_t> rs, ci, ds, map;
    globalStructure(rs, ci, ds, map);
    for (uint32_t s = 0; s < nBlocks_; ++s)
      if (map[s] != UINT32_MAX) std::copy(&internal[9 * size_t(s)], &internal[9 * size_t(s)] + 9, blocks + 9 * size_t(map[s]));
  }
  toDevice(cforce_.as<void>(), savedForce.data(), sizeof(float) * nRows_);
}

void HipSolver::debugSystem(float* blocks, float* grad, float* precond, uint32_t* contactFlags,
                            float* contactNormals) {
  requireBuilt();
  check(hipStreamSynchronize(stream_), "sync");
  const size_t n = model_.x.size();
  if (blocks) {
    DeviceBuffer tmp;
    tmp.allocate(sizeof(float) * 9 * nBlocks_);
    dk::launchSellToSlots(view_, tmp.as<float>(), stream_);
    std::vector<float> internal(9 * size_t(nBlocks_));
    check(hipMemcpyAsync(internal.data(), tmp.as<void>(), sizeof(float) * internal.size(), hipMemcpyDeviceToHost, stream_),
          "copy");
    check(hipStreamSynchronize(stream_), "sync");
    std::vector<uint32_t> rs, ci, ds, map;
    globalStructure(rs, ci, ds, map);
    for (uint32_t s = 0; s < nBlocks_; ++s)
      if (map[s] != UINT32_MAX) std::copy(&internal[9 * size_t(s)], &internal[9 * size_t(s)] + 9, blocks + 9 * size_t(map[s]));
  }
  if (grad) {
    std::vector<dk::F3> g(nRows_ + 1);
    fromDevice(g.data(), grad_.as<void>(), sizeof(dk::F3) * (nRows_ + 1));
    for (size_t i = 0; i < n; ++i) {
      grad[3 * i] = g[rowOf_[i]].x;
      grad[3 * i + 1] = g[rowOf_[i]].y;
      grad[3 * i + 2] = g[rowOf_[i]].z;
    }
  }
  if (precond) {  // stored as the upper triangle (00, 01, 02, 11, 12, 22), exported as the full block
    std::vector<float> m(dk::kPrecondFloats * size_t(nRows_ + 1));
    fromDevice(m.data(), precond_.as<void>(), sizeof(float) * m.size());
    static const int kFull[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (size_t i = 0; i < n; ++i)
      for (int c = 0; c < 9; ++c) precond[9 * i + c] = m[dk::kPrecondFloats * size_t(rowOf_[i]) + kFull[c]];
  }
  if (contactFlags) {
    std::vector<uint32_t> f(nRows_ + 1);
    fromDevice(f.data(), cflags_.as<void>(), sizeof(uint32_t) * (nRows_ + 1));
    for (size_t i = 0; i < n; ++i) contactFlags[i] = f[rowOf_[i]];
  }
  if (contactNormals) {
    std::vector<float> c(4 * size_t(nRows_ + 1));
    fromDevice(c.data(), cnormal_.as<void>(), sizeof(float) * c.size());
    for (size_t i = 0; i < n; ++i) std::copy(&c[4 * size_t(rowOf_[i])], &c[4 * size_t(rowOf_[i])] + 4, contactNormals + 4 * i);
  }
}

void HipSolver::setWarmStart(ws::Net* net) {
  requireBuilt();
  check(hipStreamSynchronize(stream_), "sync");
  net_ = net;
  if (!net_) {
    check(hipMemsetAsync(warm_.as<void>(), 0, warm_.bytes(), stream_), "memset");
    check(hipStreamSynchronize(stream_), "sync");
  }
}

void HipSolver::frameInputs(float* x, float* v, float* dvPrev, float* phiNormal) {
  requireBuilt();
  dk::launchWarmFeatures(view_, stream_);
  check(hipStreamSynchronize(stream_), "sync");
  const size_t n = model_.x.size();
  auto copyF3 = [&](const DeviceBuffer& b, float* out) {
    if (!out) return;
    std::vector<dk::F3> h(nRows_ + 1);
    fromDevice(h.data(), b.as<void>(), sizeof(dk::F3) * (nRows_ + 1));
    for (size_t i = 0; i < n; ++i) {
      const dk::F3 q = h[rowOf_[i]];
      out[3 * i] = q.x;
      out[3 * i + 1] = q.y;
      out[3 * i + 2] = q.z;
    }
  };
  copyF3(x_, x);
  copyF3(v_, v);
  copyF3(dvLast_, dvPrev);
  if (phiNormal) {
    std::vector<float4> h(nRows_ + 1);
    fromDevice(h.data(), phiN_.as<void>(), sizeof(float4) * (nRows_ + 1));
    for (size_t i = 0; i < n; ++i) {
      const float4 q = h[rowOf_[i]];
      phiNormal[4 * i] = q.x;
      phiNormal[4 * i + 1] = q.y;
      phiNormal[4 * i + 2] = q.z;
      phiNormal[4 * i + 3] = q.w;
    }
  }
}

void HipSolver::meshGraph(std::vector<uint32_t>& rowStart, std::vector<uint32_t>& nbr) const {
  requireBuilt();
  const size_t n = model_.x.size();
  std::vector<std::vector<uint32_t>> adj(n);
  for (const StretchTerm& t : model_.stretch)
    for (int a = 0; a < 3; ++a) {
      adj[t.v[a]].push_back(t.v[(a + 1) % 3]);
      adj[t.v[(a + 1) % 3]].push_back(t.v[a]);
    }
  rowStart.assign(n + 1, 0);
  nbr.clear();
  for (size_t i = 0; i < n; ++i) {
    std::sort(adj[i].begin(), adj[i].end());
    adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
    nbr.insert(nbr.end(), adj[i].begin(), adj[i].end());
    rowStart[i + 1] = static_cast<uint32_t>(nbr.size());
  }
}

void HipSolver::debugSpmv(const float* x, float* y) {
  requireBuilt();
  std::vector<dk::F3> xs(nRows_ + 1, dk::F3{0, 0, 0}), ys(nRows_ + 1);
  for (size_t i = 0; i < model_.x.size(); ++i) xs[rowOf_[i]] = dk::F3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  check(hipMemcpyAsync(scratchF3_.as<void>(), xs.data(), sizeof(dk::F3) * (nRows_ + 1), hipMemcpyHostToDevice, stream_),
        "copy");
  dk::launchSpmvPlain(view_, maxK_, scratchF3_.as<dk::F3>(), scratchF3b_.as<dk::F3>(), stream_);
  check(hipMemcpyAsync(ys.data(), scratchF3b_.as<void>(), sizeof(dk::F3) * (nRows_ + 1), hipMemcpyDeviceToHost, stream_),
        "copy");
  check(hipStreamSynchronize(stream_), "sync");
  for (size_t i = 0; i < model_.x.size(); ++i) {
    y[3 * i] = ys[rowOf_[i]].x;
    y[3 * i + 1] = ys[rowOf_[i]].y;
    y[3 * i + 2] = ys[rowOf_[i]].z;
  }
}

}  // namespace dh
