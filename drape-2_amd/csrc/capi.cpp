// capi.cpp — the extern "C" boundary (include/drape_hip.h) over HipSolver and the standalone kernel entry points.
// Every entry point catches C++ exceptions and returns the status code of the reference exception type, with the
// message in a thread-local string (drape_last_error).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "drape_hip.h"
#include "bake.hpp"
#include "kernels.hpp"
#include "normals.hpp"
#include "solver.hpp"
#include "warmstart.hpp"

struct drape_solver {
  dh::HipSolver impl;
};

// A network with the device buffers its host-array entry points stage through.
struct drape_net {
  explicit drape_net(uint64_t seed) : impl(seed) {
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess)
      throw std::runtime_error("hip: stream creation failed");
  }
  ~drape_net() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
  ws::Net impl;
  hipStream_t stream = nullptr;
  dh::DeviceBuffer x, v, dv, pn, target, out;
  uint32_t cap = 0;
};

struct drape_graph {
  drape_graph(uint32_t n, const uint32_t* rs, const uint32_t* nbr, const double* rest2) : impl(n, rs, nbr, rest2) {}
  ws::Graph impl;
};

namespace {

thread_local std::string gError;

template <class F>
int guard(F&& f) {
  try {
    gError.clear();
    f();
    return DRAPE_OK;
  } catch (const std::invalid_argument& e) {
    gError = e.what();
    return DRAPE_ERR_INVALID_ARGUMENT;
  } catch (const std::out_of_range& e) {
    gError = e.what();
    return DRAPE_ERR_OUT_OF_RANGE;
  } catch (const std::logic_error& e) {
    gError = e.what();
    return DRAPE_ERR_LOGIC;
  } catch (const std::exception& e) {
    gError = e.what();
    return DRAPE_ERR_RUNTIME;
  } catch (...) {
    gError = "hip: unknown error";
    return DRAPE_ERR_RUNTIME;
  }
}

void need(const void* p, const char* what) {
  if (!p) throw std::invalid_argument(std::string("null ") + what);
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("hip: ") + what + ": " + hipGetErrorString(e));
}

// Scoped device allocation for the standalone entry points.
struct Dev {
  void* p = nullptr;
  explicit Dev(size_t bytes) { check(hipMalloc(&p, std::max<size_t>(bytes, 16)), "hipMalloc"); }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Stream {
  hipStream_t s = nullptr;
  Stream() { check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
  ~Stream() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

void requireDevice() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw std::runtime_error("hip: no GPU device available");
}

dk::Fabric32 fabric32(const drape_fabric& f) {
  dk::Fabric32 g{};
  g.stretchWarp = float(f.stretch_warp);
  g.stretchWeft = float(f.stretch_weft);
  g.stretchBias = float(f.stretch_bias);
  g.strainLimitWarp = float(f.strain_limit_warp);
  g.strainLimitWeft = float(f.strain_limit_weft);
  g.weight = float(f.weight);
  g.thickness = float(f.thickness);
  g.friction = float(f.friction);
  g.damping = float(f.damping);
  return g;
}

}  // namespace

extern "C" {

const char* drape_last_error(void) { return gError.c_str(); }

int drape_settings_default(drape_settings* out) {
  return guard([&] {
    need(out, "output");
    *out = drape_settings{};
    out->gravity[1] = -9.81;
    out->collision_offset = 0.003;
    out->assembly_closing_speed = 0.25;
    out->assembly_extra_damping = 10.0;
    out->damping_scale = 5.0;
    out->strain_limit_stiffness_scale = 100.0;
    out->seam_stiffness = 1e4;
    out->friction_slip_speed = 3e-4;
    out->max_cg_iterations = 50;
    out->cg_tolerance = 1e-2;
  });
}

int drape_hip_available(void) {
  try {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
      gError = "hip: no GPU device available";
      return 0;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      gError = "hip: device 0 is not gfx950";
      return 0;
    }
    return 1;
  } catch (...) {
    return 0;
  }
}

int drape_solver_create(drape_solver** out) {
  return guard([&] {
    need(out, "output");
    *out = new drape_solver();
  });
}

void drape_solver_destroy(drape_solver* s) { delete s; }

int drape_solver_build(drape_solver* s, const drape_scene_desc* scene) {
  return guard([&] {
    need(s, "solver");
    need(scene, "scene");
    s->impl.build(scene, 1);
  });
}

int drape_solver_build_batch(drape_solver* s, uint32_t n, const drape_scene_desc* scenes) {
  return guard([&] {
    need(s, "solver");
    if (n) need(scenes, "scenes");
    s->impl.build(scenes, n);
  });
}

int drape_solver_update_fabric(drape_solver* s, const char* garment, uint32_t n, const drape_fabric* f) {
  return guard([&] {
    need(s, "solver");
    std::vector<dh::Fabric> fabrics;
    for (uint32_t k = 0; k < n; ++k) fabrics.push_back(dh::fabricFromDesc(f[k]));
    s->impl.updateFabric(garment ? garment : "", fabrics);
  });
}

int drape_solver_set_phase(drape_solver* s, const char* garment, int32_t phase) {
  return guard([&] {
    need(s, "solver");
    if (phase != DRAPE_PHASE_ASSEMBLING && phase != DRAPE_PHASE_SETTLING) throw std::invalid_argument("unknown phase");
    s->impl.setPhase(garment ? garment : "", phase);
  });
}

int drape_solver_range(const drape_solver* s, const char* garment, uint32_t* offset, uint32_t* count) {
  return guard([&] {
    need(s, "solver");
    need(offset, "offset");
    need(count, "count");
    s->impl.range(garment ? garment : "", *offset, *count);
  });
}

int drape_solver_particle_count(const drape_solver* s, size_t* out) {
  return guard([&] {
    need(s, "solver");
    need(out, "output");
    *out = s->impl.particleCount();
  });
}

int drape_solver_segment_count(const drape_solver* s, uint32_t* out) {
  return guard([&] {
    need(s, "solver");
    need(out, "output");
    *out = s->impl.segmentCount();
  });
}

int drape_solver_segment_range(const drape_solver* s, uint32_t segment, uint32_t* offset, uint32_t* count) {
  return guard([&] {
    need(s, "solver");
    need(offset, "offset");
    need(count, "count");
    s->impl.segmentRange(segment, *offset, *count);
  });
}

int drape_solver_step(drape_solver* s, double dt, int32_t substeps, int32_t iterations) {
  return guard([&] {
    need(s, "solver");
    s->impl.step(dt, substeps, iterations);
  });
}

int drape_solver_read_positions(const drape_solver* s, float* out, size_t n) {
  return guard([&] {
    need(s, "solver");
    if (n) need(out, "output");
    s->impl.readPositions(out, n);
  });
}

int drape_solver_read_velocities(const drape_solver* s, float* out, size_t n) {
  return guard([&] {
    need(s, "solver");
    if (n) need(out, "output");
    s->impl.readVelocities(out, n);
  });
}

int drape_solver_stats(const drape_solver* s, drape_stats* out) {
  return guard([&] {
    need(s, "solver");
    need(out, "output");
    *out = s->impl.stats();
  });
}

int drape_solver_read_positions_normals(const drape_solver* s, float* positions, float* normals, size_t n) {
  return guard([&] {
    need(s, "solver");
    if (n) need(normals, "normals");
    s->impl.readPositionsNormals(positions, normals, n);
  });
}

int drape_solver_segment_stats(const drape_solver* s, uint32_t first, uint32_t count, drape_stats* out) {
  return guard([&] {
    need(s, "solver");
    if (count) need(out, "output");
    s->impl.segmentStats(first, count, out);
  });
}

int drape_solver_snapshot_size(const drape_solver* s, size_t* bytes) {
  return guard([&] {
    need(s, "solver");
    need(bytes, "output");
    *bytes = s->impl.snapshotSize();
  });
}

int drape_solver_snapshot_save(const drape_solver* s, void* out, size_t bytes) {
  return guard([&] {
    need(s, "solver");
    need(out, "output");
    s->impl.snapshotSave(out, bytes);
  });
}

int drape_solver_snapshot_restore(drape_solver* s, const void* in, size_t bytes) {
  return guard([&] {
    need(s, "solver");
    need(in, "snapshot");
    s->impl.snapshotRestore(in, bytes);
  });
}

int drape_solver_counters(const drape_solver* s, drape_counters* out) {
  return guard([&] {
    need(s, "solver");
    need(out, "output");
    *out = s->impl.counters();
  });
}

int drape_solver_set_refinement(drape_solver* s, int32_t restarts) {
  return guard([&] {
    need(s, "solver");
    if (restarts < 0) throw std::invalid_argument("refinement restarts must be >= 0");
    s->impl.setRefinement(restarts);
  });
}

int drape_solver_pcg_report(const drape_solver* s, drape_pcg_report* out) {
  return guard([&] {
    need(s, "solver");
    need(out, "out");
    *out = s->impl.pcgReport();
  });
}

int drape_solver_reset_counters(drape_solver* s) {
  return guard([&] {
    need(s, "solver");
    s->impl.resetCounters();
  });
}

int drape_solver_set_graph(drape_solver* s, int32_t enabled) {
  return guard([&] {
    need(s, "solver");
    s->impl.setGraph(enabled != 0);
  });
}

int drape_solver_set_pcg_variant(drape_solver* s, int32_t variant) {
  return guard([&] {
    need(s, "solver");
    s->impl.setPcgVariant(variant);
  });
}

int drape_solver_set_pcg_shape(drape_solver* s, int32_t waves, int32_t slots, int32_t halo) {
  return guard([&] {
    need(s, "solver");
    s->impl.setPcgShape(waves, slots, halo);
  });
}

int drape_solver_pcg_shape(const drape_solver* s, int32_t* waves, int32_t* slots, int32_t* halo) {
  return guard([&] {
    need(s, "solver");
    need(waves, "output");
    need(slots, "output");
    need(halo, "output");
    int w = 0, k = 0, h = 0;
    s->impl.pcgShape(w, k, h);
    *waves = w;
    *slots = k;
    *halo = h;
  });
}

int drape_solver_set_solve_spin_cap(drape_solver* s, uint32_t polls) {
  return guard([&] {
    need(s, "solver");
    s->impl.setSolveSpinCap(polls);
  });
}

int drape_solver_persistent_available(const drape_solver* s, int32_t* available) {
  return guard([&] {
    need(s, "solver");
    need(available, "output");
    *available = s->impl.persistentAvailable() ? 1 : 0;
  });
}

int drape_solver_set_spmv_sampling(drape_solver* s, int32_t every) {
  return guard([&] {
    need(s, "solver");
    s->impl.setSpmvSampling(every);
  });
}

int drape_solver_spmv_timing(const drape_solver* s, uint64_t* launches, double* mean_us, double* bytes_per_launch) {
  return guard([&] {
    need(s, "solver");
    uint64_t l = 0;
    double m = 0, b = 0;
    s->impl.spmvTiming(l, m, b);
    if (launches) *launches = l;
    if (mean_us) *mean_us = m;
    if (bytes_per_launch) *bytes_per_launch = b;
  });
}

int drape_solver_structure_sizes(const drape_solver* s, drape_structure_sizes* out) {
  return guard([&] {
    need(s, "solver");
    need(out, "output");
    *out = s->impl.structureSizes();
  });
}

int drape_solver_read_structure(const drape_solver* s, uint32_t* row_start, uint32_t* col_index, uint32_t* diag_slot,
                                uint32_t* stretch_slots, uint32_t* hinge_slots, uint32_t* seam_slots) {
  return guard([&] {
    need(s, "solver");
    s->impl.readStructure(row_start, col_index, diag_slot, stretch_slots, hinge_slots, seam_slots);
  });
}

int drape_solver_gather_sizes(const drape_solver* s, uint32_t* n_grad, uint32_t* n_hess) {
  return guard([&] {
    need(s, "solver");
    std::vector<uint32_t> gs, gi, bs, bi;
    s->impl.gatherLists(gs, gi, bs, bi);
    if (n_grad) *n_grad = static_cast<uint32_t>(gi.size());
    if (n_hess) *n_hess = static_cast<uint32_t>(bi.size());
  });
}

int drape_solver_read_gather(const drape_solver* s, uint32_t* grad_start, uint32_t* grad_idx, uint32_t* block_start,
                             uint32_t* block_idx) {
  return guard([&] {
    need(s, "solver");
    std::vector<uint32_t> gs, gi, bs, bi;
    s->impl.gatherLists(gs, gi, bs, bi);
    if (grad_start) std::copy(gs.begin(), gs.end(), grad_start);
    if (grad_idx) std::copy(gi.begin(), gi.end(), grad_idx);
    if (block_start) std::copy(bs.begin(), bs.end(), block_start);
    if (block_idx) std::copy(bi.begin(), bi.end(), block_idx);
  });
}

int drape_solver_read_terms(const drape_solver* s, uint32_t* stretch_v, uint32_t* hinge_v, double* hinge_k,
                            uint32_t* seam_ab) {
  return guard([&] {
    need(s, "solver");
    s->impl.readTerms(stretch_v, hinge_v, hinge_k, seam_ab);
  });
}

int drape_solver_debug_assemble(drape_solver* s, double h, const float* u, double* energy, float* grad,
                                float* blocks) {
  return guard([&] {
    need(s, "solver");
    need(u, "displacement");
    if (!(h > 0)) throw std::invalid_argument("h must be positive");
    s->impl.debugAssemble(h, u, energy, grad, blocks);
  });
}

int drape_solver_debug_spmv(drape_solver* s, const float* x, float* y) {
  return guard([&] {
    need(s, "solver");
    need(x, "input");
    need(y, "output");
    s->impl.debugSpmv(x, y);
  });
}

int drape_solver_debug_system(drape_solver* s, float* blocks, float* grad, float* precond, uint32_t* contact_flags,
                              float* contact_normals) {
  return guard([&] {
    need(s, "solver");
    s->impl.debugSystem(blocks, grad, precond, contact_flags, contact_normals);
  });
}

}  // extern "C"

namespace {
std::vector<bake::Component> bakeComponents(const drape_body_mesh* components, uint32_t n) {
  need(components, "components");
  std::vector<bake::Component> c(n);
  for (uint32_t i = 0; i < n; ++i) {
    need(components[i].vertices, "vertices");
    need(components[i].triangles, "triangles");
    for (size_t k = 0; k < 3 * size_t(components[i].n_triangles); ++k)
      if (components[i].triangles[k] >= components[i].n_vertices)
        throw std::out_of_range("bake: triangle vertex out of range");
    c[i] = bake::Component{components[i].vertices, components[i].n_vertices, components[i].triangles,
                           components[i].n_triangles};
  }
  return c;
}
bake::Options bakeOptions(const drape_bake_options* o) {
  bake::Options b;
  if (o) b = bake::Options{o->cell_size, o->padding, o->band};
  return b;
}
}  // namespace

extern "C" {

int drape_bake_field_grid(const drape_body_mesh* components, uint32_t n, const drape_bake_options* options,
                          drape_field_desc* out) {
  return guard([&] {
    need(out, "output");
    const auto c = bakeComponents(components, n);
    const bake::Options o = bakeOptions(options);
    const bake::Grid g = bake::gridOf(c.data(), n, o);
    *out = drape_field_desc{};
    for (int a = 0; a < 3; ++a) {
      out->origin[a] = g.origin[a];
      out->dims[a] = g.dims[a];
    }
    out->cell_size = o.cellSize;
    out->band = o.band;
    out->values = nullptr;
  });
}

int drape_bake_collision_field(const drape_body_mesh* components, uint32_t n, const drape_bake_options* options,
                               float* values) {
  return guard([&] {
    need(values, "values");
    const auto c = bakeComponents(components, n);
    Stream st;
    bake::bakeField(c.data(), n, bakeOptions(options), values, st.s);
  });
}

int drape_vertex_normals(const float* positions, uint32_t n, const uint32_t* triangles, uint32_t n_triangles,
                         const uint32_t* weld_pairs, uint32_t n_welds, float* out) {
  return guard([&] {
    if (n) {
      need(positions, "positions");
      need(out, "output");
    }
    if (n_triangles) need(triangles, "triangles");
    if (n_welds) need(weld_pairs, "weld pairs");
    Stream st;
    const nrm::Mesh mesh(n, triangles, n_triangles, weld_pairs, n_welds);
    Dev dp(sizeof(dk::F3) * std::max<uint32_t>(n, 1)), dn(sizeof(dk::F3) * std::max<uint32_t>(n, 1));
    if (n) check(hipMemcpyAsync(dp.p, positions, sizeof(float) * 3 * size_t(n), hipMemcpyHostToDevice, st.s), "copy");
    mesh.compute(dp.as<dk::F3>(), dn.as<dk::F3>(), st.s);
    if (n) check(hipMemcpyAsync(out, dn.p, sizeof(float) * 3 * size_t(n), hipMemcpyDeviceToHost, st.s), "copy");
    check(hipStreamSynchronize(st.s), "normals");
  });
}

int drape_vertex_normals_noalloc(const float* positions, uint32_t n, const uint32_t* triangles, uint32_t n_triangles,
                                 const uint32_t* weld_pairs, uint32_t n_welds, float* out, void* stream) {
  return guard([&] {
    if (n) {
      need(positions, "positions");
      need(out, "output");
    }
    if (n_triangles) need(triangles, "triangles");
    if (n_welds) need(weld_pairs, "weld pairs");
    requireDevice();
    nrm::vertexNormalsNoAlloc(positions, n, triangles, n_triangles, weld_pairs, n_welds, out,
                              static_cast<hipStream_t>(stream));
  });
}

int drape_bsr_spmv(uint32_t n, uint32_t n_blocks, const uint32_t* row_start, const uint32_t* col_index,
                   const float* blocks, const uint8_t* pinned, const float* x, float* y, int32_t reps,
                   double* mean_us) {
  return guard([&] {
    requireDevice();
    need(row_start, "row_start");
    if (n_blocks) {
      need(col_index, "col_index");
      need(blocks, "blocks");
    }
    need(x, "x");
    need(y, "y");
    if (row_start[n] != n_blocks) throw std::invalid_argument("row_start[n] must equal n_blocks");
    for (uint32_t i = 0; i < n; ++i)
      if (row_start[i] > row_start[i + 1]) throw std::invalid_argument("row_start must be non-decreasing");
    for (uint32_t k = 0; k < n_blocks; ++k)
      if (col_index[k] >= n) throw std::out_of_range("column index out of range");
    const uint32_t nRows = std::max<uint32_t>(dk::kTile, (n + dk::kTile - 1) / dk::kTile * dk::kTile);
    const uint32_t nSlices = nRows / dk::kSlice, nTiles = nRows / dk::kTile;
    std::vector<uint32_t> flags(nRows, dk::kRowPinned | dk::kRowPad), sellK(nSlices, 0), sellBase(nSlices + 1, 0);
    for (uint32_t i = 0; i < n; ++i) flags[i] = (pinned && pinned[i]) ? dk::kRowPinned : 0u;
    for (uint32_t t = 0; t < nSlices; ++t) {
      uint32_t K = 0;
      for (uint32_t l = 0; l < dk::kSlice; ++l) {
        const uint32_t r = t * dk::kSlice + l;
        if (r < n) K = std::max(K, row_start[r + 1] - row_start[r]);
      }
      sellK[t] = K;
      sellBase[t + 1] = sellBase[t] + K * 9u * dk::kSlice;
    }
    uint32_t maxK = 0;
    for (uint32_t t = 0; t < nSlices; ++t) maxK = std::max(maxK, sellK[t]);
    std::vector<uint32_t> sellCol(size_t(sellBase[nSlices] / 9) + dk::kSellColSlack, nRows);  // see kSellColSlack
    std::vector<float> sellVal(size_t(sellBase[nSlices]) + dk::kSellSlack, 0.0f);  // zero slack: see kSellSlack
    for (uint32_t r = 0; r < n; ++r) {
      const uint32_t t = r / dk::kSlice, l = r % dk::kSlice;
      for (uint32_t s = row_start[r]; s < row_start[r + 1]; ++s) {
        const uint32_t k = s - row_start[r];
        const uint32_t c = col_index[s];
        sellCol[sellBase[t] / 9 + k * dk::kSlice + l] = (pinned && pinned[c]) ? nRows : c;
        for (int q = 0; q < 9; ++q) sellVal[sellBase[t] + (k * dk::kSlice + l) * 9 + q] = blocks[9 * size_t(s) + q];
      }
    }
    Stream st;
    Dev dFlags(4 * nRows), dK(4 * nSlices), dBase(4 * (nSlices + 1)), dCol(4 * sellCol.size()),
        dVal(4 * sellVal.size()), dx(sizeof(dk::F3) * (nRows + 1)), dy(sizeof(dk::F3) * (nRows + 1));
    std::vector<dk::F3> xs(nRows + 1, dk::F3{0, 0, 0});
    for (uint32_t i = 0; i < n; ++i) xs[i] = dk::F3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    check(hipMemcpyAsync(dFlags.p, flags.data(), 4 * nRows, hipMemcpyHostToDevice, st.s), "copy");
    check(hipMemcpyAsync(dK.p, sellK.data(), 4 * nSlices, hipMemcpyHostToDevice, st.s), "copy");
    check(hipMemcpyAsync(dBase.p, sellBase.data(), 4 * (nSlices + 1), hipMemcpyHostToDevice, st.s), "copy");
    check(hipMemcpyAsync(dCol.p, sellCol.data(), 4 * sellCol.size(), hipMemcpyHostToDevice, st.s), "copy");
    check(hipMemcpyAsync(dVal.p, sellVal.data(), 4 * sellVal.size(), hipMemcpyHostToDevice, st.s), "copy");
    check(hipMemcpyAsync(dx.p, xs.data(), sizeof(dk::F3) * (nRows + 1), hipMemcpyHostToDevice, st.s), "copy");
    dk::View v{};
    v.nRows = nRows;
    v.nTiles = nTiles;
    v.nSlices = nSlices;
    v.rowFlags = dFlags.as<uint32_t>();
    v.sellK = dK.as<uint32_t>();
    v.sellBase = dBase.as<uint32_t>();
    v.sellCol = dCol.as<uint32_t>();
    v.sellVal = dVal.as<float>();
    dk::launchSpmvPlain(v, maxK, dx.as<dk::F3>(), dy.as<dk::F3>(), st.s);  // warm-up
    const int R = std::max(1, reps);
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    check(hipEventRecord(e0, st.s), "event");
    for (int r = 0; r < R; ++r) dk::launchSpmvPlain(v, maxK, dx.as<dk::F3>(), dy.as<dk::F3>(), st.s);
    check(hipEventRecord(e1, st.s), "event");
    check(hipEventSynchronize(e1), "event");
    float ms = 0;
    check(hipEventElapsedTime(&ms, e0, e1), "event");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (mean_us) *mean_us = 1000.0 * ms / R;
    std::vector<dk::F3> ys(nRows + 1);
    check(hipMemcpy(ys.data(), dy.p, sizeof(dk::F3) * (nRows + 1), hipMemcpyDeviceToHost), "copy");
    for (uint32_t i = 0; i < n; ++i) {
      y[3 * i] = ys[i].x;
      y[3 * i + 1] = ys[i].y;
      y[3 * i + 2] = ys[i].z;
    }
  });
}

int drape_eval_terms(int32_t kind, uint32_t n, const float* x, const float* params, const drape_fabric* fabrics,
                     float limit_scale, float* energy, float* grad, float* hess) {
  return guard([&] {
    requireDevice();
    if (kind < 0 || kind > 3) throw std::invalid_argument("kind must be 0..3");
    if (n == 0) return;
    need(x, "x");
    need(params, "params");
    need(energy, "energy");
    need(grad, "grad");
    need(hess, "hess");
    if (kind == 0) need(fabrics, "fabrics");
    const int N = kind == 0 ? 3 : kind == 1 ? 4 : kind == 2 ? 2 : 1;
    const int np = kind == 0 ? 5 : kind == 1 ? 2 : kind == 2 ? 2 : 5;
    Stream st;
    Dev dx(sizeof(float) * 3 * N * size_t(n)), dp(sizeof(float) * np * size_t(n)), df(sizeof(dk::Fabric32) * size_t(n)),
        de(sizeof(float) * n), dg(sizeof(float) * 3 * N * size_t(n)), dh(sizeof(float) * 9 * N * N * size_t(n));
    check(hipMemcpyAsync(dx.p, x, sizeof(float) * 3 * N * size_t(n), hipMemcpyHostToDevice, st.s), "copy");
    check(hipMemcpyAsync(dp.p, params, sizeof(float) * np * size_t(n), hipMemcpyHostToDevice, st.s), "copy");
    if (kind == 0) {
      std::vector<dk::Fabric32> f(n);
      for (uint32_t i = 0; i < n; ++i) f[i] = fabric32(fabrics[i]);
      check(hipMemcpyAsync(df.p, f.data(), sizeof(dk::Fabric32) * n, hipMemcpyHostToDevice, st.s), "copy");
    }
    dk::launchEvalTerms(kind, n, dx.as<float>(), dp.as<float>(), df.as<dk::Fabric32>(), limit_scale, de.as<float>(),
                        dg.as<float>(), dh.as<float>(), st.s);
    check(hipStreamSynchronize(st.s), "eval");
    check(hipMemcpy(energy, de.p, sizeof(float) * n, hipMemcpyDeviceToHost), "copy");
    check(hipMemcpy(grad, dg.p, sizeof(float) * 3 * N * size_t(n), hipMemcpyDeviceToHost), "copy");
    check(hipMemcpy(hess, dh.p, sizeof(float) * 9 * N * N * size_t(n), hipMemcpyDeviceToHost), "copy");
  });
}

int drape_field_query(const drape_field_desc* field, uint32_t n, const float* p, float* phi, float* normal) {
  return guard([&] {
    requireDevice();
    need(field, "field");
    if (n == 0) return;
    need(p, "points");
    need(phi, "phi");
    need(normal, "normal");
    dk::Field32 f{};
    size_t count = 0;
    if (field->values) {
      if (field->dims[0] < 2 || field->dims[1] < 2 || field->dims[2] < 2)
        throw std::invalid_argument("body field needs >= 2 samples per axis");
      f.ox = float(field->origin[0]);
      f.oy = float(field->origin[1]);
      f.oz = float(field->origin[2]);
      f.cell = float(field->cell_size);
      f.dx = uint32_t(field->dims[0]);
      f.dy = uint32_t(field->dims[1]);
      f.dz = uint32_t(field->dims[2]);
      f.present = 1u;
      count = size_t(f.dx) * f.dy * f.dz;
    }
    f.band = float(field->band);
    Stream st;
    Dev dv(sizeof(float) * std::max<size_t>(count, 1)), dp(sizeof(float) * 3 * size_t(n)), dphi(sizeof(float) * n),
        dn(sizeof(float) * 3 * size_t(n));
    if (count) check(hipMemcpyAsync(dv.p, field->values, sizeof(float) * count, hipMemcpyHostToDevice, st.s), "copy");
    check(hipMemcpyAsync(dp.p, p, sizeof(float) * 3 * size_t(n), hipMemcpyHostToDevice, st.s), "copy");
    dk::launchFieldQuery(dv.as<float>(), f, n, dp.as<float>(), dphi.as<float>(), dn.as<float>(), st.s);
    check(hipStreamSynchronize(st.s), "query");
    check(hipMemcpy(phi, dphi.p, sizeof(float) * n, hipMemcpyDeviceToHost), "copy");
    check(hipMemcpy(normal, dn.p, sizeof(float) * 3 * size_t(n), hipMemcpyDeviceToHost), "copy");
  });
}

int drape_net_create(drape_net** out, uint64_t seed) {
  return guard([&] {
    need(out, "output");
    requireDevice();
    *out = new drape_net(seed);
  });
}

void drape_net_destroy(drape_net* net) { delete net; }

int drape_net_param_count(const drape_net* net, size_t* out) {
  return guard([&] {
    need(net, "network");
    need(out, "output");
    *out = net->impl.paramCount();
  });
}

int drape_net_get_params(const drape_net* net, float* out, size_t n) {
  return guard([&] {
    need(net, "network");
    need(out, "output");
    if (n != net->impl.paramCount()) throw std::invalid_argument("parameter count mismatch");
    net->impl.getParams(out);
  });
}

int drape_net_set_params(drape_net* net, const float* in, size_t n) {
  return guard([&] {
    need(net, "network");
    need(in, "parameters");
    if (n != net->impl.paramCount()) throw std::invalid_argument("parameter count mismatch");
    net->impl.setParams(in);
  });
}

int drape_net_reset_optimizer(drape_net* net) {
  return guard([&] {
    need(net, "network");
    net->impl.resetOptimizer();
  });
}

int drape_graph_create(drape_graph** out, uint32_t n, const uint32_t* row_start, const uint32_t* nbr,
                       const double* rest2) {
  return guard([&] {
    need(out, "output");
    need(row_start, "row_start");
    need(rest2, "rest2");
    if (row_start[n] && !nbr) throw std::invalid_argument("null nbr");
    for (uint32_t i = 0; i < n; ++i)
      if (row_start[i + 1] < row_start[i]) throw std::invalid_argument("row_start must be non-decreasing");
    requireDevice();
    *out = new drape_graph(n, row_start, nbr, rest2);
  });
}

void drape_graph_destroy(drape_graph* g) { delete g; }

namespace {

// Stage a frame of host arrays (n vertices) into the network's device buffers (graph rows, zero padded).
ws::FrameView stageFrame(drape_net* net, const drape_graph* g, const float* x, const float* v, const float* dvPrev,
                         const float* phiNormal, const float* target) {
  const uint32_t rows = g->impl.rows(), n = g->impl.particles();
  if (rows > net->cap) {
    for (dh::DeviceBuffer* b : {&net->x, &net->v, &net->dv, &net->target, &net->out}) b->allocate(sizeof(dk::F3) * rows);
    net->pn.allocate(sizeof(float4) * rows);
    net->cap = rows;
  }
  auto put = [&](dh::DeviceBuffer& b, const float* src, size_t floatsPerRow) {
    check(hipMemsetAsync(b.as<void>(), 0, sizeof(float) * floatsPerRow * rows, net->stream), "memset");
    if (src)
      check(hipMemcpyAsync(b.as<void>(), src, sizeof(float) * floatsPerRow * n, hipMemcpyHostToDevice, net->stream),
            "copy");
  };
  put(net->x, x, 3);
  put(net->v, v, 3);
  put(net->dv, dvPrev, 3);
  put(net->pn, phiNormal, 4);
  put(net->target, target, 3);
  return ws::FrameView{net->x.as<dk::F3>(), net->v.as<dk::F3>(), net->dv.as<dk::F3>(), net->pn.as<float4>()};
}

}  // namespace

int drape_net_predict(drape_net* net, const drape_graph* g, const float* x, const float* v, const float* dv_prev,
                      const float* phi_normal, float* dv_out) {
  return guard([&] {
    need(net, "network");
    need(g, "graph");
    need(x, "x");
    need(v, "v");
    need(dv_prev, "dv_prev");
    need(phi_normal, "phi_normal");
    need(dv_out, "output");
    const ws::FrameView f = stageFrame(net, g, x, v, dv_prev, phi_normal, nullptr);
    net->impl.predict(g->impl.view(), f, net->out.as<dk::F3>(), net->stream);
    check(hipMemcpyAsync(dv_out, net->out.as<void>(), sizeof(float) * 3 * g->impl.particles(), hipMemcpyDeviceToHost,
                         net->stream),
          "copy");
    check(hipStreamSynchronize(net->stream), "sync");
  });
}

int drape_net_train_step(drape_net* net, const drape_graph* g, const float* x, const float* v, const float* dv_prev,
                         const float* phi_normal, const float* target, float lr, double* loss, float* grad) {
  return guard([&] {
    need(net, "network");
    need(g, "graph");
    need(x, "x");
    need(v, "v");
    need(dv_prev, "dv_prev");
    need(phi_normal, "phi_normal");
    need(target, "target");
    if (!(lr >= 0.0f)) throw std::invalid_argument("learning rate must be >= 0");
    const ws::FrameView f = stageFrame(net, g, x, v, dv_prev, phi_normal, target);
    const double l = net->impl.lossAndGradient(g->impl.view(), f, net->target.as<dk::F3>(), lr, net->stream);
    if (loss) *loss = l;
    if (grad) net->impl.getGradient(grad);
  });
}

int drape_solver_set_warm_start(drape_solver* s, drape_net* net) {
  return guard([&] {
    need(s, "solver");
    s->impl.setWarmStart(net ? &net->impl : nullptr);
  });
}

int drape_solver_frame_inputs(drape_solver* s, float* x, float* v, float* dv_prev, float* phi_normal) {
  return guard([&] {
    need(s, "solver");
    s->impl.frameInputs(x, v, dv_prev, phi_normal);
  });
}

int drape_solver_mesh_graph_size(const drape_solver* s, size_t* n_entries) {
  return guard([&] {
    need(s, "solver");
    need(n_entries, "output");
    std::vector<uint32_t> rs, nb;
    s->impl.meshGraph(rs, nb);
    *n_entries = nb.size();
  });
}

int drape_solver_mesh_graph(const drape_solver* s, uint32_t* row_start, uint32_t* nbr) {
  return guard([&] {
    need(s, "solver");
    std::vector<uint32_t> rs, nb;
    s->impl.meshGraph(rs, nb);
    if (row_start) std::copy(rs.begin(), rs.end(), row_start);
    if (nbr) std::copy(nb.begin(), nb.end(), nbr);
  });
}

}  // extern "C"
