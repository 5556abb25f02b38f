// hip_solver.hpp — header-only C++ wrapper of the HIP C-ABI (drape_hip.h) with the reference's Solver surface.
//
// It mirrors drape::sim::Solver (libs/sim/include/drape/sim/solver.hpp:85-98) method for method, over plain
// arrays instead of Eigen types, and turns the C status codes back into the reference's exceptions:
// std::invalid_argument, std::out_of_range, std::runtime_error (device failures, message starting "hip:") and
// std::logic_error. A reference build derives from its own Solver and forwards to this class.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "drape_hip.h"

namespace drape::sim::hip {

inline void check(int code) {
  if (code == DRAPE_OK) return;
  const std::string msg = drape_last_error();
  switch (code) {
    case DRAPE_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case DRAPE_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
    case DRAPE_ERR_LOGIC: throw std::logic_error(msg);
    default: throw std::runtime_error(msg);
  }
}

// isAvailable(SolverBackend::Hip) (solver_factory.hpp:17-18).
inline bool isAvailable() { return drape_hip_available() != 0; }

class HipSolver {
 public:
  HipSolver() { check(drape_solver_create(&s_)); }
  ~HipSolver() { drape_solver_destroy(s_); }
  HipSolver(const HipSolver&) = delete;
  HipSolver& operator=(const HipSolver&) = delete;

  void build(const drape_scene_desc& scene) { check(drape_solver_build(s_, &scene)); }
  void buildBatch(const std::vector<drape_scene_desc>& scenes) {
    check(drape_solver_build_batch(s_, static_cast<uint32_t>(scenes.size()), scenes.data()));
  }
  void updateFabric(const std::string& garment, const std::vector<drape_fabric>& pieceFabric) {
    check(drape_solver_update_fabric(s_, garment.c_str(), static_cast<uint32_t>(pieceFabric.size()),
                                     pieceFabric.data()));
  }
  void setPhase(const std::string& garment, int32_t phase) { check(drape_solver_set_phase(s_, garment.c_str(), phase)); }
  void step(double dt, int32_t substeps, int32_t iterations) { check(drape_solver_step(s_, dt, substeps, iterations)); }
  void readPositions(float* out, std::size_t n) const { check(drape_solver_read_positions(s_, out, n)); }
  drape_stats stats() const {
    drape_stats st{};
    check(drape_solver_stats(s_, &st));
    return st;
  }
  std::vector<drape_stats> segmentStats(uint32_t first, uint32_t count) const {
    std::vector<drape_stats> st(count);
    check(drape_solver_segment_stats(s_, first, count, st.data()));
    return st;
  }
  std::vector<unsigned char> snapshot() const {
    std::size_t bytes = 0;
    check(drape_solver_snapshot_size(s_, &bytes));
    std::vector<unsigned char> blob(bytes);
    check(drape_solver_snapshot_save(s_, blob.data(), bytes));
    return blob;
  }
  void restore(const std::vector<unsigned char>& blob) {
    check(drape_solver_snapshot_restore(s_, blob.data(), blob.size()));
  }
  void range(const std::string& garment, uint32_t& offset, uint32_t& count) const {
    check(drape_solver_range(s_, garment.c_str(), &offset, &count));
  }
  std::size_t particleCount() const {
    std::size_t n = 0;
    check(drape_solver_particle_count(s_, &n));
    return n;
  }
  drape_counters counters() const {
    drape_counters c{};
    check(drape_solver_counters(s_, &c));
    return c;
  }

 private:
  drape_solver* s_ = nullptr;
};

}  // namespace drape::sim::hip
