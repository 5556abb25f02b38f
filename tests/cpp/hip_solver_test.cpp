// Compiled C++ test of the drop-in boundary: the reference-shaped wrapper drape::sim::hip::HipSolver
// (include/drape/sim/hip_solver.hpp) over libdrape_hip.so, as a reference build would call it through its Solver
// interface (libs/sim/include/drape/sim/solver.hpp:85-98; the backend-parameterised suite, backends.hpp:14-47).
//
//   hip_solver_test             build, step, readPositions, stats, counters, snapshot -> restore replay, and the
//                               reference's exception types for misuse (logic_error, out_of_range,
//                               invalid_argument)
//   hip_solver_test no-device   run with the GPU hidden: creating a solver raises std::runtime_error naming "hip"
//                               (solver_factory.cpp:57: a backend that is not available throws on construction)
// Prints one line per check and "boundary test passed"; exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "drape/sim/hip_solver.hpp"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  std::printf("%s: %s\n", ok ? "ok" : "FAILED", what);
  std::fflush(stdout);
  if (!ok) ++failures;
}

template <class E, class F>
void expectThrows(F&& f, const char* what, const char* mustContain = nullptr) {
  bool ok = false;
  std::string msg;
  try {
    f();
  } catch (const E& e) {
    msg = e.what();
    ok = mustContain == nullptr || msg.find(mustContain) != std::string::npos;
  } catch (const std::exception& e) {
    msg = std::string("other exception: ") + e.what();
  }
  std::printf("%s: %s [%s]\n", ok ? "ok" : "FAILED", what, msg.c_str());
  std::fflush(stdout);
  if (!ok) ++failures;
}

// A 16 x 16 sheet at 1 cm in the XZ plane at y = 0.3 m, pinned at two corners, over a sphere of radius 0.15 m
// sampled on a 4 mm grid (the reference's sphere scene, sphere_scene.hpp), testJersey fabric.
struct Scene {
  static constexpr uint32_t N = 16;
  std::vector<double> rest, area, grain{1.0, 0.0}, pos, pinPos, fieldOrigin;
  std::vector<uint32_t> vpiece, tris, tpiece, pins;
  std::vector<float> field;
  drape_fabric fabric{};
  drape_garment_desc garment{};
  drape_field_desc body{};
  drape_scene_desc desc{};

  Scene() {
    const double h = 0.01;
    for (uint32_t j = 0; j < N; ++j)
      for (uint32_t i = 0; i < N; ++i) {
        rest.push_back(i * h);
        rest.push_back(j * h);
        pos.push_back(-0.075 + i * h);
        pos.push_back(0.3);
        pos.push_back(-0.075 + j * h);
        vpiece.push_back(0);
        area.push_back(0.0);
      }
    for (uint32_t j = 0; j + 1 < N; ++j)
      for (uint32_t i = 0; i + 1 < N; ++i) {
        const uint32_t a = j * N + i, b = a + 1, c = a + N + 1, d = a + N;
        const uint32_t t[6] = {a, b, c, a, c, d};
        for (uint32_t k : t) tris.push_back(k);
        tpiece.push_back(0);
        tpiece.push_back(0);
        for (uint32_t k : t) area[k] += 0.5 * h * h / 3.0;
      }
    fabric = drape_fabric{400, 200, 100, 0.5, 0.8, 2e-6, 1.5e-6, 0.160, 0.0006, 0.3, 0.3};  // testJersey
    garment.id = "sheet";
    garment.mesh = drape_mesh_desc{N * N,  rest.data(), vpiece.data(), uint32_t(tpiece.size()), tris.data(),
                                   tpiece.data(), area.data(), nullptr, 1, grain.data(), 0, nullptr};
    garment.initial_positions = pos.data();
    garment.n_piece_fabrics = 1;
    garment.piece_fabric = &fabric;
    pins = {0, N - 1};
    for (uint32_t p : pins)
      for (int k = 0; k < 3; ++k) pinPos.push_back(pos[3 * p + k]);
    const double cell = 0.004, lo[3] = {-0.2, -0.2, -0.2};
    const int dims[3] = {101, 101, 101};
    for (int k = 0; k < dims[2]; ++k)
      for (int j = 0; j < dims[1]; ++j)
        for (int i = 0; i < dims[0]; ++i) {
          const double x = lo[0] + i * cell, y = lo[1] + j * cell, z = lo[2] + k * cell;
          field.push_back(static_cast<float>(std::sqrt(x * x + y * y + z * z) - 0.15));
        }
    body = drape_field_desc{{lo[0], lo[1], lo[2]}, cell, {dims[0], dims[1], dims[2]}, 0, 0.02, field.data()};
    desc.n_garments = 1;
    desc.garments = &garment;
    desc.body = &body;
    desc.n_pins = uint32_t(pins.size());
    desc.pin_particles = pins.data();
    desc.pin_positions = pinPos.data();
    drape_settings_default(&desc.settings);
  }
};

std::vector<float> positions(const drape::sim::hip::HipSolver& s) {
  std::vector<float> x(3 * s.particleCount());
  s.readPositions(x.data(), s.particleCount());
  return x;
}

}  // namespace

int main(int argc, char** argv) {
  using drape::sim::hip::HipSolver;
  if (argc > 1 && std::strcmp(argv[1], "no-device") == 0) {
    expect(!drape::sim::hip::isAvailable(), "isAvailable() is false without a device");
    expectThrows<std::runtime_error>([] { HipSolver s; }, "construction without a device", "hip");
    std::printf(failures ? "boundary test FAILED\n" : "boundary test passed\n");
    return failures ? 1 : 0;
  }
  expect(drape::sim::hip::isAvailable(), "isAvailable() on the GPU");
  Scene sc;
  HipSolver s;
  expectThrows<std::logic_error>([&] { s.step(1.0 / 60.0, 1, 2); }, "step before build is a logic_error", "not built");
  s.build(sc.desc);
  expect(s.particleCount() == Scene::N * Scene::N, "particleCount() after build");
  uint32_t off = 99, cnt = 0;
  s.range("sheet", off, cnt);
  expect(off == 0 && cnt == Scene::N * Scene::N, "range(garment)");
  expectThrows<std::out_of_range>([&] { s.range("zzz", off, cnt); }, "unknown garment is out_of_range", "zzz");
  expectThrows<std::invalid_argument>([&] {
    std::vector<float> x(3 * (s.particleCount() + 1));
    s.readPositions(x.data(), s.particleCount() + 1);
  }, "readPositions with a wrong span size is invalid_argument");
  const std::vector<float> x0 = positions(s);
  for (int f = 0; f < 5; ++f) s.step(1.0 / 60.0, 1, 2);
  const std::vector<unsigned char> snap = s.snapshot();
  for (int f = 0; f < 10; ++f) s.step(1.0 / 60.0, 1, 2);
  const std::vector<float> a = positions(s);
  const drape_stats st = s.stats();
  expect(st.has_non_finite == 0, "stats: no non-finite values");
  expect(st.kinetic_energy > 0.0 && st.rms_speed > 0.0, "stats: the falling sheet moves");
  bool pinned = true;
  for (uint32_t p : sc.pins)
    for (int k = 0; k < 3; ++k) pinned = pinned && a[3 * p + k] == static_cast<float>(sc.pos[3 * p + k]);
  expect(pinned, "pins hold their positions");
  const uint32_t mid = (Scene::N / 2) * Scene::N + Scene::N / 2;
  expect(a[3 * mid + 1] < x0[3 * mid + 1] - 0.01f, "the sheet's middle fell under gravity");
  const drape_counters c = s.counters();
  expect(c.substeps == 15 && c.newton_iterations >= 15 && c.pcg_iterations > 0, "counters after 15 frames");
  s.restore(snap);
  for (int f = 0; f < 10; ++f) s.step(1.0 / 60.0, 1, 2);
  const std::vector<float> b = positions(s);
  expect(std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0, "snapshot -> restore replays bit for bit");
  expectThrows<std::invalid_argument>([&] {
    std::vector<unsigned char> other(snap.begin(), snap.end() - 8);
    s.restore(other);
  }, "restore of a snapshot of another size is invalid_argument");
  std::printf(failures ? "boundary test FAILED\n" : "boundary test passed\n");
  return failures ? 1 : 0;
}
