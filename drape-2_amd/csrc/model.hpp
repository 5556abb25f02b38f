// model.hpp — host-side scene model of the HIP solver: garments, particles and energy terms built from the
// C-ABI scene descriptors (the job of SimModel::build, libs/sim/src/sim_model.cpp:76-197), plus the row layout of
// a batched build (each scene one segment, padded to whole 256-row tiles).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "drape_hip.h"

namespace dh {

struct Vec3 {
  double x = 0, y = 0, z = 0;
};
struct Vec3f {
  float x = 0, y = 0, z = 0;
};

struct Fabric {
  double stretchWarp = 0, stretchWeft = 0, stretchBias = 0, strainLimitWarp = 0, strainLimitWeft = 0;
  double bendWarp = 0, bendWeft = 0, weight = 0, thickness = 0, friction = 0, damping = 0;
};

struct Garment {
  std::string id;
  uint32_t offset = 0, count = 0;  // global particle range
  uint32_t segment = 0;
  int32_t phase = DRAPE_PHASE_SETTLING;
  double assemblyTime = 0;
  uint32_t fabricOffset = 0, pieceCount = 0;
};

struct StretchTerm {
  uint32_t v[3];
  double dmInv[4];  // (0,0), (1,0), (0,1), (1,1)
  double area;
  uint32_t fabric;
};

struct HingeSide {
  uint32_t fabric;
  double edge[2];
  double grain[2];
};

struct HingeTerm {
  uint32_t v[4];
  double K = 0;
  HingeSide side[2];
  double geom = 0;  // |e|^2 / (A1 + A2)
};

struct SeamTerm {
  uint32_t a, b, garment;
  float startGap;
};

struct Field {
  double origin[3] = {0, 0, 0};
  double cell = 0.004;
  int32_t dims[3] = {0, 0, 0};
  double band = 0.02;
  std::vector<float> values;
  bool present() const { return !values.empty(); }
  double sample(const Vec3& p) const;  // CollisionField::sample (collision_field.cpp:8-27), FP64 on the host
};

struct Settings {
  double gravity[3] = {0, -9.81, 0};
  double collisionOffset = 0.003, closingSpeed = 0.25, extraDamping = 10.0, dampingScale = 5.0;
  double limitScale = 100.0, seamStiffness = 1e4, frictionSlip = 3e-4, cgTolerance = 1e-2;
  int32_t maxCg = 50;
};

struct Segment {
  uint32_t particleBegin = 0, particleCount = 0;  // global particles
  uint32_t rowBegin = 0, rowCount = 0;            // internal rows (rowCount a multiple of 256)
  uint32_t garmentBegin = 0, garmentCount = 0;
  Settings settings;
  Field body;
};

// Everything the solver needs on the host. Particle-indexed arrays use global (scene-order) indices; terms store
// global indices too and are mapped to internal rows on upload.
struct Model {
  std::vector<Segment> segments;
  std::vector<Garment> garments;
  std::vector<Fabric> fabrics;
  std::vector<Vec3f> x, v;
  std::vector<double> mass, vertexArea, contactForce;
  std::vector<double> rest;  // pattern-space rest position, 2 per particle
  std::vector<Vec3> fExt;
  std::vector<uint32_t> pFabric, pGarment, pSegment;
  std::vector<uint8_t> pinned;
  std::vector<std::pair<uint32_t, Vec3f>> pins;
  std::vector<StretchTerm> stretch;
  std::vector<HingeTerm> hinges;
  std::vector<SeamTerm> seams;

  static Model build(const drape_scene_desc* scenes, uint32_t nScenes);

  size_t garmentIndex(const std::string& id) const;  // std::out_of_range for unknown ids (sim_model.cpp:12-16)
  uint32_t internalRow(uint32_t particle) const {
    const Segment& s = segments[pSegment[particle]];
    return s.rowBegin + (particle - s.particleBegin);
  }
  uint32_t rowCount() const { return segments.empty() ? 0 : segments.back().rowBegin + segments.back().rowCount; }
  double hingeStiffness(const HingeTerm& h) const;  // sim_model.cpp:20-27
  void updateMass(uint32_t i) { mass[i] = std::max(vertexArea[i] * fabrics[pFabric[i]].weight, 1e-12); }
  double contactRadius(uint32_t i) const {
    return segments[pSegment[i]].settings.collisionOffset + 0.5 * fabrics[pFabric[i]].thickness;
  }
};

Fabric fabricFromDesc(const drape_fabric& f);
void bendingCoefficients(double bendWarp, double bendWeft, double& p, double& q);  // bending.cpp:7-10
double edgeBendStiffness(double p, double q, const double* edge, const double* grain);  // bending.cpp:12-18

}  // namespace dh
