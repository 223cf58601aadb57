// ParticleSystem.h -- particle state and material of the MPM substep (host side, no Eigen).
// Mirrors AnisotropicElastoplasticity/ParticleSystem.h:13-79. State is FP32 and planar (structure of
// arrays): positions[3n] = {x0[n], x1[n], x2[n]}; affineMomenta_1..3 of the reference are the rows
// of B, kept here as one 9-plane row-major array; plastic holds J_P = det F_P for snow (the only use
// of plasticDeformationGradients, HybridSolver.cpp:283) and plasticAmount q for sand.
#pragma once
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "../mpm_c_abi.h"

namespace mpm {

class ParticleSystem {
 public:
  std::vector<float> masses;       // [n]
  std::vector<float> volumes;      // [n]
  std::vector<float> plastic;      // [n] J_P (snow) or plasticAmount (sand)
  std::vector<float> frictionAngles;  // [n] per-particle friction angle in degrees (optional)
  std::vector<float> velocities;   // [3n]
  std::vector<float> positions;    // [3n]
  std::vector<float> affineMomenta;  // [9n] B row-major
  std::vector<float> elasticDeformationGradients;  // [9n] F_E row-major

  int material = MPM_SAND;
  double youngsModulus = 1.4e5;
  double poissonRatio = 0.2;
  double criticalCompression = 2.5e-2;
  double criticalStretch = 7.5e-3;
  double friction = 0.0;  // kept for the reference's field; unused by its solver (ParticleSystem.h:41)

  ParticleSystem() = default;

  // Mirrors ParticleSystem::ParticleSystem (ParticleSystem.cpp:81-117); F_E defaults to I, B to 0.
  ParticleSystem(const std::vector<float>& velocities_, const std::vector<float>& positions_,
                 const std::vector<float>& masses_, const std::vector<float>& volumes_, int material_,
                 double E, double nu, double thetaC = 2.5e-2, double thetaS = 7.5e-3)
      : masses(masses_), volumes(volumes_), velocities(velocities_), positions(positions_), material(material_),
        youngsModulus(E), poissonRatio(nu), criticalCompression(thetaC), criticalStretch(thetaS) {
    const size_t n = masses.size();
    if (positions.size() != 3 * n || velocities.size() != 3 * n || volumes.size() != n)
      throw std::invalid_argument("ParticleSystem: inconsistent array sizes");
    affineMomenta.assign(9 * n, 0.0f);
    elasticDeformationGradients.assign(9 * n, 0.0f);
    for (size_t p = 0; p < n; ++p)
      for (int a = 0; a < 3; ++a) elasticDeformationGradients[(4 * a) * n + p] = 1.0f;
    plastic.assign(n, material == MPM_SNOW ? 1.0f : 0.0f);
  }

  size_t size() const { return masses.size(); }

  // Seeded replacement for the reference factories (ParticleSystem.cpp:119-401): a box of
  // cells[0] x cells[1] x cells[2] lattice cells with ppa^3 jittered particles each, density rho,
  // initial velocity v0.
  static ParticleSystem Block(uint64_t seed, const double lo[3], const double hi[3], const int32_t cells[3], int ppa,
                              double rho, const float v0[3], int material, double E, double nu) {
    const int64_t n = (int64_t)cells[0] * cells[1] * cells[2] * ppa * ppa * ppa;
    std::vector<float> x(3 * n), v(3 * n), m(n), vol(n);
    if (mpm_generate_block(seed, lo, hi, cells, ppa, n, x.data()) != MPM_OK)
      throw std::runtime_error("mpm_generate_block failed");
    double V = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]) / double(n);
    for (int64_t p = 0; p < n; ++p) {
      m[p] = float(rho * V);
      vol[p] = float(V);
      for (int a = 0; a < 3; ++a) v[a * n + p] = v0[a];
    }
    return ParticleSystem(v, x, m, vol, material, E, nu);
  }

  // Seeded equivalents of the reference's shape factories (ParticleSystem.cpp:119-401), drawn by
  // mpm_generate_shape. Masses are the reference's totals divided by n (its constants kept,
  // e.g. 100 * 3.14 r^3 for SnowBall); its volumes are placeholders for the first P2G's density
  // estimate (HybridSolver::solve with estimateVolumes), so here V = shape volume / n until then.
  static ParticleSystem SnowBall(const double center[3], double radius, int64_t n, uint64_t seed = 1) {
    const double a[4] = {center[0], center[1], center[2], radius};
    return Shape(seed, MPM_SHAPE_BALL, a, n, 100.0 * 3.14 * radius * radius * radius,
                 4.0 / 3.0 * M_PI * radius * radius * radius, -1, MPM_SNOW, 1.4e5, 0.2);
  }
  static ParticleSystem SandBall(const double center[3], double radius, int64_t n, uint64_t seed = 1) {
    const double a[4] = {center[0], center[1], center[2], radius};
    return Shape(seed, MPM_SHAPE_BALL, a, n, 1300.0 * 3.14 * radius * radius * radius,
                 4.0 / 3.0 * M_PI * radius * radius * radius, -1, MPM_SAND, 3.537e5, 0.3);
  }
  static ParticleSystem SandBlock(const double bmin[3], const double bmax[3], double holeRadius, int64_t n,
                                  uint64_t seed = 1) {
    const double a[7] = {bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2], holeRadius};
    const double vol = (bmax[0] - bmin[0]) * (bmax[1] - bmin[1]) * (bmax[2] - bmin[2]) -
                       0.25 * M_PI * holeRadius * holeRadius * holeRadius;
    return Shape(seed, MPM_SHAPE_BLOCK, a, n, 1300.0 * vol, vol, -1, MPM_SAND, 3.537e5, 0.3);
  }
  // Height along `axis` (2 = z, the reference's); initial velocity -1 along the axis.
  static ParticleSystem SandCylinder(const double baseCenter[3], double radius, double height, int64_t n,
                                     uint64_t seed = 1, int axis = 2) {
    const double a[6] = {baseCenter[0], baseCenter[1], baseCenter[2], radius, height, double(axis)};
    const double vol = M_PI * radius * radius * height;
    return Shape(seed, MPM_SHAPE_CYLINDER, a, n, 1300.0 * vol, vol, axis, MPM_SAND, 3.537e5, 0.3);
  }

 private:
  static ParticleSystem Shape(uint64_t seed, int shape, const double* a, int64_t n, double totalMass,
                              double totalVolume, int velocityAxis, int material, double E, double nu) {
    if (n < 1) throw std::invalid_argument("ParticleSystem: n must be positive");
    std::vector<float> x(3 * n), v(3 * n, 0.0f), m(n, float(totalMass / double(n))), vol(n, float(totalVolume / double(n)));
    if (mpm_generate_shape(seed, shape, a, n, x.data()) != MPM_OK) throw std::runtime_error("mpm_generate_shape failed");
    if (velocityAxis >= 0)
      for (int64_t p = 0; p < n; ++p) v[velocityAxis * n + p] = -1.0f;
    return ParticleSystem(v, x, m, vol, material, E, nu);
  }
};

}  // namespace mpm
