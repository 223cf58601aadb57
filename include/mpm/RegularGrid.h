// RegularGrid.h -- Eulerian grid of the MPM substep (host side, no Eigen).
// Mirrors AnisotropicElastoplasticity/RegularGrid.h:15-61: node (i,j,k) sits at min + (i,j,k)*h with
// h = (max-min)/res (RegularGrid.cpp:144-146, 147-159) and linear index k*r0*r1 + j*r0 + i (:164-168).
// Grid arrays live on the device; masses/velocities here are host copies filled by
// HybridSolver::syncToHost.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <tuple>
#include <vector>

namespace mpm {

class RegularGrid {
 public:
  RegularGrid(const std::array<double, 3>& minBound, const std::array<double, 3>& maxBound,
              const std::array<int, 3>& resolution)
      : minBound_(minBound), maxBound_(maxBound), resolution_(resolution) {
    for (int a = 0; a < 3; ++a) {
      // the reference checks maxBound[2] < maxBound[2] (RegularGrid.cpp:129-131), i.e. never the z axis
      if (!(maxBound[a] > minBound[a]) || resolution[a] < 1)
        throw std::invalid_argument("RegularGrid: maxBound must be bigger than minBound and resolution positive");
      h_[a] = (maxBound_[a] - minBound_[a]) / resolution_[a];
    }
  }

  int toIndex(int i, int j, int k) const { return k * resolution_[0] * resolution_[1] + j * resolution_[0] + i; }
  std::tuple<int, int, int> toCoordinate(int index) const {
    int k = index / (resolution_[0] * resolution_[1]);
    int j = (index % (resolution_[0] * resolution_[1])) / resolution_[0];
    int i = index - k * resolution_[0] * resolution_[1] - j * resolution_[0];
    return std::make_tuple(i, j, k);
  }
  int gridNumber() const { return resolution_[0] * resolution_[1] * resolution_[2]; }
  double gridVolume() const { return h_[0] * h_[1] * h_[2]; }
  const std::array<double, 3>& minBound() const { return minBound_; }
  const std::array<double, 3>& maxBound() const { return maxBound_; }
  const std::array<double, 3>& h() const { return h_; }
  const std::array<int, 3>& resolution() const { return resolution_; }

  // host copies (planar): masses[n], velocities[3n]; valid after HybridSolver::syncToHost
  std::vector<float> masses;
  std::vector<float> velocities;

  // RegularGrid::max_velocity (RegularGrid.cpp:188-200) over the host copy
  double max_velocity() const {
    const size_t n = masses.size();
    double m = 0.0;
    for (size_t i = 0; i < n && velocities.size() == 3 * n; ++i) {
      double v0 = velocities[i], v1 = velocities[n + i], v2 = velocities[2 * n + i];
      double s = v0 * v0 + v1 * v1 + v2 * v2;
      if (s > m) m = s;
    }
    return std::sqrt(m);
  }
  double CFL_condition() const {
    double hm = h_[0] < h_[1] ? (h_[0] < h_[2] ? h_[0] : h_[2]) : (h_[1] < h_[2] ? h_[1] : h_[2]);
    return max_velocity() / hm;
  }

 private:
  std::array<double, 3> minBound_, maxBound_;
  std::array<int, 3> resolution_;
  std::array<double, 3> h_{};
};

}  // namespace mpm
