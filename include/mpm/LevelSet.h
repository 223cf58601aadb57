// LevelSet.h -- collision level sets as analytic primitives (host side).
// Replaces the std::function pair LevelSet / DLevelSet (AnisotropicElastoplasticity/LevelSet.h:5-6,
// HybridSolver.h:18-19), which cannot be evaluated on the GPU. The composite level set is the
// minimum over primitives; the normal comes from the primitive that attains it.
#pragma once
#include <vector>

#include "../mpm_c_abi.h"

namespace mpm {

struct LevelSet {
  std::vector<mpm_levelset> prims;

  // groundLevelSet(x, groundZ) = x[2] - groundZ, normal (0,0,1) (LevelSet.cpp:8-16)
  static LevelSet ground(double groundZ) { return plane(0.0, 0.0, 1.0, groundZ); }
  // phi = n.x - d with unit normal n
  static LevelSet plane(double n0, double n1, double n2, double d) {
    LevelSet l;
    l.prims.push_back(make(MPM_LS_PLANE, {float(n0), float(n1), float(n2), float(d)}, {0, 0, 0, 0}));
    return l;
  }
  // wall2groundLevelSet(x, wallX, wallY, groundZ) (LevelSet.cpp:18-42)
  static LevelSet wall2ground(double wallX, double wallY, double groundZ) {
    LevelSet l;
    l.prims.push_back(make(MPM_LS_WALL2GROUND, {float(wallX), float(wallY), float(groundZ), 0}, {0, 0, 0, 0}));
    return l;
  }
  // material inside the box [lo, hi]
  static LevelSet box(const double lo[3], const double hi[3]) {
    LevelSet l;
    l.prims.push_back(make(MPM_LS_BOX, {float(lo[0]), float(lo[1]), float(lo[2]), 0},
                           {float(hi[0]), float(hi[1]), float(hi[2]), 0}));
    return l;
  }
  // solid sphere obstacle
  static LevelSet sphere(const double c[3], double r) {
    LevelSet l;
    l.prims.push_back(make(MPM_LS_SPHERE, {float(c[0]), float(c[1]), float(c[2]), float(r)}, {0, 0, 0, 0}));
    return l;
  }
  LevelSet operator+(const LevelSet& o) const {
    LevelSet l = *this;
    l.prims.insert(l.prims.end(), o.prims.begin(), o.prims.end());
    return l;
  }

 private:
  static mpm_levelset make(int type, std::initializer_list<float> a, std::initializer_list<float> b) {
    mpm_levelset p{};
    p.type = type;
    int i = 0;
    for (float v : a) p.a[i++] = v;
    i = 0;
    for (float v : b) p.b[i++] = v;
    return p;
  }
};

}  // namespace mpm
