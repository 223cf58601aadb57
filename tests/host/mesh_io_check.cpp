// mesh_io_check.cpp -- CPU checks of include/mpm/LagrangianMesh.h (no GPU calls): the OBJ reader of
// LagrangianMesh::ObjMesh (LagrangianMesh.cpp:198-280), the grid factory, the frame writer
// (HybridSolver.cpp:1006-1030), bindConstraints and the derived material constants
// (LagrangianMesh.cpp:331-351). Usage: mesh_io_check <scratch directory> <obj file to parse>;
// prints the parsed mesh of the second argument as "nv nf" then one line per vertex and face.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/mpm/LagrangianMesh.h"

using mpm::LagrangianMesh;

static int fails = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      ++fails;                                                     \
    }                                                              \
  } while (0)

template <class F>
static bool throws(F f) {
  try {
    f();
  } catch (const std::runtime_error&) {
    return true;
  }
  return false;
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string dir = argv[1];
  // grid factory -> writer -> reader round trip (positions written with 6 significant digits)
  LagrangianMesh g = LagrangianMesh::Grid(5, 4, 0.125, 0.25, 0.5, 0.75, 2e3, 0.04, 200, 0.3, 1e3, 4e4, 30.0);
  CHECK(g.vertexCount() == 20 && g.faceCount() == 2 * 4 * 3);
  g.writeObj(dir + "/grid.obj");
  LagrangianMesh r = LagrangianMesh::ObjMesh(dir + "/grid.obj", 2e3, 0.04, 200, 0.3, 1e3, 4e4, 30.0);
  CHECK(r.vertexCount() == g.vertexCount() && r.faces == g.faces);
  for (size_t i = 0; i < g.vertexPositions.size(); ++i) CHECK(r.vertexPositions[i] == g.vertexPositions[i]);
  for (float v : r.vertexVelocities) CHECK(v == 0.0f);
  // every triangle of the grid is counter-clockwise seen from +y: (d1 x d2).y > 0
  const size_t nv = g.vertexCount();
  for (size_t e = 0; e < g.faceCount(); ++e) {
    const int a = g.faces[3 * e], b = g.faces[3 * e + 1], c = g.faces[3 * e + 2];
    const double d1x = g.vertexPositions[b] - g.vertexPositions[a], d1z = g.vertexPositions[2 * nv + b] - g.vertexPositions[2 * nv + a];
    const double d2x = g.vertexPositions[c] - g.vertexPositions[a], d2z = g.vertexPositions[2 * nv + c] - g.vertexPositions[2 * nv + a];
    CHECK(d1z * d2x - d1x * d2z > 0.0);
  }
  // material constants: E = 200, nu = 0.3 -> mu = 76.923..., lambda = 115.384...; tan 30 deg
  CHECK(std::fabs(g.mu() - 200.0 / 2.6) < 1e-12);
  CHECK(std::fabs(g.lambda() - 200.0 * 0.3 / 1.3 / 0.4) < 1e-12);
  CHECK(std::fabs(g.frictionCoeff() - 1.0 / std::sqrt(3.0)) < 1e-12);
  const mpm_mesh_params mp = g.params();
  CHECK(mp.youngs_modulus == 200.0f && mp.poisson_ratio == 0.3f && mp.shear_stiffness == 1e3f && mp.stiffness == 4e4f &&
        mp.friction_angle == 30.0f && mp.density == 2e3f && mp.thickness == 0.04f && mp.reference_normal_row == 0);
  // constraints
  std::vector<double> fixedOk(nv, 0.0);
  fixedOk[0] = fixedOk[1] = 1.0;
  g.bindConstraints(fixedOk);
  CHECK(g.isFixed(0) && g.isFixed(1) && !g.isFixed(2));
  CHECK(throws([&] { g.bindConstraints(std::vector<double>(nv + 1, 0.0)); }));
  // malformed files
  {
    std::ofstream(dir + "/bad0.obj") << "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 0\n";
    std::ofstream(dir + "/bad4.obj") << "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n";
    std::ofstream(dir + "/bad2.obj") << "v 0 0 0\nv 1 0 0\nf 1 2\n";
  }
  for (const char* b : {"bad0.obj", "bad4.obj", "bad2.obj", "missing.obj"})
    CHECK(throws([&] { LagrangianMesh::ObjMesh(dir + "/" + b, 1, 1, 1, 0.3, 0, 0, 0); }));
  CHECK(throws([&] { LagrangianMesh::Grid(1, 4, 0.1, 0, 0, 0, 1, 1, 1, 0.3, 0, 0, 0); }));
  // the file named on the command line, parsed and printed for the caller to compare
  LagrangianMesh m = LagrangianMesh::ObjMesh(argv[2], 2e3, 0.04, 200, 0.3, 0.0, 4e4, 0.0);
  const size_t n = m.vertexCount();
  std::printf("%zu %zu\n", n, m.faceCount());
  for (size_t v = 0; v < n; ++v)
    std::printf("%.9g %.9g %.9g\n", m.vertexPositions[v], m.vertexPositions[n + v], m.vertexPositions[2 * n + v]);
  for (size_t e = 0; e < m.faceCount(); ++e) std::printf("%d %d %d\n", m.faces[3 * e], m.faces[3 * e + 1], m.faces[3 * e + 2]);
  if (fails) std::fprintf(stderr, "%d checks failed\n", fails);
  return fails ? 1 : 0;
}
