// cloth_example.cpp -- the reference's cloth calling pattern (main.cpp:41-105 without the viewer):
// LagrangianMesh::ObjMesh from an OBJ file, vertices 0 and 1 bound as fixed, a RegularGrid, no
// ParticleSystem, HybridSolver::setLagrangianMesh and an adaptive solve that writes
// mesh/mesh_N.obj frames. The OBJ file is a 24 x 24-vertex sheet this program writes first.
// Usage: cloth_example <output directory>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <string>

#include "../../../include/mpm/HybridSolver.h"

using namespace mpm;

int main(int argc, char** argv) {
  const std::string out = argc > 1 ? argv[1] : "cloth_out";
  std::filesystem::create_directories(out);
  const int side = 24;
  const double edge = 0.4 / (side - 1);
  const std::string obj = out + "/sheet.obj";
  LagrangianMesh::Grid(side, side, edge, 0.3, 0.3, 0.6, 2e3, 0.04, 200, 0.3, 0.0, 4e4, 0.0).writeObj(obj);

  // main.cpp:77-78: ObjMesh(file, density 2e3, thickness 0.04, E 200, nu 0.3, gamma 0, k 4e4, phi 0)
  LagrangianMesh mesh = LagrangianMesh::ObjMesh(obj, 2e3, 0.04, 200, 0.3, 0.0, 4e4, 0.0);
  std::vector<double> constraints(mesh.vertexCount(), 0.0);
  constraints[0] = 1.0;
  constraints[1] = 1.0;
  mesh.bindConstraints(constraints);
  const std::vector<float> x0 = mesh.vertexPositions;

  RegularGrid rg({0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {64, 64, 64});
  HybridSolver solver(nullptr, &rg);
  solver.setLagrangianMesh(&mesh);
  mpm_params& p = solver.params();
  p.kernel_order = MPM_KERNEL_QUADRATIC;
  p.friction_model = MPM_FRICTION_COULOMB;
  p.friction = 0.4f;
  p.gravity[0] = 0.0f; p.gravity[1] = -9.8f; p.gravity[2] = 0.0f;
  // main.cpp:86-90 binds a ground plane; here the plane y = 0.1 (gravity along -y) and a sphere
  const double centre[3] = {0.5, 0.45, 0.5};
  solver.setLevelSet(LevelSet::plane(0.0, 1.0, 0.0, 0.1) + LevelSet::sphere(centre, 0.1));
  solver.setOutputDirectory(out);
  solver.solve(0.3, 1.5 / 60.0, 0.95, /*estimateVolumes*/ true);  // two frames of 1/60 s
  const mpm_stats st = solver.stats();

  const size_t nv = mesh.vertexCount();
  double ymean0 = 0.0, ymean1 = 0.0;
  bool finite = true;
  for (size_t v = 0; v < nv; ++v) {
    ymean0 += x0[nv + v];
    ymean1 += mesh.vertexPositions[nv + v];
    for (int d = 0; d < 3; ++d) finite = finite && std::isfinite(mesh.vertexPositions[d * nv + v]);
  }
  ymean0 /= double(nv);
  ymean1 /= double(nv);
  bool pinned = true;
  for (int v : {0, 1})
    for (int d = 0; d < 3; ++d) pinned = pinned && mesh.vertexPositions[d * nv + v] == x0[d * nv + v];
  std::printf("vertices %zu faces %zu substeps %lld frames %d mean_y %.6f -> %.6f pinned %s status %d\n", nv,
              mesh.faceCount(), (long long)st.substeps, solver.framesWritten(), ymean0, ymean1, pinned ? "yes" : "no",
              st.status);
  return (st.status == 0 && finite && pinned && ymean1 < ymean0 && solver.framesWritten() == 2) ? 0 : 1;
}
