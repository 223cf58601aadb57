// sand_column_example.cpp -- the reference's calling pattern (main.cpp:41-105 without the viewer):
// build a ParticleSystem and a RegularGrid, bind them to a HybridSolver with a level set and run it.
// Prints the particle count, substeps run and the centre of mass before and after.
#include <cstdio>

#include "../../../include/mpm/HybridSolver.h"

int main() {
  using namespace mpm;
  const double lo[3] = {0.4, 0.05, 0.4}, hi[3] = {0.6, 0.35, 0.6};
  const int32_t cells[3] = {8, 12, 8};
  const float v0[3] = {0.0f, 0.0f, 0.0f};
  ParticleSystem ps = ParticleSystem::Block(2, lo, hi, cells, 2, 2200.0, v0, MPM_SAND, 3.5e5, 0.3);
  RegularGrid rg({0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {64, 64, 64});
  HybridSolver solver(&ps, &rg);
  mpm_params& p = solver.params();
  p.kernel_order = MPM_KERNEL_QUADRATIC;
  p.stress_at_predicted_F = 0;
  p.friction_model = MPM_FRICTION_COULOMB;
  p.friction = 0.4f;
  p.gravity[0] = 0.0f; p.gravity[1] = -9.8f; p.gravity[2] = 0.0f;
  p.dp_h[0] = 30.0f; p.dp_h[1] = 0.0f; p.dp_h[2] = 0.0f; p.dp_h[3] = 0.0f;
  const double blo[3] = {0.05, 0.05, 0.05}, bhi[3] = {0.95, 0.95, 0.95};
  solver.setLevelSet(LevelSet::box(blo, bhi));
  auto com = [&]() {
    double c = 0.0;
    const size_t n = ps.size();
    for (size_t i = 0; i < n; ++i) c += ps.positions[n + i];
    return c / double(n);
  };
  double y0 = com();
  solver.solve(1e-4, 0.02);
  mpm_stats st = solver.stats();
  double y1 = com();
  std::printf("particles %zu substeps %lld com_y %.6f -> %.6f status %d\n", ps.size(), (long long)st.substeps, y0, y1,
              st.status);
  return (st.status == 0 && y1 < y0) ? 0 : 1;
}
