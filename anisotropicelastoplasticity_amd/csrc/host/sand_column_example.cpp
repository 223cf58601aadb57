// sand_column_example.cpp -- the reference's calling pattern (main.cpp:41-105 without the viewer):
// build a ParticleSystem and a RegularGrid, bind them to a HybridSolver with a level set and solve
// with the adaptive step, writing particle_N.obj frames; then checkpoint, continue, and check that a
// second solver restored from the checkpoint continues bit for bit.
// Usage: sand_column_example <output directory>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../../include/mpm/HybridSolver.h"

using namespace mpm;

static void configure(HybridSolver& solver) {
  mpm_params& p = solver.params();
  p.kernel_order = MPM_KERNEL_QUADRATIC;
  p.stress_at_predicted_F = 0;
  p.friction_model = MPM_FRICTION_COULOMB;
  p.friction = 0.4f;
  p.gravity[0] = 0.0f; p.gravity[1] = -9.8f; p.gravity[2] = 0.0f;
  p.dp_h[0] = 30.0f; p.dp_h[1] = 0.0f; p.dp_h[2] = 0.0f; p.dp_h[3] = 0.0f;
  const double blo[3] = {0.05, 0.05, 0.05}, bhi[3] = {0.95, 0.95, 0.95};
  solver.setLevelSet(LevelSet::box(blo, bhi));
}

int main(int argc, char** argv) {
  const std::string out = argc > 1 ? argv[1] : "sand_column_out";
  const double lo[3] = {0.4, 0.05, 0.4}, hi[3] = {0.6, 0.35, 0.6};
  const int32_t cells[3] = {8, 12, 8};
  const float v0[3] = {0.0f, 0.0f, 0.0f};
  ParticleSystem ps = ParticleSystem::Block(2, lo, hi, cells, 2, 2200.0, v0, MPM_SAND, 3.5e5, 0.3);
  RegularGrid rg({0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {64, 64, 64});
  HybridSolver solver(&ps, &rg);
  configure(solver);
  solver.setOutputDirectory(out);
  auto com = [](const ParticleSystem& q) {
    double c = 0.0;
    const size_t n = q.size();
    for (size_t i = 0; i < n; ++i) c += q.positions[n + i];
    return c / double(n);
  };
  const double y0 = com(ps);
  solver.solve(0.3, 1.5 / 60.0, 0.95, /*estimateVolumes*/ true);  // two frames of 1/60 s
  mpm_stats st = solver.stats();
  const double y1 = com(ps);
  std::printf("particles %zu substeps %lld frames %d com_y %.6f -> %.6f status %d\n", ps.size(),
              (long long)st.substeps, solver.framesWritten(), y0, y1, st.status);
  // checkpoint, then continue here and in a solver restored from the checkpoint
  const std::string ck = out + "/checkpoint.bin";
  solver.saveCheckpoint(ck);
  solver.run(5e-5, 10);
  solver.syncToHost();
  ParticleSystem ps2;
  RegularGrid rg2({0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {64, 64, 64});
  HybridSolver resumed(&ps2, &rg2);
  configure(resumed);
  resumed.loadCheckpoint(ck);
  resumed.run(5e-5, 10);
  resumed.syncToHost();
  const bool same = ps2.positions.size() == ps.positions.size() &&
                    std::memcmp(ps2.positions.data(), ps.positions.data(), sizeof(float) * ps.positions.size()) == 0 &&
                    std::memcmp(ps2.elasticDeformationGradients.data(), ps.elasticDeformationGradients.data(),
                                sizeof(float) * ps.elasticDeformationGradients.size()) == 0;
  std::printf("checkpoint resume bitwise %s\n", same ? "yes" : "no");
  return (st.status == 0 && y1 < y0 && solver.framesWritten() == 2 && same) ? 0 : 1;
}
