// HybridSolver.h -- the reference's solver API over the C ABI of libmpm_hip.so (header-only C++).
// Mirrors AnisotropicElastoplasticity/HybridSolver.h:27-95: a solver bound to caller-owned
// ParticleSystem / RegularGrid objects (raw, non-owning pointers, :30-32, :86-88) with
// setParticleSystem / setRegularGrid / setLevelSet / solve. Errors raise std::runtime_error
// instead of the reference's cerr + while(1).
#pragma once
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mpm_c_abi.h"
#include "LevelSet.h"
#include "ParticleSystem.h"
#include "RegularGrid.h"

namespace mpm {

class LagrangianMesh;  // cloth (LagrangianMesh.h:13-117): not available in this build

class HybridSolver {
 public:
  explicit HybridSolver(ParticleSystem* ps = nullptr, RegularGrid* rg = nullptr, int device = 0) : ps_(ps), rg_(rg) {
    if (mpm_create(&ctx_, device) != MPM_OK) throw std::runtime_error(std::string("mpm_create: ") + mpm_last_error(nullptr));
    mpm_params_default(&params_);
  }
  ~HybridSolver() { mpm_destroy(ctx_); }
  HybridSolver(const HybridSolver&) = delete;
  HybridSolver& operator=(const HybridSolver&) = delete;

  void setParticleSystem(ParticleSystem* ps) { ps_ = ps; dirty_ = true; }
  void setRegularGrid(RegularGrid* rg) { rg_ = rg; dirty_ = true; }
  void setLagrangianMesh(LagrangianMesh*) { throw std::runtime_error("cloth (LagrangianMesh) is not built yet"); }
  void setLevelSet(const LevelSet& ls) { ls_ = ls; dirty_ = true; }
  // The constants the reference hard-codes (HybridSolver.cpp:264-267, 457, 465, 641-644) and the
  // switches the configurations need (kernel order, stress at F^n, friction model, gravity).
  void setParams(const mpm_params& p) { params_ = p; dirty_ = true; }
  mpm_params& params() { dirty_ = true; return params_; }

  // HybridSolver::solve (HybridSolver.cpp:827-1034) with a fixed time step: the reference's first
  // P2G with the density estimate (:854, when estimateVolumes), then substeps until t > maxt, then
  // the particle state is copied back into the ParticleSystem. `alpha` is unused, as in the reference.
  void solve(double dt, double maxt, double /*alpha*/ = 0.0, bool estimateVolumes = false) {
    upload();
    if (estimateVolumes) check(mpm_estimate_volumes(ctx_), "mpm_estimate_volumes");
    int n = 0;
    for (double t = 0.0; t <= maxt; t += dt) ++n;
    check(mpm_run(ctx_, float(dt), n), "mpm_run");
    syncToHost();
  }
  void substep(double dt) { upload(); check(mpm_substep(ctx_, float(dt)), "mpm_substep"); }
  void run(double dt, int n) { upload(); check(mpm_run(ctx_, float(dt), n), "mpm_run"); }

  // Copy particle state (and the grid's masses and velocities) back to the host objects.
  void syncToHost() {
    if (ps_) {
      const size_t n = ps_->size();
      ps_->positions.resize(3 * n);
      ps_->velocities.resize(3 * n);
      ps_->affineMomenta.resize(9 * n);
      ps_->elasticDeformationGradients.resize(9 * n);
      ps_->plastic.resize(n);
      ps_->volumes.resize(n);
      check(mpm_get_particles(ctx_, ps_->positions.data(), ps_->velocities.data(), ps_->affineMomenta.data(),
                              ps_->elasticDeformationGradients.data(), ps_->plastic.data(), ps_->volumes.data()),
            "mpm_get_particles");
    }
    if (rg_) {
      const size_t ng = rg_->gridNumber();
      rg_->masses.resize(ng);
      rg_->velocities.resize(3 * ng);
      check(mpm_get_grid(ctx_, rg_->masses.data(), nullptr, rg_->velocities.data(), nullptr), "mpm_get_grid");
    }
  }

  mpm_stats stats() {
    mpm_stats s{};
    check(mpm_get_stats(ctx_, &s), "mpm_get_stats");
    return s;
  }
  mpm_ctx* context() { return ctx_; }

 private:
  void check(int rc, const char* what) {
    if (rc != MPM_OK) throw std::runtime_error(std::string(what) + ": " + mpm_last_error(ctx_));
  }
  void upload() {
    if (!dirty_) return;
    if (!rg_ || !ps_) throw std::runtime_error("HybridSolver needs a RegularGrid and a ParticleSystem");
    const double mn[3] = {rg_->minBound()[0], rg_->minBound()[1], rg_->minBound()[2]};
    const double mx[3] = {rg_->maxBound()[0], rg_->maxBound()[1], rg_->maxBound()[2]};
    const int32_t res[3] = {rg_->resolution()[0], rg_->resolution()[1], rg_->resolution()[2]};
    check(mpm_set_grid(ctx_, mn, mx, res), "mpm_set_grid");
    mpm_params p = params_;
    p.material = ps_->material;
    p.youngs_modulus = float(ps_->youngsModulus);
    p.poisson_ratio = float(ps_->poissonRatio);
    p.critical_compression = float(ps_->criticalCompression);
    p.critical_stretch = float(ps_->criticalStretch);
    p.per_particle_phi = ps_->frictionAngles.size() == ps_->size() ? 1 : 0;
    check(mpm_set_params(ctx_, &p), "mpm_set_params");
    check(mpm_set_levelset(ctx_, ls_.prims.data(), (int32_t)ls_.prims.size()), "mpm_set_levelset");
    const int32_t n = (int32_t)ps_->size();
    check(mpm_set_particles(ctx_, n, ps_->positions.data(), ps_->velocities.data(),
                            ps_->affineMomenta.empty() ? nullptr : ps_->affineMomenta.data(),
                            ps_->elasticDeformationGradients.empty() ? nullptr : ps_->elasticDeformationGradients.data(),
                            ps_->plastic.empty() ? nullptr : ps_->plastic.data(), ps_->masses.data(),
                            ps_->volumes.data(), p.per_particle_phi ? ps_->frictionAngles.data() : nullptr),
          "mpm_set_particles");
    dirty_ = false;
  }

  mpm_ctx* ctx_ = nullptr;
  ParticleSystem* ps_ = nullptr;
  RegularGrid* rg_ = nullptr;
  LevelSet ls_;
  mpm_params params_{};
  bool dirty_ = true;
};

}  // namespace mpm
