// HybridSolver.h -- the reference's solver API over the C ABI of libmpm_hip.so (header-only C++).
// Mirrors AnisotropicElastoplasticity/HybridSolver.h:27-95: a solver bound to caller-owned
// ParticleSystem / RegularGrid objects (raw, non-owning pointers, :30-32, :86-88) with
// setParticleSystem / setRegularGrid / setLevelSet / solve. Errors raise std::runtime_error
// instead of the reference's cerr + while(1).
#pragma once
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mpm_c_abi.h"
#include "LagrangianMesh.h"
#include "LevelSet.h"
#include "ParticleSystem.h"
#include "RegularGrid.h"

namespace mpm {

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
  // HybridSolver::setLagrangianMesh (HybridSolver.h:88): a cloth alone (no ParticleSystem, as in
  // main.cpp:77-84) or beside the particles; null detaches it.
  void setLagrangianMesh(LagrangianMesh* mesh) { mesh_ = mesh; dirty_ = true; }
  void setLevelSet(const LevelSet& ls) { ls_ = ls; dirty_ = true; }
  // The constants the reference hard-codes (HybridSolver.cpp:264-267, 457, 465, 641-644) and the
  // switches the configurations need (kernel order, stress at F^n, friction model, gravity).
  void setParams(const mpm_params& p) { params_ = p; dirty_ = true; }
  mpm_params& params() { dirty_ = true; return params_; }

  // HybridSolver::solve (HybridSolver.cpp:827-1034): the first P2G with the density estimate
  // (:854, when estimateVolumes), then the adaptive loop Dt = CFL / max(300, max|v_i| / h_min)
  // clamped to 1/60 s frames until t > maxt. After every frame the particle positions go to
  // <outputDirectory>/particle/particle_<frame>.obj as in the reference (:990-1005), if an output
  // directory is set, and the cloth's vertices and faces to <outputDirectory>/mesh/mesh_<frame>.obj
  // (:1006-1030) when a mesh is bound. `alpha` is unused, as in the reference.
  void solve(double CFL, double maxt, double /*alpha*/ = 0.0, bool estimateVolumes = true) {
    upload();
    if (estimateVolumes) check(mpm_estimate_volumes(ctx_), "mpm_estimate_volumes");
    if (!outdir_.empty()) {
      std::filesystem::create_directories(std::filesystem::path(outdir_) / "particle");
      if (mesh_) std::filesystem::create_directories(std::filesystem::path(outdir_) / "mesh");
    }
    check(mpm_solve(ctx_, CFL, maxt, &HybridSolver::onFrame, this), "mpm_solve");
    syncToHost();
  }
  // Fixed-step variant: substeps of dt until t > maxt (the configurations' fixed Dt).
  void solveFixed(double dt, double maxt, bool estimateVolumes = false) {
    upload();
    if (estimateVolumes) check(mpm_estimate_volumes(ctx_), "mpm_estimate_volumes");
    int n = 0;
    for (double t = 0.0; t <= maxt; t += dt) ++n;
    check(mpm_run(ctx_, float(dt), n), "mpm_run");
    syncToHost();
  }
  // Frame files of solve(): <dir>/particle/particle_N.obj (empty: no output).
  void setOutputDirectory(const std::string& dir) { outdir_ = dir; }
  int framesWritten() const { return frames_; }

  // Resumable binary snapshot of the particle state (positions, velocities, B, F, plastic, masses,
  // volumes, friction angles, the stress state of the next P2G) and the material; loadCheckpoint
  // restores it into the bound ParticleSystem and the device, and the following substeps repeat the
  // original run bit for bit.
  void saveCheckpoint(const std::string& path) {
    if (mesh_) throw std::runtime_error("saveCheckpoint covers particle systems without a cloth");
    syncToHost();
    const size_t n = ps_->size();
    std::vector<float> K(6 * n, 0.0f);
    const bool haveK = params_.stress_at_predicted_F == 0;
    if (haveK) check(mpm_get_stress(ctx_, K.data()), "mpm_get_stress");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    const uint64_t hdr[3] = {0x50414e534d504d31ull /* "1MPMSNAP" */, (uint64_t)n, haveK ? 1ull : 0ull};
    f.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    const int32_t mat = ps_->material;
    const double el[4] = {ps_->youngsModulus, ps_->poissonRatio, ps_->criticalCompression, ps_->criticalStretch};
    f.write(reinterpret_cast<const char*>(&mat), sizeof(mat));
    f.write(reinterpret_cast<const char*>(el), sizeof(el));
    auto put = [&](const std::vector<float>& v, size_t count) {
      std::vector<float> tmp(count, 0.0f);
      for (size_t i = 0; i < count && i < v.size(); ++i) tmp[i] = v[i];
      f.write(reinterpret_cast<const char*>(tmp.data()), sizeof(float) * count);
    };
    put(ps_->positions, 3 * n);
    put(ps_->velocities, 3 * n);
    put(ps_->affineMomenta, 9 * n);
    put(ps_->elasticDeformationGradients, 9 * n);
    put(ps_->plastic, n);
    put(ps_->masses, n);
    put(ps_->volumes, n);
    put(ps_->frictionAngles, n);
    put(K, 6 * n);
    if (!f) throw std::runtime_error("write failed: " + path);
  }
  void loadCheckpoint(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + path);
    uint64_t hdr[3];
    f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    if (!f || hdr[0] != 0x50414e534d504d31ull) throw std::runtime_error(path + " is not a checkpoint");
    const size_t n = (size_t)hdr[1];
    int32_t mat = 0;
    double el[4];
    f.read(reinterpret_cast<char*>(&mat), sizeof(mat));
    f.read(reinterpret_cast<char*>(el), sizeof(el));
    auto get = [&](std::vector<float>& v, size_t count) {
      v.resize(count);
      f.read(reinterpret_cast<char*>(v.data()), sizeof(float) * count);
    };
    std::vector<float> phi, K;
    get(ps_->positions, 3 * n);
    get(ps_->velocities, 3 * n);
    get(ps_->affineMomenta, 9 * n);
    get(ps_->elasticDeformationGradients, 9 * n);
    get(ps_->plastic, n);
    get(ps_->masses, n);
    get(ps_->volumes, n);
    get(phi, n);
    get(K, 6 * n);
    if (!f) throw std::runtime_error("truncated checkpoint " + path);
    ps_->material = mat;
    ps_->youngsModulus = el[0];
    ps_->poissonRatio = el[1];
    ps_->criticalCompression = el[2];
    ps_->criticalStretch = el[3];
    if (ps_->frictionAngles.size() == n || params_.per_particle_phi) ps_->frictionAngles = phi;
    dirty_ = true;
    upload();
    if (hdr[2]) check(mpm_set_stress(ctx_, K.data()), "mpm_set_stress");
  }
  void substep(double dt) { upload(); check(mpm_substep(ctx_, float(dt)), "mpm_substep"); }
  void run(double dt, int n) { upload(); check(mpm_run(ctx_, float(dt), n), "mpm_run"); }

  // Copy particle state, the cloth's vertices and element directions, and the grid's masses and
  // velocities back to the host objects.
  void syncToHost() {
    if (mesh_) {
      const size_t nv = mesh_->vertexCount(), nf = mesh_->faceCount();
      mesh_->vertexPositions.resize(3 * nv);
      mesh_->vertexVelocities.resize(3 * nv);
      mesh_->elementDirections.resize(9 * nf);
      check(mpm_get_mesh(ctx_, mesh_->vertexPositions.data(), mesh_->vertexVelocities.data(),
                         mesh_->elementDirections.data()),
            "mpm_get_mesh");
    }
    if (ps_ && ps_->size() > 0) {
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
    if (!rg_ || (!ps_ && !mesh_)) throw std::runtime_error("HybridSolver needs a RegularGrid and a ParticleSystem or a LagrangianMesh");
    const double mn[3] = {rg_->minBound()[0], rg_->minBound()[1], rg_->minBound()[2]};
    const double mx[3] = {rg_->maxBound()[0], rg_->maxBound()[1], rg_->maxBound()[2]};
    const int32_t res[3] = {rg_->resolution()[0], rg_->resolution()[1], rg_->resolution()[2]};
    check(mpm_set_grid(ctx_, mn, mx, res), "mpm_set_grid");
    mpm_params p = params_;
    const bool haveParticles = ps_ && ps_->size() > 0;
    if (haveParticles) {
      p.material = ps_->material;
      p.youngs_modulus = float(ps_->youngsModulus);
      p.poisson_ratio = float(ps_->poissonRatio);
      p.critical_compression = float(ps_->criticalCompression);
      p.critical_stretch = float(ps_->criticalStretch);
      p.per_particle_phi = ps_->frictionAngles.size() == ps_->size() ? 1 : 0;
    } else {
      p.per_particle_phi = 0;
    }
    check(mpm_set_params(ctx_, &p), "mpm_set_params");
    check(mpm_set_levelset(ctx_, ls_.prims.data(), (int32_t)ls_.prims.size()), "mpm_set_levelset");
    if (haveParticles) {
      const int32_t n = (int32_t)ps_->size();
      check(mpm_set_particles(ctx_, n, ps_->positions.data(), ps_->velocities.data(),
                              ps_->affineMomenta.empty() ? nullptr : ps_->affineMomenta.data(),
                              ps_->elasticDeformationGradients.empty() ? nullptr : ps_->elasticDeformationGradients.data(),
                              ps_->plastic.empty() ? nullptr : ps_->plastic.data(), ps_->masses.data(),
                              ps_->volumes.data(), p.per_particle_phi ? ps_->frictionAngles.data() : nullptr),
            "mpm_set_particles");
    }
    if (mesh_) {
      mesh_->validate();
      const mpm_mesh_params mp = mesh_->params();
      check(mpm_set_mesh(ctx_, (int32_t)mesh_->vertexCount(), mesh_->vertexPositions.data(),
                         mesh_->vertexVelocities.size() == mesh_->vertexPositions.size() ? mesh_->vertexVelocities.data()
                                                                                         : nullptr,
                         (int32_t)mesh_->faceCount(), mesh_->faces.data(),
                         mesh_->vertexIsFixed.empty() ? nullptr : mesh_->vertexIsFixed.data(), &mp),
            "mpm_set_mesh");
    }
    dirty_ = false;
  }

  static void onFrame(mpm_ctx* /*ctx*/, int32_t frame, double /*t*/, void* user) {
    HybridSolver* self = static_cast<HybridSolver*>(user);
    if (self->outdir_.empty()) return;
    self->syncToHost();
    if (self->ps_) {
      const std::string path = self->outdir_ + "/particle/particle_" + std::to_string(frame) + ".obj";
      std::ofstream out(path);
      const size_t n = self->ps_->size();
      const std::vector<float>& x = self->ps_->positions;  // planar: x0[n], x1[n], x2[n]
      for (size_t p = 0; p < n; ++p) out << "v " << x[p] << ' ' << x[n + p] << ' ' << x[2 * n + p] << '\n';
    }
    if (self->mesh_) self->mesh_->writeObj(self->outdir_ + "/mesh/mesh_" + std::to_string(frame) + ".obj");
    ++self->frames_;
  }

  std::string outdir_;
  int frames_ = 0;
  mpm_ctx* ctx_ = nullptr;
  ParticleSystem* ps_ = nullptr;
  RegularGrid* rg_ = nullptr;
  LagrangianMesh* mesh_ = nullptr;
  LevelSet ls_;
  mpm_params params_{};
  bool dirty_ = true;
};

}  // namespace mpm
