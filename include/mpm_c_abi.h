/*
 * mpm_c_abi.h -- C ABI of the MI355X (gfx950) explicit MPM substep library (libmpm_hip.so).
 *
 * This is the drop-in boundary for the hot path of the reference solver:
 * the body of the `while (t <= maxt)` loop of HybridSolver::solve
 * (AnisotropicElastoplasticity/HybridSolver.cpp:867-989) and the private methods it calls.
 * The reference has no C ABI of its own: its callers hold raw pointers to ParticleSystem /
 * RegularGrid / LagrangianMesh objects and call HybridSolver methods (HybridSolver.h:83-93).
 * The C++ host classes in include/mpm/ keep that API and forward to the functions below;
 * each function names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns an int status (MPM_OK = 0). No C++ exception crosses the ABI.
 *     mpm_last_error() returns a message for the last failure on a context.
 *     (The reference reports errors with cerr and exit()/while(1), e.g. RegularGrid.cpp:129-134.)
 *   - The caller owns host arrays; they are copied in and out. The context owns device memory.
 *   - Particle arrays are structure-of-arrays planes of n floats: x[3n] = {x0[n], x1[n], x2[n]};
 *     3x3 matrices are 9 row-major planes: M[(3*a+b)*n + p] = M_ab of particle p.
 *     (The reference's MatrixX3d members are column-major n x 3, i.e. already one plane per
 *     component; affineMomenta_1..3 are the rows of B; ParticleSystem.h:27-35.)
 *   - Particle order seen by the caller never changes: getters return arrays in the order given
 *     to mpm_set_particles, whatever order the device keeps them in.
 *   - One context per host thread. Calls are asynchronous on the context's HIP stream;
 *     getters synchronise.
 *   - Device state is FP32. FP64 is used on the device only in the 3x3 SVD of near-degenerate
 *     deformation gradients and in the Drucker-Prager projection near the cone apex.
 */
#ifndef MPM_C_ABI_H
#define MPM_C_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MPM_OK 0
#define MPM_ERR_INVALID 1     /* bad argument or size */
#define MPM_ERR_HIP 2         /* a HIP runtime call failed */
#define MPM_ERR_STATE 3       /* call order (e.g. substep before set_particles) */
#define MPM_ERR_NOMEM 4       /* device allocation failed */
#define MPM_ERR_UNSUPPORTED 5 /* option not implemented */

/* ---- enums ---- */
/* Interpolation kernel. Reference: cubic B-spline (interpolation.cpp:9-33), D^-1 = 3/h^2
 * (HybridSolver.cpp:175-177). Quadratic B-spline with D^-1 = 4/h^2 is the standard APIC choice. */
#define MPM_KERNEL_QUADRATIC 2
#define MPM_KERNEL_CUBIC 3

/* Material. Reference enum MaterialType {SNOW, SAND} (HybridSolver.h:21-25). */
#define MPM_SNOW 0
#define MPM_SAND 1

/* Grid collision friction model (HybridSolver.cpp:460-511).
 * REFERENCE: the reference as written: on approach, v_BF = v_t and v = 0 if |v_t| < -mu v_n else
 *            v = v_t (the Coulomb term at :500-501 is lost to a stray ';').
 * COULOMB:   the intended rule: v = v_t + mu v_n v_t/|v_t| (or 0), v_BF = v_t.
 * STICKY:    every node inside the level set gets v = v_BF = 0. */
#define MPM_FRICTION_REFERENCE 0
#define MPM_FRICTION_COULOMB 1
#define MPM_FRICTION_STICKY 2

/* Level-set primitives; the composite level set is the minimum over primitives and the normal
 * is taken from the primitive that attains it. Replaces the std::function pair LevelSet/DLevelSet
 * (LevelSet.h:5-6), which cannot cross to the GPU.
 *   PLANE       phi = n.x - d                  a = {n0,n1,n2,d}      (groundLevelSet: n=(0,0,1), d=groundZ,
 *                                                                   LevelSet.cpp:8-16)
 *   WALL2GROUND phi = min(x2-g, wx-x0, wy-x1), normal of the nearest plane by |distance|
 *                                              a = {wx, wy, g, 0}    (LevelSet.cpp:18-42)
 *   BOX         material inside [lo,hi]: phi = min_i min(x_i-lo_i, hi_i-x_i), inward face normal
 *                                              a = {lo0,lo1,lo2,0} b = {hi0,hi1,hi2,0}
 *   SPHERE      solid obstacle: phi = |x-c| - r, normal (x-c)/|x-c|
 *                                              a = {c0,c1,c2,r} */
#define MPM_LS_PLANE 0
#define MPM_LS_WALL2GROUND 1
#define MPM_LS_BOX 2
#define MPM_LS_SPHERE 3

/* Stages of one substep, in order (HybridSolver.cpp:867-989 reordered so that P2G opens the
 * substep; the reference's initial P2G at :854 becomes the first substep's P2G). */
#define MPM_STAGE_SORT 0  /* cell keys -> stable radix sort -> bins                         */
#define MPM_STAGE_P2G 1   /* particleToGrid_ :113-250 (+ computeGridForces_ :252-368)        */
#define MPM_STAGE_GRID 2  /* updateGridVelocities_ :725-737, gravity :457, collision :460-551 */
#define MPM_STAGE_G2P 3   /* :739-825, advection :942-950, :553-609, updatePlasticity_ :612-681 */
#define MPM_NUM_STAGES 4

/* Status flags reported by mpm_get_stats (sticky until mpm_clear_status). */
#define MPM_STATUS_NONFINITE 1      /* a particle state value became NaN/Inf */
#define MPM_STATUS_OUT_OF_DOMAIN 2  /* a particle cell was clamped to the grid */
#define MPM_STATUS_CFL 4            /* a particle moved more than one cell in a substep */

/* Simulation parameters. Defaults (mpm_params_default) reproduce the reference's hard-coded
 * values; every deviation the configurations need is one switch here, not a second code path. */
typedef struct mpm_params {
  int32_t kernel_order;          /* MPM_KERNEL_*; reference = CUBIC */
  int32_t material;              /* MPM_SNOW / MPM_SAND; reference solve() hard-codes SAND (:873,955,959) */
  int32_t stress_at_predicted_F; /* 1: stress at F^ = (I + dt grad v)F from post-P2G grid velocity, times
                                    F^T (reference, :269-306,321-338); 0: stress at F^n (standard explicit) */
  int32_t friction_model;        /* MPM_FRICTION_*; reference = REFERENCE */
  float gravity[3];              /* reference: (0,0,-9.8) as a force m*g (:457) */
  float youngs_modulus;          /* E (ParticleSystem::youngsModulus) */
  float poisson_ratio;           /* nu */
  float critical_compression;    /* snow theta_c (ParticleSystem.cpp:173-176) */
  float critical_stretch;        /* snow theta_s */
  float hardening;               /* snow xi, lambda,mu scale exp(xi(1-J_P)) (:267,285-286) */
  float dp_h[4];                 /* sand friction angle in degrees: phi(q) = h0 + (h1 q - h3) exp(-h2 q)
                                    (:641-647); reference (35, 9, 0.2, 10); constant phi <=> h1 = h3 = 0 */
  int32_t per_particle_phi;      /* 1: a per-particle friction angle (degrees) replaces h0 */
  float friction;                /* grid collision friction coefficient; reference 0.2 (:465) */
  float apic_damping;            /* C <- skew + (1-d) sym (:808-824); reference d = 0 for particles */
  float fp64_svd_ratio;          /* redo the SVD in FP64 when sigma_min < ratio * sigma_max */
  float fp64_apex_radius;        /* run the DP cone projection in FP64 when |eps_trial| < radius */
  int32_t reserved[4];
} mpm_params;

typedef struct mpm_levelset {
  int32_t type; /* MPM_LS_* */
  float a[4];
  float b[4];
} mpm_levelset;

#define MPM_MAX_LEVELSETS 16

/* One scene of a batched context (mpm_set_scenes): its grid, parameters, level sets and particle
 * count. The fields mean what the arguments of mpm_set_grid, mpm_set_params and mpm_set_levelset
 * mean for a single scene. */
typedef struct mpm_scene_desc {
  double min_bound[3];
  double max_bound[3];
  int32_t res[3];
  int32_t n_particles;
  mpm_params params;
  int32_t n_levelsets;
  mpm_levelset levelsets[MPM_MAX_LEVELSETS];
} mpm_scene_desc;

typedef struct mpm_stats {
  int64_t substeps;          /* substeps run since mpm_set_particles */
  int32_t status;            /* OR of MPM_STATUS_* */
  int32_t n_particles;
  int32_t n_active_tiles;    /* 4x4x4-cell tiles holding particles, last substep */
  int32_t n_active_blocks;   /* 4x4x4-node grid blocks updated, last substep */
  int64_t n_active_nodes;    /* nodes with m > 0, last substep (counted on request) */
  int64_t fp64_svd_count;    /* SVDs redone in FP64, cumulative */
  int64_t fp64_apex_count;   /* DP projections run in FP64, cumulative */
  int64_t surrogate_candidates; /* cone-branch particles handed to the surrogate, cumulative */
  int64_t surrogate_accepted;   /* of which the prediction passed the gate, cumulative */
  int64_t pairs_logged;         /* (trial, projected) pairs logged since mpm_set_pair_logging */
} mpm_stats;

typedef struct mpm_ctx mpm_ctx; /* opaque: one HIP stream + device buffers */

/* Fill p with the reference's defaults: cubic kernel, sand, stress at predicted F, reference
 * friction with mu = 0.2, gravity (0,0,-9.8), snow (1.4e5, 0.2, 2.5e-2, 7.5e-3, 10) from
 * ParticleSystem.cpp:173-176 and HybridSolver.cpp:267, DP hardening (35, 9, 0.2, 10) from
 * HybridSolver.cpp:641-644. */
void mpm_params_default(mpm_params* p);

/* Create a context on HIP device `device`. Replaces HybridSolver::HybridSolver (HybridSolver.h:83-84). */
int mpm_create(mpm_ctx** out, int32_t device);
/* Release all device memory and the stream. */
int mpm_destroy(mpm_ctx* ctx);
/* Message of the last failure on ctx (or of the last failed mpm_create when ctx is NULL). */
const char* mpm_last_error(const mpm_ctx* ctx);

/* Node-centred grid: node (i,j,k) sits at min + (i,j,k)*h with h = (max-min)/res, linear index
 * k*res0*res1 + j*res0 + i. Mirrors RegularGrid::RegularGrid (RegularGrid.cpp:117-162),
 * RegularGrid::toIndex (:164-168) and HybridSolver::setRegularGrid (HybridSolver.h:87). */
int mpm_set_grid(mpm_ctx* ctx, const double min_bound[3], const double max_bound[3], const int32_t res[3]);

/* Replaces the constants hard-coded in HybridSolver.cpp (:264-267, :465, :641-644, :457). */
int mpm_set_params(mpm_ctx* ctx, const mpm_params* params);

/* Replaces HybridSolver::setLevelSet(LevelSet, DLevelSet) (HybridSolver.h:89). n <= 16. */
int mpm_set_levelset(mpm_ctx* ctx, const mpm_levelset* prims, int32_t n);

/* Batched scenes (no counterpart in the reference, which simulates one ParticleSystem per
 * HybridSolver): nscenes independent scenes in one context. Every substep stage covers all scenes
 * in one launch (one radix sort over scene-major keys, one P2G, one grid update, one G2P); each
 * scene evolves exactly as it would alone in its own context, bit for bit. The scenes must share
 * kernel_order and stress_at_predicted_F; mpm_set_particles then takes all particles concatenated
 * scene by scene (n = sum of n_particles), and mpm_get_particles returns them in that order.
 * mpm_set_grid / mpm_set_params / mpm_set_levelset apply to single-scene contexts only. */
int mpm_set_scenes(mpm_ctx* ctx, int32_t nscenes, const mpm_scene_desc* scenes);

/* Mirrors the ParticleSystem constructor (ParticleSystem.cpp:81-117) and
 * HybridSolver::setParticleSystem (HybridSolver.h:86). plastic = J_P = det F_P for snow (the only
 * use of F_P, HybridSolver.cpp:283) and plasticAmount q for sand (:646,665,672). phi (degrees)
 * may be NULL unless params.per_particle_phi. B may be NULL (zero) -- the reference leaves it
 * uninitialised (ParticleSystem.cpp:113-115). */
int mpm_set_particles(mpm_ctx* ctx, int32_t n, const float* x, const float* v, const float* B,
                      const float* F, const float* plastic, const float* mass, const float* volume,
                      const float* phi);

/* Cloth material of a LagrangianMesh (LagrangianMesh.h:66-73, LagrangianMesh::ObjMesh
 * LagrangianMesh.cpp:300-352): in-plane fixed-corotated Young's modulus and Poisson ratio, the
 * shear stiffness gamma and normal stiffness k of the normal/shear response, and the friction angle
 * whose tangent bounds the shear in the return mapping (HybridSolver.cpp:684-722). */
typedef struct mpm_mesh_params {
  float youngs_modulus;
  float poisson_ratio;
  float shear_stiffness;  /* gamma */
  float stiffness;        /* k */
  float friction_angle;   /* degrees */
  float density;
  float thickness;        /* element volume = 0.25 x area x thickness (LagrangianMesh.cpp:312) */
  int32_t reference_normal_row; /* 1: the normal-force direction vector is (D_rest^T).col(2) as in
                                   HybridSolver.cpp:432 (equal to D_rest^-T e3 for a mesh in the
                                   z = const plane only); 0: D_rest^-T e3 */
} mpm_mesh_params;

/* Attach a triangle cloth to a single-scene context after mpm_set_particles and before the first
 * substep, or to a context without particles (a cloth alone, as in the reference's main.cpp:77-84;
 * a later call then replaces that cloth) (replaces HybridSolver::setLagrangianMesh,
 * HybridSolver.h:88, and the LagrangianMesh constructor, LagrangianMesh.cpp:127-195). x and v are planar 3 x nv (v may be NULL), faces are
 * nf x 3 vertex indices, fixed (nv bytes, may be NULL) marks vertices whose surrounding nodes are
 * stopped (LagrangianMesh::bindConstraints, HybridSolver.cpp:513-550). The mesh's vertices and
 * triangles move with the particles in every substep; mpm_get_particles keeps returning the
 * particles only. */
int mpm_set_mesh(mpm_ctx* ctx, int32_t nv, const float* x, const float* v, int32_t nf, const int32_t* faces,
                 const uint8_t* fixed, const mpm_mesh_params* params);
/* Vertex positions and velocities (3 x nv planar) and the element directions D = [d1 d2 d3]
 * (9 x nf, row-major per element); any pointer may be NULL. */
int mpm_get_mesh(mpm_ctx* ctx, float* vx, float* vv, float* D);

/* Copy particle state out in caller order; any pointer may be NULL. */
int mpm_get_particles(mpm_ctx* ctx, float* x, float* v, float* B, float* F, float* plastic,
                      float* volume);
/* The volume-weighted Kirchhoff stress V tau (6 x n planes: xx yy zz xy xz yz, caller order) that
 * the next P2G scatters (stress at F^n). A checkpoint that stores it next to mpm_get_particles and
 * restores both (mpm_set_particles, then mpm_set_stress) resumes bit for bit; without it the next
 * substep recomputes the stress from F. No counterpart in the reference, which keeps no such state. */
int mpm_get_stress(mpm_ctx* ctx, float* K);
int mpm_set_stress(mpm_ctx* ctx, const float* K);

/* The reference's first P2G with evaluateVolumesAndDensities = true (HybridSolver.cpp:242-249):
 * rho_p = sum_i w_ip m_i / h^3, V_p = m_p / rho_p, at the current positions. */
int mpm_estimate_volumes(mpm_ctx* ctx);

/* One explicit substep with fixed dt (HybridSolver.cpp:873-987). */
int mpm_substep(mpm_ctx* ctx, float dt);
/* nsteps substeps with fixed dt, replayed from one captured HIP graph. */
int mpm_run(mpm_ctx* ctx, float dt, int32_t nsteps);
/* Called by mpm_solve on the calling thread after the substep that completes frame `frame`
 * (0-based) at simulated time t; the particle state may be read from inside the callback. */
typedef void (*mpm_frame_callback)(mpm_ctx* ctx, int32_t frame, double t, void* user);
/* The reference's adaptive loop, HybridSolver::solve(CFL, maxt, alpha) (HybridSolver.cpp:827-1034;
 * alpha is unused there as here): Dt = CFL / max(300, max|v_i| / h_min) from the grid velocities of
 * a first P2G, then from the velocities after each velocity update; each substep is clamped to the
 * end of its 1/60 s frame, t advances by 1/60 per frame, and the loop ends after the frame that
 * takes t past maxt. Dt is evaluated on the device; cb (may be null) runs at every frame end. */
int mpm_solve(mpm_ctx* ctx, double cfl, double maxt, mpm_frame_callback cb, void* user);
/* Launch shape of the P2G and grid kernels. None of the shapes changes an order of summation, so
 * results are bitwise identical for every valid shape. */
#define MPM_SORT_COUNTING 0 /* one-digit counting sort over the cells, stable rank inside each cell */
#define MPM_SORT_RADIX 1    /* LSD radix sort, 8-bit digits */
#define MPM_P2G_ROWS 0      /* P2G: a wave walks cell rows, 16 lanes per cell */
#define MPM_P2G_LAYERS 1    /* P2G: a wave holds a whole cell layer, 4 lanes per cell */
typedef struct mpm_tuning {
  int32_t p2g_ranks;   /* particles per cell staged per P2G chunk (the bin tile): 4 or 8 (rows), 2 or 4 (layers) */
  int32_t p2g_blocks;  /* P2G workgroups, persistent over the active 4x4x4-cell tiles */
  int32_t grid_blocks; /* grid-update workgroups, persistent over the active node blocks */
  float p2g_ms;        /* times of the chosen shapes measured by mpm_autotune (0 otherwise) */
  float grid_ms;
  int32_t sort_mode;   /* MPM_SORT_*: both give the same order, bit for bit */
  float sort_ms;
  int32_t p2g_pass;    /* MPM_P2G_*: the same sums in the same order, bit for bit */
} mpm_tuning;
/* Autotuner (no counterpart in the reference): time every candidate shape `reps` times on the
 * current particle state (sort, P2G and grid update are idempotent, the state does not advance),
 * keep the fastest and report it in *chosen (may be NULL). */
int mpm_autotune(mpm_ctx* ctx, float dt, int32_t reps, mpm_tuning* chosen);
int mpm_get_tuning(mpm_ctx* ctx, mpm_tuning* tuning);
int mpm_set_tuning(mpm_ctx* ctx, const mpm_tuning* tuning);
/* Clock of the last mpm_solve: simulated time, last Dt, frames completed, substeps run. */
int mpm_get_clock(mpm_ctx* ctx, double* t, double* dt, int32_t* frames, int64_t* substeps);
/* Run a single stage (MPM_STAGE_*) of the current substep; stages must be issued in order
 * SORT, P2G, GRID, G2P. Used by per-stage parity tests. */
int mpm_stage(mpm_ctx* ctx, int32_t stage, float dt);

/* Result of the last SORT stage: keys in sorted order (cell index k*r0*r1 + j*r0 + i of
 * trunc((x-min)/h)), the caller-order particle id at each sorted position, and bin offsets
 * (n_cells+1 exclusive prefix of the per-cell counts). Equals std::stable_sort of the keys by
 * caller index. Any pointer may be NULL. */
int mpm_get_sorted(mpm_ctx* ctx, uint32_t* keys, int32_t* ids, int32_t* bins);

/* Dense grid arrays after the last P2G / GRID stage. mom = momentum accumulated by P2G
 * (plus dt * stress force when stress is evaluated at F^n); v, vbf = velocity after the grid
 * update and collision, and before friction (HybridSolver.cpp:463,492). Nodes outside the
 * blocks touched by particles read 0. Any pointer may be NULL; m has n_nodes floats, others 3*n_nodes
 * (planar). */
int mpm_get_grid(mpm_ctx* ctx, float* m, float* mom, float* v, float* vbf);

int mpm_get_stats(mpm_ctx* ctx, mpm_stats* stats);
int mpm_clear_status(mpm_ctx* ctx);
int mpm_synchronize(mpm_ctx* ctx);

/* Timing helper for benchmarks: run nsteps substeps, recording HIP events on the context stream
 * around every stage; stage_ms[MPM_NUM_STAGES] receives the summed stage times and *total_ms the
 * whole-span time. */
int mpm_time_substeps(mpm_ctx* ctx, float dt, int32_t nsteps, float* total_ms, float* stage_ms);
/* As mpm_time_substeps, and kernel_ms[MPM_NUM_STAGES] receives the summed time of each stage's main
 * kernel: for G2P, the k_g2p dispatch alone (its start and end, as the launch records them into HIP
 * events), without the deferred FP64 apex projections that follow it in the stage; for the other
 * stages, the stage time. No counterpart in the reference. */
int mpm_time_kernels(mpm_ctx* ctx, float dt, int32_t nsteps, float* total_ms, float* stage_ms, float* kernel_ms);
/* Measurement helper with no counterpart in the reference: nsteps eagerly launched substeps with the
 * surrogate enabled; kernel_ms receives the summed dispatch times (HIP events) of the surrogate kernel
 * (k_surrogate: network, gate and state update), candidates the number of candidates it processed. */
int mpm_time_surrogate(mpm_ctx* ctx, float dt, int32_t nsteps, float* kernel_ms, int64_t* candidates);

/* ---- learned return-map surrogate (no counterpart in the reference; replaces the cone branch of
 * the Drucker-Prager projection, HybridSolver.cpp:667-673, behind a gate that falls back to it) ----
 * A 5 -> 128 -> 128 -> 128 -> 128 -> 3 SiLU MLP with BF16 weights and activations and FP32
 * accumulation (gfx950 MFMA). Input row: trial Hencky principal strains eps_1 >= eps_2 >= eps_3,
 * friction angle (degrees), K/mu. Output: projected principal strains H. Weights are one flat FP32
 * vector of mpm_surrogate_nparam() entries (the master copy; padding entries are kept at zero). */
int32_t mpm_surrogate_nparam(void);
int mpm_surrogate_init(mpm_ctx* ctx, uint64_t seed);
int mpm_surrogate_set_weights(mpm_ctx* ctx, const float* w, int32_t nparam);
int mpm_surrogate_get_weights(mpm_ctx* ctx, float* w, int32_t nparam);
/* H (n x 3) predicted for inputs in (n x 5), row-major. */
int mpm_surrogate_predict(mpm_ctx* ctx, const float* in, float* H, int32_t n);
/* Gradient of L = sum (y - t)^2 / n over normalised outputs y = (H - eps)/1e-3, with the same kernels
 * the trainer uses; *loss receives L. */
int mpm_surrogate_gradient(mpm_ctx* ctx, const float* in, const float* H, int32_t n, float* grad, float* loss);
/* Allocation-free variant of mpm_surrogate_gradient. It has no counterpart in the reference; it is
 * for callers that cannot allocate device memory: code under HIP graph capture, tight serving or
 * online-training loops, harnesses that forbid hipMalloc. in (n x 5), H (n x 3), grad
 * (mpm_surrogate_nparam() floats) and loss (1 float, the mean loss) are DEVICE pointers. The call
 * only enqueues kernels on the context's stream (results are ready after mpm_synchronize) and
 * allocates nothing: its intermediates live in one fixed library-owned device array of about
 * 9.3 MB per GPU, which keeps the last call's contents. n must be in [1, 2048] (MPM_ERR_INVALID
 * otherwise). Calls that overlap on one GPU (from different contexts or threads) race on that
 * array. Results equal mpm_surrogate_gradient's bit for bit. */
int mpm_surrogate_gradient_noalloc(mpm_ctx* ctx, const float* in, const float* H, int32_t n, float* grad, float* loss);
/* Adam on the GPU over `epochs` passes of minibatches of `batch` rows in a seeded shuffled order;
 * epoch_loss[epochs] receives the mean loss of each pass. */
int mpm_surrogate_train(mpm_ctx* ctx, const float* in, const float* H, int32_t n, int32_t epochs, int32_t batch,
                        float lr, uint64_t seed, float* epoch_loss);
/* Log (trial, projected) cone-branch pairs emitted by the exact return map in later substeps:
 * capacity pairs of 8 floats (eps(3), phi, K/mu, H(3)); 0 disables logging. */
int mpm_set_pair_logging(mpm_ctx* ctx, int64_t capacity);
int mpm_get_pairs(mpm_ctx* ctx, float* pairs, int64_t max_pairs, int64_t* count);
/* Enable the surrogate in the substep: cone-branch particles take the network's H when it passes the
 * gate (yield equality, volume and flow direction within gate_tol), else the exact projection. */
int mpm_set_surrogate(mpm_ctx* ctx, int32_t enable, float gate_tol);
/* Diagnostic with no counterpart in the reference: record the gate decision of each of the first
 * `capacity` surrogate candidates of every substep (0 disables), 10 floats per candidate: sorted
 * slot (int32 bits), eps(3), friction angle, K/mu, predicted H(3), accepted (1 or 0). The getter
 * returns the records of the last substep run with the surrogate enabled. */
int mpm_set_surrogate_log(mpm_ctx* ctx, int64_t capacity);
int mpm_get_surrogate_log(mpm_ctx* ctx, float* records, int64_t max_records, int64_t* count);

/* Measurement helper with no counterpart in the reference: copies nbytes (a multiple of 16) between
 * two device buffers `reps` times with a plain float4 copy kernel (k_copy). Under rocprofv3 --pmc its
 * FETCH_SIZE and WRITE_SIZE readings calibrate the counters against a known byte count. */
int mpm_bench_copy(mpm_ctx* ctx, int64_t nbytes, int32_t reps, float* ms);

/* Seeded jittered particle generator (splitmix64 counter hash). The box [lo,hi) is cut into a
 * lattice of cells[0] x cells[1] x cells[2] cells; each cell holds ppa^3 particles on a stratified
 * sub-lattice, each jittered uniformly within its sub-cell. Particle id
 * = ((k*cells1 + j)*cells0 + i)*ppa^3 + (sk*ppa + sj)*ppa + si; jitter component c of particle id
 * is u = (splitmix64(seed*0x632BE59BD9B4E019 ^ (4*id + c)) >> 40) * 2^-24 and the coordinate is
 * lo + (hi-lo) * (cell + (sub + u)/ppa) / cells, evaluated in double and rounded to float.
 * n must equal cells0*cells1*cells2*ppa^3. Replaces the unseeded samplers of the reference
 * factories (ParticleSystem.cpp:119-401). */
int mpm_generate_block(uint64_t seed, const double lo[3], const double hi[3], const int32_t cells[3],
                       int32_t ppa, int64_t n, float* x);

/* Seeded rejection samplers of the reference's shape factories (ParticleSystem.cpp:119-401), which
 * draw from an unseeded Eigen Random / default_random_engine(time(0)). Candidate k (k = 0, 1, ...)
 * has the uniforms u_c = (splitmix64(s ^ (4*k + c)) >> 40) * 2^-24, c = 0, 1, 2, with
 * s = splitmix64(seed*0x632BE59BD9B4E019 ^ shape); the accepted candidates fill x (3 x n planar)
 * in order. Arithmetic is in double, rounded to float at the end.
 *   MPM_SHAPE_BALL      a = {cx, cy, cz, r}: p = 2u - 1, accepted if |p|^2 <= 1, x = c + r p
 *                       (SnowBall :119-178, SandBall :180-238)
 *   MPM_SHAPE_BLOCK     a = {min x, y, z, max x, y, z, hole radius}: x = min + (max - min) u,
 *                       rejected if |x - (min x, min y, (min z + max z)/2)| < hole radius; the hole
 *                       diameter must be below max z - min z (SandBlock :240-323)
 *   MPM_SHAPE_CYLINDER  a = {base x, y, z, r, height, axis}: d = (2u_0 - 1, 2u_1 - 1) accepted if
 *                       |d|^2 <= 1; x = base + r d on the two other axes in ascending order and
 *                       base + height u_2 along `axis` (2 = z as in SandCylinder :325-401)
 * Fails with MPM_ERR_INVALID on bad arguments or when 64 n + 4096 candidates do not fill n. */
#define MPM_SHAPE_BALL 0
#define MPM_SHAPE_BLOCK 1
#define MPM_SHAPE_CYLINDER 2
int mpm_generate_shape(uint64_t seed, int32_t shape, const double* a, int64_t n, float* x);

#ifdef __cplusplus
}
#endif
#endif /* MPM_C_ABI_H */
