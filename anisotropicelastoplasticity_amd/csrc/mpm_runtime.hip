// mpm_runtime.hip -- context, device buffers, stage sequencing, HIP-graph replay and the C ABI of
// libmpm_hip.so (declared in include/mpm_c_abi.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "../../include/mpm_c_abi.h"
#include "mpm_kernels.h"
#include "mpm_mlp.h"
#include "mpm_noalloc.h"

using namespace mpm;

static_assert(sizeof(mpm_params) == 100 && sizeof(mpm_levelset) == 36, "C ABI struct layout");
static_assert(sizeof(mpm_scene_desc) == 48 + 16 + sizeof(mpm_params) + 4 + 16 * 36, "C ABI struct layout");

namespace {
thread_local std::string g_create_error;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};
}  // namespace

// Host description of one scene (grid geometry, parameters, level sets, particle count).
struct SceneHost {
  double gmin[3] = {0, 0, 0}, gmax[3] = {1, 1, 1}, h[3] = {1, 1, 1};
  int res[3] = {0, 0, 0};
  int tiles[3] = {0, 0, 0};
  int64_t ng = 0, ntiles = 0;
  mpm_params params{};
  std::vector<mpm_levelset> ls;
  int n = 0;  // all particles of the scene, mesh vertices and elements included
  // cloth: caller ids [mesh_v0, mesh_e0) are vertices, [mesh_e0, mesh_e1) elements (none: all equal)
  int mesh_v0 = INT32_MAX, mesh_e0 = INT32_MAX, mesh_e1 = INT32_MAX;
  float c_mu = 0, c_lambda = 0, c_gamma = 0, c_k = 0, c_cf = 0;
};

// Host copy of the cloth of a single-scene context (mpm_set_mesh).
struct MeshHost {
  int np = 0, nv = 0, ne = 0;  // particles before the mesh, vertices, elements
  int* faces = nullptr;        // device: [e][3] vertex caller ids
  float* rest = nullptr;       // device: [e][9]
  int* vcsr_off = nullptr;     // device: [nv+1]
  int* vcsr = nullptr;         // device: [3 ne]
  float* elem_fv = nullptr;    // device: [e][9]
  float* elem_stress = nullptr;
  float* vert_force = nullptr; // device: [v][3]
  int* fixed = nullptr;        // device: fixed vertex caller ids
  int nfixed = 0;
};

struct mpm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // scenes: one for the single-scene calls (mpm_set_grid / _params / _levelset), S for mpm_set_scenes
  bool have_grid = false;
  bool batch = false;
  std::vector<SceneHost> scenes;
  int64_t ng = 0;      // nodes (= cells) of all scenes
  int64_t ntiles = 0;  // tiles of all scenes
  int order = 3, predicted = 0, any_sand = 0;
  // device scene table (SceneTable points into it)
  unsigned char* table_buf = nullptr;
  SceneTable T{};
  int nblocks_p = 0;    // 256-particle blocks over all scenes
  int sort_passes = 3;  // 8-bit digit passes covering the largest global cell key
  int sort_mode = SORT_COUNTING;  // SORT_COUNTING or SORT_RADIX (mpm_autotune / mpm_set_tuning)
  uint32_t* ks_sorted = nullptr;  // sorted keys / ids of the last sort (k1/v1 or k2/v2)
  int* vs_sorted = nullptr;

  float* thr = nullptr;      // per-axis cell thresholds
  float* grid = nullptr;     // m, mom(3), v(3), vbf(3)
  int* cell_start = nullptr;
  int* cell_count = nullptr;  // per-cell particle counts of the counting sort (zero between substeps)
  int* cell_end = nullptr;  // cell_start + 1 (not a separate allocation)
  // per scan block of cells: 1 where a cell holds a particle; two sets, used by alternate substeps
  // (each substep's count clears the other set, whose P2G has run)
  uint32_t* occ[2] = {nullptr, nullptr};
  int occ_par = 0;
  const uint32_t* occ_last = nullptr;  // the set of the last sort
  int* tile_slot = nullptr;
  int* active_tiles = nullptr;
  int* block_slot = nullptr;
  int* active_blocks = nullptr;
  float* partials = nullptr;
  int64_t partial_slots = 0;

  // particles (all kinds); n_api: the particles of mpm_set_particles (without a mesh's vertices
  // and elements), which mpm_get_particles / mpm_get_stress return
  int n = 0;
  int n_api = 0;
  MeshHost mesh;
  int64_t cap = 0;
  float* state[2] = {nullptr, nullptr};
  int cur = 0;
  int sort_buf = 0;
  uint32_t* keys = nullptr;  // next cell key by caller id
  int* pos[2] = {nullptr, nullptr};  // caller id -> physical index in state[b]
  uint32_t *k1 = nullptr, *k2 = nullptr;
  int *v1 = nullptr, *v2 = nullptr;
  int* src = nullptr;  // sorted position -> physical index of the particle in state[cur]
  uint32_t* sortws = nullptr;
  int64_t sortws_words = 0;
  int nb_sort = 0;
  bool keys_valid = false;
  // pos[cur] maps every caller id to its slot in state[cur]. G2P maintains it only for contexts with a
  // cloth mesh (the cloth kernels address vertices and elements by caller id); otherwise it is rebuilt
  // from the state's id plane when something needs it (ensure_pos).
  bool pos_ok = false;
  int* tslot = nullptr;  // counting-sort scratch: each binned id's physical slot (n)
  bool stress_valid = false;  // PK planes of state[cur] hold V tau of the current F

  // counters and statistics
  int* counters = nullptr;
  int* status = nullptr;
  unsigned long long* fallbacks = nullptr;  // [0] svd, [1] apex, [2] node count scratch
  int64_t substeps = 0;
  int next_stage = MPM_STAGE_SORT;

  // graph of two substeps, per starting buffer
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  float graph_dt[2] = {0.0f, 0.0f};

  int p2g_grid = 512;
  int p2g_ranks = 8;  // P2G staging chunk (mpm_autotune / mpm_set_tuning)
  int p2g_pass = P2G_ROWS;  // P2G schedule (mpm_autotune / mpm_set_tuning)
  // adaptive time step (mpm_solve): the device clock, and the pointer the kernels get (null for
  // fixed-step runs)
  Clock* clock = nullptr;
  Clock* clk_active = nullptr;
  hipGraphExec_t graph_adapt[2] = {nullptr, nullptr};
  int grid_grid = 2048;
  int n_cu = 256;

  // learned return-map surrogate
  float* mlp_master = nullptr;
  float* mlp_m1 = nullptr;
  float* mlp_m2 = nullptr;
  float* mlp_grad = nullptr;
  float* scratch = nullptr;  // the fixed scratch of the allocation-free entry point (this device's address)
  uint16_t* mlp_wbf = nullptr;
  uint16_t* mlp_wbft = nullptr;
  int mlp_step = 0;
  bool mlp_ready = false;
  float* pair_log = nullptr;
  int64_t pair_cap = 0;
  // per G2P wave (4 per particle block): sorted position of lane 0, and the lanes that deferred to
  // the surrogate or to the FP64 apex pass; the lists are built from them without atomics
  int* wave_pos = nullptr;
  uint64_t* cand_mask = nullptr;
  uint64_t* apex_mask = nullptr;
  int* list_gsum = nullptr;  // per list and group of 64 waves: count, then offset
  int64_t wave_cap = 0;
  float* dec_log = nullptr;  // per-candidate gate records of the last surrogate substep
  int64_t dec_cap = 0;
  unsigned long long* sur_counts = nullptr;  // [0] candidates, [1] accepted, [2] pairs logged
  int sur_enabled = 0;
  float gate_tol = 1e-5f;
  float* cand = nullptr;
  int* cand_slot = nullptr;
  int* cand_idx = nullptr;  // the listed candidates' record indices in cand
  // set by mpm_time_kernels: the start and end of a stage's main kernel (taken by its launch)
  hipEvent_t ev_main[2] = {nullptr, nullptr};
  hipEvent_t ev_sur[2] = {nullptr, nullptr};  // the surrogate kernel's start and stop (timing only)
  int* apex_slot = nullptr;  // sorted slots of the deferred FP64 apex projections (n)
  int* apex_idx = nullptr;    // their record indices in apex_usv (n)
  float* apex_usv = nullptr;  // their trial SVD (CAND_SVD_PLANES x n)
};

static void drop_mesh(mpm_ctx* ctx);

#define CK(call)                                                                     \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                  \
      return MPM_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

static int fail(mpm_ctx* ctx, int code, const std::string& msg) {
  ctx->err = msg;
  return code;
}

template <class T>
static int dalloc(mpm_ctx* ctx, T*& p, size_t count) {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
  }
  if (count == 0) return MPM_OK;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess) return fail(ctx, MPM_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  p = static_cast<T*>(q);
  return MPM_OK;
}

template <class T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

static void drop_graphs(mpm_ctx* ctx) {
  for (int i = 0; i < 2; ++i) {
    if (ctx->graph[i]) {
      (void)hipGraphExecDestroy(ctx->graph[i]);
      ctx->graph[i] = nullptr;
    }
    if (ctx->graph_adapt[i]) {
      (void)hipGraphExecDestroy(ctx->graph_adapt[i]);
      ctx->graph_adapt[i] = nullptr;
    }
  }
}

static SceneHost& scene0(mpm_ctx* ctx) {
  if (ctx->scenes.empty()) {
    ctx->scenes.resize(1);
    mpm_params_default(&ctx->scenes[0].params);
  }
  return ctx->scenes[0];
}

// ------------------------------------------------------------------ parameters
// Multiplier m and shift l of the exact unsigned division by d >= 1 (Granlund and Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994): with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1, floor(n / d) = (t + ((n - t) >> 1)) >> (l - 1) for every 32-bit
// n, where t = floor(n m / 2^32). d = 1 is marked by l = 0.
static void div_magic(uint32_t d, uint32_t& m, int& l) {
  l = 0;
  while ((1ull << l) < d) ++l;
  m = l == 0 ? 0u : (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
}

static DevParams devparams_of(const SceneHost& sc) {
  DevParams d{};
  const mpm_params& p = sc.params;
  d.order = p.kernel_order == MPM_KERNEL_QUADRATIC ? 2 : 3;
  d.material = p.material;
  d.predicted = p.stress_at_predicted_F;
  d.friction_model = p.friction_model;
  for (int a = 0; a < 3; ++a) d.gravity[a] = p.gravity[a];
  double E = p.youngs_modulus, nu = p.poisson_ratio;
  d.lambda0 = float(E * nu / (1.0 + nu) / (1.0 - 2.0 * nu));
  d.mu0 = float(E / 2.0 / (1.0 + nu));
  d.theta_c = p.critical_compression;
  d.theta_s = p.critical_stretch;
  d.xi = p.hardening;
  d.h0 = p.dp_h[0];
  d.h1 = p.dp_h[1];
  d.h2 = p.dp_h[2];
  d.h3 = p.dp_h[3];
  d.per_particle_phi = p.per_particle_phi;
  d.alpha_f = -1.0f;
  d.alpha_d = -1.0;
  if (p.dp_h[1] == 0.0f && p.dp_h[3] == 0.0f && !p.per_particle_phi) {
    // phi(q) = h0 + (0 q - 0) exp(-h2 q) = h0 for every q (HybridSolver.cpp:646-647)
    const double sf = std::sin(double(p.dp_h[0]) * 3.14159265358979323846 / 180.0);
    d.alpha_d = std::sqrt(2.0 / 3.0) * 2.0 * sf / (3.0 - sf);
    d.alpha_f = float(d.alpha_d);
  }
  d.friction = p.friction;
  d.damping = p.apic_damping;
  d.svd_ratio = p.fp64_svd_ratio;
  d.apex_radius = p.fp64_apex_radius;
  double hmin = std::min(sc.h[0], std::min(sc.h[1], sc.h[2]));
  d.dinv = float((d.order == 3 ? 3.0 : 4.0) / hmin / hmin);
  for (int a = 0; a < 3; ++a) {
    d.gmin[a] = float(sc.gmin[a]);
    d.h[a] = float(sc.h[a]);
    d.inv_h[a] = float(1.0 / sc.h[a]);
    d.res[a] = sc.res[a];
    d.tiles[a] = sc.tiles[a];
  }
  for (int a = 0; a < 2; ++a) div_magic((uint32_t)sc.res[a], d.div_m[a], d.div_s[a]);
  d.nls = (int)sc.ls.size();
  for (int k = 0; k < d.nls; ++k) d.ls[k] = sc.ls[k];
  d.mesh_v0 = sc.mesh_v0;
  d.mesh_e0 = sc.mesh_e0;
  d.mesh_e1 = sc.mesh_e1;
  d.c_mu = sc.c_mu;
  d.c_lambda = sc.c_lambda;
  d.c_gamma = sc.c_gamma;
  d.c_k = sc.c_k;
  d.c_cf = sc.c_cf;
  return d;
}

// Scene table on the device: per-scene DevParams and the offset tables of SceneTable.
static int alloc_wave_tables(mpm_ctx* ctx);

static int upload_table(mpm_ctx* ctx) {
  const int S = (int)ctx->scenes.size();
  if (S == 0) return MPM_OK;
  for (int k = 1; k < S; ++k) {
    if (ctx->scenes[k].params.kernel_order != ctx->scenes[0].params.kernel_order ||
        ctx->scenes[k].params.stress_at_predicted_F != ctx->scenes[0].params.stress_at_predicted_F)
      return fail(ctx, MPM_ERR_UNSUPPORTED, "the scenes of one context share kernel_order and stress_at_predicted_F");
  }
  std::vector<DevParams> dp(S);
  std::vector<int> part_off(S + 1, 0), pblock_off(S + 1, 0), tile_off(S + 1, 0), thr_off(S, 0);
  std::vector<int64_t> node_off(S + 1, 0);
  int thr = 0;
  ctx->any_sand = 0;
  for (int k = 0; k < S; ++k) {
    const SceneHost& sc = ctx->scenes[k];
    dp[k] = devparams_of(sc);
    ctx->any_sand |= sc.params.material == MPM_SAND;
    part_off[k + 1] = part_off[k] + sc.n;
    pblock_off[k + 1] = pblock_off[k] + (sc.n + 255) / 256;
    tile_off[k + 1] = tile_off[k] + (int)sc.ntiles;
    node_off[k + 1] = node_off[k] + sc.ng;
    thr_off[k] = thr;
    thr += sc.res[0] + sc.res[1] + sc.res[2] + 3;
  }
  ctx->order = dp[0].order;
  ctx->predicted = dp[0].predicted;
  ctx->nblocks_p = pblock_off[S];
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t o_dp = 0, o_part = al(o_dp + sizeof(DevParams) * S), o_pb = al(o_part + sizeof(int) * (S + 1));
  const size_t o_tile = al(o_pb + sizeof(int) * (S + 1)), o_node = al(o_tile + sizeof(int) * (S + 1));
  const size_t o_thr = al(o_node + sizeof(int64_t) * (S + 1)), total = al(o_thr + sizeof(int) * S);
  std::vector<unsigned char> h(total, 0);
  std::memcpy(h.data() + o_dp, dp.data(), sizeof(DevParams) * S);
  std::memcpy(h.data() + o_part, part_off.data(), sizeof(int) * (S + 1));
  std::memcpy(h.data() + o_pb, pblock_off.data(), sizeof(int) * (S + 1));
  std::memcpy(h.data() + o_tile, tile_off.data(), sizeof(int) * (S + 1));
  std::memcpy(h.data() + o_node, node_off.data(), sizeof(int64_t) * (S + 1));
  std::memcpy(h.data() + o_thr, thr_off.data(), sizeof(int) * S);
  CK(hipStreamSynchronize(ctx->stream));
  int rc = dalloc(ctx, ctx->table_buf, total);
  if (rc) return rc;
  CK(hipMemcpy(ctx->table_buf, h.data(), total, hipMemcpyHostToDevice));
  SceneTable& T = ctx->T;
  T.nscenes = S;
  T.params = reinterpret_cast<const DevParams*>(ctx->table_buf + o_dp);
  T.part_off = reinterpret_cast<const int*>(ctx->table_buf + o_part);
  T.pblock_off = reinterpret_cast<const int*>(ctx->table_buf + o_pb);
  T.tile_off = reinterpret_cast<const int*>(ctx->table_buf + o_tile);
  T.node_off = reinterpret_cast<const int64_t*>(ctx->table_buf + o_node);
  T.thr_off = reinterpret_cast<const int*>(ctx->table_buf + o_thr);
  ctx->stress_valid = false;
  drop_graphs(ctx);
  if (ctx->wave_cap < 4 * (int64_t)ctx->nblocks_p) return alloc_wave_tables(ctx);
  return MPM_OK;
}

// smallest float f with (double(f) - gmin)/h >= c, evaluated exactly as HybridSolver.cpp:33-36
static float threshold(double gmin, double h, int c) {
  auto ok = [&](float f) { return (static_cast<double>(f) - gmin) / h >= static_cast<double>(c); };
  float f = static_cast<float>(gmin + c * h);
  while (ok(f)) f = std::nextafter(f, -INFINITY);
  while (!ok(f)) f = std::nextafter(f, INFINITY);
  return f;
}

// Sort workspace words: 4 x 256 digit histograms, 4 digit-pass tickets and 4 cell-scan tickets, the
// cell-scan look-back words (64-bit, one per scan block), then the digit passes' look-back words.
static constexpr int64_t SORTWS_TICKETS = 1024, SORTWS_SCAN = 1032;
// [SORTWS_SCAN, radix status): the scan blocks' totals
static int64_t scan_blocks(const mpm_ctx* ctx) { return (ctx->ng + SCAN_TILE - 1) / SCAN_TILE; }
static int64_t sortws_radix_status(const mpm_ctx* ctx) { return SORTWS_SCAN + scan_blocks(ctx) + 1; }

// ------------------------------------------------------------------ substep stages
// Rebuild pos[cur] from the id plane of state[cur] if G2P did not maintain it.
static void ensure_pos(mpm_ctx* ctx) {
  if (ctx->pos_ok || ctx->n <= 0) return;
  launch_pos_of_ids(ctx->stream, reinterpret_cast<const int*>(ctx->state[ctx->cur] + (size_t)PID * ctx->cap), ctx->n,
                    ctx->pos[ctx->cur]);
  ctx->pos_ok = true;
}

static int stage_sort(mpm_ctx* ctx) {
  hipStream_t s = ctx->stream;
  float* st = ctx->state[ctx->cur];
  if (ctx->mesh.ne > 0 || ctx->mesh.nv > 0) ensure_pos(ctx);  // the cloth kernels address by caller id
  if (!ctx->keys_valid) launch_keys(s, ctx->T, ctx->nblocks_p, st, ctx->cap, ctx->n, ctx->thr, ctx->keys, ctx->status);
  ResetArgs r;
  if (ctx->sort_mode == SORT_COUNTING) {  // nothing to zero: the scan totals are overwritten
    r.zero_words = ctx->sortws + SORTWS_TICKETS;
    r.n_zero = 0;
  } else {
    r.zero_words = ctx->sortws;
    r.n_zero = ctx->sortws_words;
  }
  r.counters = ctx->counters;
  r.active_tiles = ctx->active_tiles;
  r.tile_slot = ctx->tile_slot;
  r.active_blocks = ctx->active_blocks;
  r.block_slot = ctx->block_slot;
  r.max_tiles = ctx->ntiles;
  r.clk = ctx->clk_active;
  // the counting sort's first kernel does the reset work itself (one launch fewer)
  if (ctx->sort_mode != SORT_COUNTING) launch_reset(s, r);
  SortBufs b;
  b.reset = r;
  b.k1 = ctx->k1;
  b.k2 = ctx->k2;
  b.v1 = ctx->v1;
  b.v2 = ctx->v2;
  b.hist = ctx->sortws;
  b.tickets = ctx->sortws + SORTWS_TICKETS;
  b.scan_agg = ctx->sortws + SORTWS_SCAN;
  b.status = ctx->sortws + sortws_radix_status(ctx);
  b.cell_count = ctx->cell_count;
  b.cell_start = ctx->cell_start;
  b.ncells = ctx->ng;
  b.T = ctx->T;
  b.tile_slot = ctx->tile_slot;
  b.pid = reinterpret_cast<const int*>(st + (size_t)PID * ctx->cap);
  b.tslot = ctx->tslot;
  b.pos = ctx->pos[ctx->cur];
  b.src = ctx->src;
  b.counters = ctx->counters;
  b.clk = ctx->clk_active;
  b.occ = ctx->occ[ctx->occ_par];
  b.occ_clear = ctx->occ[1 - ctx->occ_par];
  ctx->occ_last = b.occ;
  ctx->occ_par = 1 - ctx->occ_par;
  launch_sort(s, b, ctx->keys, ctx->n, ctx->sort_passes, ctx->sort_mode);
  if (ctx->sort_mode != SORT_COUNTING) ctx->pos_ok = true;  // the radix mode's key conversion wrote pos[cur]
  launch_bounds(s, ctx->T, ctx->nblocks_p, ctx->ks_sorted, ctx->n, ctx->ntiles, ctx->cell_count,
                ctx->tile_slot, ctx->active_tiles, ctx->block_slot, ctx->active_blocks, ctx->counters, ctx->vs_sorted,
                ctx->pos[ctx->cur], ctx->src, ctx->clk_active, ctx->sort_mode);
  ctx->sort_buf = ctx->cur;
  return MPM_OK;
}

// Cloth kernel arguments on buffer b (state[b] with its caller-id map pos[b]).
static ClothArgs cloth_args(mpm_ctx* ctx, int b) {
  ClothArgs c{};
  const MeshHost& m = ctx->mesh;
  c.P = devparams_of(ctx->scenes[0]);
  c.st = ctx->state[b];
  c.cap = ctx->cap;
  c.n = ctx->n;
  c.pos = ctx->pos[b];
  c.nv = m.nv;
  c.ne = m.ne;
  c.faces = m.faces;
  c.rest = m.rest;
  c.vcsr_off = m.vcsr_off;
  c.vcsr = m.vcsr;
  c.elem_fv = m.elem_fv;
  c.elem_stress = m.elem_stress;
  c.vert_force = m.vert_force;
  c.fixed = m.fixed;
  c.nfixed = m.nfixed;
  c.gvel = reinterpret_cast<float4*>(ctx->grid + GRID_VEL * ctx->ng);
  c.keys = ctx->keys;
  c.thr = ctx->thr;
  c.status = ctx->status;
  c.fallback = ctx->fallbacks;
  c.clk = ctx->clk_active;
  return c;
}

static P2GArgs p2g_args(mpm_ctx* ctx, float dt) {
  P2GArgs a;
  a.T = ctx->T;
  a.state = ctx->state[ctx->cur];
  a.cap = ctx->cap;
  a.n = ctx->n;
  a.src = ctx->src;
  a.cell_start = ctx->cell_start;
  a.cell_end = ctx->cell_end;
  a.occ = ctx->occ_last;
  a.counters = ctx->counters;
  a.active_tiles = ctx->active_tiles;
  a.partials = ctx->partials;
  a.gvel = reinterpret_cast<const float4*>(ctx->grid + GRID_VEL * ctx->ng);
  a.dt = dt;
  a.fallback_svd = ctx->fallbacks;
  a.clk = ctx->clk_active;
  a.vert_force = ctx->mesh.vert_force;
  a.elem_stress = ctx->mesh.elem_stress;
  return a;
}

static GridArgs grid_args(mpm_ctx* ctx, float dt) {
  GridArgs g;
  g.T = ctx->T;
  g.counters = ctx->counters;
  g.active_blocks = ctx->active_blocks;
  g.tile_slot = ctx->tile_slot;
  g.partials = ctx->partials;
  g.gm = ctx->grid + GRID_M * ctx->ng;
  g.gmom = ctx->grid + GRID_MOM * ctx->ng;
  g.gvel = reinterpret_cast<float4*>(ctx->grid + GRID_VEL * ctx->ng);
  g.ng = ctx->ng;
  g.dt = dt;
  g.clk = ctx->clk_active;
  g.vmax_out = nullptr;
  return g;
}

static int stage_p2g(mpm_ctx* ctx, float dt) {
  hipStream_t s = ctx->stream;
  P2GArgs a = p2g_args(ctx, dt);
  // cloth forces from the current directions (HybridSolver.cpp:370-455), read by the P2G staging
  if (ctx->mesh.ne > 0) launch_cloth_forces(s, cloth_args(ctx, ctx->cur));
  if (!ctx->predicted) {
    if (!ctx->stress_valid) {
      launch_stress_init(s, ctx->T, ctx->nblocks_p, ctx->state[ctx->cur], ctx->cap, ctx->n, ctx->fallbacks);
      ctx->stress_valid = true;
    }
    launch_p2g(s, a, ctx->order, P2G_MOM_STRESS, ctx->p2g_grid, ctx->p2g_ranks, ctx->p2g_pass);
  } else {
    launch_p2g(s, a, ctx->order, P2G_MOM, ctx->p2g_grid, ctx->p2g_ranks, ctx->p2g_pass);
    launch_grid(s, grid_args(ctx, dt), GRID_NORMALIZE, ctx->grid_grid);
    launch_p2g(s, a, ctx->order, P2G_FORCE, ctx->p2g_grid, ctx->p2g_ranks, ctx->p2g_pass);
  }
  return MPM_OK;
}

static int stage_grid(mpm_ctx* ctx, float dt) {
  GridArgs g = grid_args(ctx, dt);
  if (ctx->clk_active) g.vmax_out = &ctx->clk_active->vmax_bits;
  launch_grid(ctx->stream, g, ctx->predicted ? GRID_FORCE_UPDATE : GRID_FUSED, ctx->grid_grid);
  // fixed cloth vertices stop the nodes around them (part of gridCollisionHandling_, :513-550)
  if (ctx->mesh.nfixed > 0) launch_cloth_fixed(ctx->stream, cloth_args(ctx, ctx->cur));
  // adaptive step: the new Dt from the updated velocities, used by G2P (HybridSolver.cpp:878)
  if (ctx->clk_active) launch_clock_update(ctx->stream, ctx->clk_active, 0);
  return MPM_OK;
}

static int stage_g2p(mpm_ctx* ctx, float dt) {
  G2PArgs a;
  a.T = ctx->T;
  a.sin = ctx->state[ctx->cur];
  a.sout = ctx->state[1 - ctx->cur];
  a.cap = ctx->cap;
  a.n = ctx->n;
  a.perm = ctx->vs_sorted;
  a.src = ctx->src;
  a.ks = ctx->ks_sorted;
  a.pos_out = ctx->mesh.ne > 0 || ctx->mesh.nv > 0 ? ctx->pos[1 - ctx->cur] : nullptr;
  a.gvel = reinterpret_cast<const float4*>(ctx->grid + GRID_VEL * ctx->ng);
  a.dt = dt;
  a.write_stress = ctx->predicted ? 0 : 1;
  a.thr = ctx->thr;
  a.keys = ctx->keys;
  a.status = ctx->status;
  a.fallback_svd = ctx->fallbacks;
  a.fallback_apex = ctx->fallbacks + 1;
  const bool surrogate = ctx->sur_enabled && ctx->mlp_ready && ctx->cand && ctx->any_sand;
  if (ctx->wave_cap < 4 * (int64_t)ctx->nblocks_p)
    return fail(ctx, MPM_ERR_STATE, "per-wave tables smaller than the particle blocks");
  a.sur_mode = surrogate ? 2 : (ctx->pair_cap > 0 ? 1 : 0);
  a.pair_log = ctx->pair_log;
  a.pair_cap = ctx->pair_cap;
  a.pair_count = ctx->sur_counts ? ctx->sur_counts + 2 : nullptr;
  a.cand = ctx->cand;
  a.cand_mask = ctx->cand_mask;
  a.wave_pos = ctx->wave_pos;
  a.apex_mask = ctx->any_sand ? ctx->apex_mask : nullptr;
  a.apex_usv = ctx->apex_usv;
  a.clk = ctx->clk_active;
  launch_g2p(ctx->stream, a, ctx->order, ctx->nblocks_p, ctx->ev_main[0], ctx->ev_main[1]);
  {
    ListArgs la{};
    la.wpos = ctx->wave_pos;
    la.gsum = ctx->list_gsum;
    la.nwaves = 4 * ctx->nblocks_p;
    if (a.apex_mask) {
      la.mask[la.nlists] = ctx->apex_mask;
      la.list[la.nlists] = ctx->apex_slot;
      la.idx[la.nlists] = ctx->apex_idx;
      la.count[la.nlists++] = ctx->counters + CNT_APEX;
    }
    if (surrogate) {
      la.mask[la.nlists] = ctx->cand_mask;
      la.list[la.nlists] = ctx->cand_slot;
      la.idx[la.nlists] = ctx->cand_idx;
      la.count[la.nlists++] = ctx->counters + CNT_CAND;
    }
    launch_position_lists(ctx->stream, la);
  }
  if (a.apex_mask) {
    ApexArgs x;
    x.T = ctx->T;
    x.st = ctx->state[1 - ctx->cur];
    x.cap = ctx->cap;
    x.n = ctx->n;
    x.apex_slot = ctx->apex_slot;
    x.apex_idx = ctx->apex_idx;
    x.apex_usv = ctx->apex_usv;
    x.apex_count = ctx->counters + CNT_APEX;
    x.write_stress = a.write_stress;
    x.log_pairs = a.sur_mode == 1;
    x.pair_log = ctx->pair_log;
    x.pair_cap = ctx->pair_cap;
    x.pair_count = a.pair_count;
    x.fallback_svd = ctx->fallbacks;
    x.fallback_apex = ctx->fallbacks + 1;
    x.clk = ctx->clk_active;
    launch_apex_finish(ctx->stream, x, 2 * ctx->n_cu);
  }
  if (surrogate) {
    SurrogateArgs sa;
    sa.T = ctx->T;
    sa.wbf = ctx->mlp_wbf;
    sa.master = ctx->mlp_master;
    sa.st = ctx->state[1 - ctx->cur];
    sa.cap = ctx->cap;
    sa.n = ctx->n;
    sa.cand = ctx->cand;
    sa.cand_slot = ctx->cand_slot;
    sa.cand_idx = ctx->cand_idx;
    sa.cand_count = ctx->counters + CNT_CAND;
    sa.gate_tol = ctx->gate_tol;
    sa.write_stress = a.write_stress;
    sa.accepted = ctx->sur_counts + 1;
    sa.candidates = ctx->sur_counts;
    sa.dec_log = ctx->dec_log;
    sa.dec_cap = ctx->dec_cap;
    sa.clk = ctx->clk_active;
    launch_surrogate(ctx->stream, sa, ctx->n_cu, ctx->ev_sur);
  }
  // cloth elements follow their vertices (positions, velocities, d1 d2) and are return-mapped
  if (ctx->mesh.ne > 0) launch_cloth_elements(ctx->stream, cloth_args(ctx, 1 - ctx->cur));
  if (ctx->clk_active) launch_clock_end(ctx->stream, ctx->clk_active);
  ctx->cur = 1 - ctx->cur;
  ctx->keys_valid = true;
  ctx->stress_valid = !ctx->predicted;
  ctx->pos_ok = a.pos_out != nullptr;
  ctx->substeps += 1;
  return MPM_OK;
}

static int run_stage(mpm_ctx* ctx, int stage, float dt) {
  switch (stage) {
    case MPM_STAGE_SORT: return stage_sort(ctx);
    case MPM_STAGE_P2G: return stage_p2g(ctx, dt);
    case MPM_STAGE_GRID: return stage_grid(ctx, dt);
    case MPM_STAGE_G2P: return stage_g2p(ctx, dt);
  }
  return fail(ctx, MPM_ERR_INVALID, "unknown stage");
}

static int ready(mpm_ctx* ctx) {
  if (!ctx) return MPM_ERR_INVALID;
  if (!ctx->have_grid) return fail(ctx, MPM_ERR_STATE, "mpm_set_grid has not been called");
  if (ctx->n <= 0) return fail(ctx, MPM_ERR_STATE, "mpm_set_particles has not been called");
  int64_t total = 0;
  for (const SceneHost& sc : ctx->scenes) total += sc.n;
  if (total != ctx->n) return fail(ctx, MPM_ERR_STATE, "mpm_set_particles must follow mpm_set_scenes");
  return MPM_OK;
}

static int one_substep(mpm_ctx* ctx, float dt) {
  for (int st = MPM_STAGE_SORT; st <= MPM_STAGE_G2P; ++st) {
    int rc = run_stage(ctx, st, dt);
    if (rc) return rc;
  }
  CK(hipGetLastError());
  return MPM_OK;
}

// ------------------------------------------------------------------ C ABI
extern "C" {

void mpm_params_default(mpm_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->kernel_order = MPM_KERNEL_CUBIC;
  p->material = MPM_SAND;
  p->stress_at_predicted_F = 1;
  p->friction_model = MPM_FRICTION_REFERENCE;
  p->gravity[0] = 0.0f;
  p->gravity[1] = 0.0f;
  p->gravity[2] = -9.8f;
  p->youngs_modulus = 1.4e5f;
  p->poisson_ratio = 0.2f;
  p->critical_compression = 2.5e-2f;
  p->critical_stretch = 7.5e-3f;
  p->hardening = 10.0f;
  p->dp_h[0] = 35.0f;
  p->dp_h[1] = 9.0f;
  p->dp_h[2] = 0.2f;
  p->dp_h[3] = 10.0f;
  p->per_particle_phi = 0;
  p->friction = 0.2f;
  p->apic_damping = 0.0f;
  p->fp64_svd_ratio = 1e-3f;
  p->fp64_apex_radius = 1e-4f;
}

int mpm_create(mpm_ctx** out, int32_t device) {
  if (!out) return MPM_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    g_create_error = "no HIP device available";
    return MPM_ERR_HIP;
  }
  if (device < 0 || device >= count) {
    g_create_error = "device index out of range";
    return MPM_ERR_INVALID;
  }
  mpm_ctx* ctx = new (std::nothrow) mpm_ctx();
  if (!ctx) return MPM_ERR_NOMEM;
  ctx->device = device;
  e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_create_error = std::string("stream creation failed: ") + hipGetErrorString(e);
    delete ctx;
    return MPM_ERR_HIP;
  }
  e = noalloc_scratch(device, &ctx->scratch);
  if (e != hipSuccess) {
    g_create_error = std::string("scratch address: ") + hipGetErrorString(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MPM_ERR_HIP;
  }
  scene0(ctx);
  int rc = dalloc(ctx, ctx->counters, CNT_NUM);
  if (!rc) rc = dalloc(ctx, ctx->status, 1);
  if (!rc) rc = dalloc(ctx, ctx->fallbacks, 4);
  if (!rc) rc = dalloc(ctx, ctx->sur_counts, 4);
  if (rc) {
    g_create_error = ctx->err;
    mpm_destroy(ctx);
    return rc;
  }
  (void)hipMemsetAsync(ctx->counters, 0, CNT_NUM * sizeof(int), ctx->stream);
  (void)hipMemsetAsync(ctx->status, 0, sizeof(int), ctx->stream);
  (void)hipMemsetAsync(ctx->fallbacks, 0, 4 * sizeof(unsigned long long), ctx->stream);
  (void)hipMemsetAsync(ctx->sur_counts, 0, 4 * sizeof(unsigned long long), ctx->stream);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ctx->p2g_grid = std::max(128, prop.multiProcessorCount * 2);  // 8-wave workgroups, two per CU
    ctx->n_cu = std::max(1, prop.multiProcessorCount);
    // 16 one-wave workgroups per CU: the autotuner's choice at configs 2 and 4 (14.2 -> 12.7 us and
    // 48.9 -> 41.9 us against 8 per CU)
    ctx->grid_grid = std::max(512, prop.multiProcessorCount * 16);
  }
  *out = ctx;
  return MPM_OK;
}

int mpm_destroy(mpm_ctx* ctx) {
  if (!ctx) return MPM_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graphs(ctx);
  dfree(ctx->table_buf);
  drop_mesh(ctx);
  dfree(ctx->thr); dfree(ctx->grid); dfree(ctx->cell_start); dfree(ctx->cell_count); dfree(ctx->occ[0]); dfree(ctx->occ[1]);
  ctx->cell_end = nullptr;
  dfree(ctx->tile_slot); dfree(ctx->active_tiles); dfree(ctx->block_slot); dfree(ctx->active_blocks);
  dfree(ctx->partials);
  dfree(ctx->state[0]); dfree(ctx->state[1]); dfree(ctx->keys); dfree(ctx->pos[0]); dfree(ctx->pos[1]); dfree(ctx->tslot);
  dfree(ctx->k1); dfree(ctx->k2); dfree(ctx->v1); dfree(ctx->v2); dfree(ctx->src); dfree(ctx->sortws); dfree(ctx->apex_slot); dfree(ctx->apex_idx); dfree(ctx->apex_usv);
  dfree(ctx->counters); dfree(ctx->status); dfree(ctx->fallbacks);
  dfree(ctx->clock);
  dfree(ctx->mlp_master); dfree(ctx->mlp_m1); dfree(ctx->mlp_m2); dfree(ctx->mlp_grad); dfree(ctx->mlp_wbf);
  dfree(ctx->mlp_wbft); dfree(ctx->pair_log); dfree(ctx->dec_log); dfree(ctx->sur_counts); dfree(ctx->cand); dfree(ctx->cand_slot); dfree(ctx->cand_idx); dfree(ctx->cand_mask); dfree(ctx->wave_pos); dfree(ctx->apex_mask); dfree(ctx->list_gsum);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MPM_OK;
}

const char* mpm_last_error(const mpm_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

static int alloc_partials(mpm_ctx* ctx) {
  if (!ctx->have_grid || ctx->n <= 0) return MPM_OK;
  int64_t slots = std::min<int64_t>(ctx->ntiles, ctx->n);
  if (slots == ctx->partial_slots && ctx->partials) return MPM_OK;
  int rc = dalloc(ctx, ctx->partials, (size_t)slots * PART_STRIDE);
  if (rc) return rc;
  ctx->partial_slots = slots;
  return MPM_OK;
}

// Radix-sort workspace: histograms, tickets and look-back words of every pass.
static int alloc_sortws(mpm_ctx* ctx) {
  if (ctx->n <= 0) return MPM_OK;
  ctx->nb_sort = (ctx->n + SORT_TILE - 1) / SORT_TILE;
  const int64_t words = sortws_radix_status(ctx) + (int64_t)ctx->sort_passes * ctx->nb_sort * 256;
  if (words != ctx->sortws_words || !ctx->sortws) {
    int rc = dalloc(ctx, ctx->sortws, words);
    if (rc) return rc;
    ctx->sortws_words = words;
  }
  CK(hipMemsetAsync(ctx->sortws, 0, sizeof(uint32_t) * ctx->sortws_words, ctx->stream));
  return MPM_OK;
}

// The last sort pass writes k1/v1 for an odd pass count, k2/v2 for an even one.
static void pick_sorted(mpm_ctx* ctx) {
  ctx->ks_sorted = (ctx->sort_passes & 1) ? ctx->k1 : ctx->k2;
  ctx->vs_sorted = (ctx->sort_passes & 1) ? ctx->v1 : ctx->v2;
}

static int set_geometry(mpm_ctx* ctx, SceneHost& sc, const double min_bound[3], const double max_bound[3],
                        const int32_t res[3]) {
  for (int a = 0; a < 3; ++a) {
    if (!(max_bound[a] > min_bound[a])) return fail(ctx, MPM_ERR_INVALID, "maxBound must be bigger than minBound");
    if (res[a] < 1) return fail(ctx, MPM_ERR_INVALID, "resolution must be positive");
  }
  // the kernels form a scene's cell and node indices in 32-bit signed arithmetic
  if ((int64_t)res[0] * res[1] * res[2] > (int64_t)INT32_MAX)
    return fail(ctx, MPM_ERR_UNSUPPORTED, "a scene has more than 2^31 - 1 cells");
  for (int a = 0; a < 3; ++a) {
    sc.gmin[a] = min_bound[a];
    sc.gmax[a] = max_bound[a];
    sc.res[a] = res[a];
    sc.h[a] = (max_bound[a] - min_bound[a]) / res[a];  // RegularGrid.cpp:144-146
    sc.tiles[a] = (res[a] + TILE - 1) / TILE;
  }
  sc.ng = (int64_t)res[0] * res[1] * res[2];
  sc.ntiles = (int64_t)sc.tiles[0] * sc.tiles[1] * sc.tiles[2];
  return MPM_OK;
}

static int check_params(mpm_ctx* ctx, const mpm_params* p) {
  if (p->kernel_order != MPM_KERNEL_QUADRATIC && p->kernel_order != MPM_KERNEL_CUBIC)
    return fail(ctx, MPM_ERR_INVALID, "kernel_order must be 2 or 3");
  if (p->material != MPM_SNOW && p->material != MPM_SAND) return fail(ctx, MPM_ERR_INVALID, "unknown material");
  if (p->friction_model < 0 || p->friction_model > 2) return fail(ctx, MPM_ERR_INVALID, "unknown friction model");
  if (!(p->poisson_ratio < 0.5f) || !(p->youngs_modulus > 0.0f)) return fail(ctx, MPM_ERR_INVALID, "bad elastic moduli");
  return MPM_OK;
}

static int check_levelsets(mpm_ctx* ctx, const mpm_levelset* prims, int32_t n) {
  if (n < 0 || n > MPM_MAX_LEVELSETS || (n > 0 && !prims)) return fail(ctx, MPM_ERR_INVALID, "0 <= n <= 16 primitives");
  for (int i = 0; i < n; ++i)
    if (prims[i].type < MPM_LS_PLANE || prims[i].type > MPM_LS_SPHERE) return fail(ctx, MPM_ERR_INVALID, "unknown level-set type");
  return MPM_OK;
}

// Grid buffers of all scenes, after their geometry changed: nodes, bins, tiles and cell thresholds
// concatenated scene-major.
static int layout(mpm_ctx* ctx) {
  int64_t ng = 0, nt = 0, nthr = 0;
  for (const SceneHost& sc : ctx->scenes) {
    ng += sc.ng;
    nt += sc.ntiles;
    nthr += sc.res[0] + sc.res[1] + sc.res[2] + 3;
  }
  if (ng > (int64_t)0xFFFFFFFFll) return fail(ctx, MPM_ERR_UNSUPPORTED, "more than 2^32 cells over all scenes");
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  ctx->ng = ng;
  ctx->ntiles = nt;
  int bits = 1;
  while ((int64_t(1) << bits) < ng) ++bits;  // keys lie in [0, ng)
  ctx->sort_passes = std::max(1, (bits + 7) / 8);
  int rc = dalloc(ctx, ctx->grid, (size_t)GRID_FLOATS * ng);
  if (!rc) rc = dalloc(ctx, ctx->cell_start, ng + 1);
  const int64_t nscan = (ng + SCAN_TILE - 1) / SCAN_TILE;
  if (!rc) rc = dalloc(ctx, ctx->occ[0], (size_t)nscan);
  if (!rc) rc = dalloc(ctx, ctx->occ[1], (size_t)nscan);
  if (!rc) rc = dalloc(ctx, ctx->cell_count, ng);
  if (!rc) rc = dalloc(ctx, ctx->tile_slot, nt);
  if (!rc) rc = dalloc(ctx, ctx->active_tiles, nt);
  if (!rc) rc = dalloc(ctx, ctx->block_slot, nt);
  if (!rc) rc = dalloc(ctx, ctx->active_blocks, nt);
  if (!rc) rc = dalloc(ctx, ctx->thr, (size_t)nthr);
  if (rc) return rc;
  CK(hipMemsetAsync(ctx->grid, 0, sizeof(float) * GRID_FLOATS * ng, ctx->stream));
  ctx->cell_end = ctx->cell_start + 1;  // cell c holds sorted positions [cell_start[c], cell_start[c + 1])
  CK(hipMemsetAsync(ctx->cell_start, 0, sizeof(int) * (ng + 1), ctx->stream));
  CK(hipMemsetAsync(ctx->occ[0], 0, sizeof(uint32_t) * nscan, ctx->stream));
  CK(hipMemsetAsync(ctx->occ[1], 0, sizeof(uint32_t) * nscan, ctx->stream));
  ctx->occ_last = ctx->occ[0];
  CK(hipMemsetAsync(ctx->cell_count, 0, sizeof(int) * ng, ctx->stream));
  launch_fill_int(ctx->stream, ctx->tile_slot, nt, -1);
  launch_fill_int(ctx->stream, ctx->block_slot, nt, -1);
  CK(hipMemsetAsync(ctx->counters, 0, CNT_NUM * sizeof(int), ctx->stream));
  std::vector<float> thr;
  thr.reserve(nthr);
  for (const SceneHost& sc : ctx->scenes)
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c <= sc.res[a]; ++c) thr.push_back(threshold(sc.gmin[a], sc.h[a], c));
  CK(hipMemcpyAsync(ctx->thr, thr.data(), thr.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  ctx->have_grid = true;
  ctx->keys_valid = false;
  ctx->partial_slots = 0;
  pick_sorted(ctx);
  if (ctx->k1) CK(hipMemsetAsync(ctx->k1, 0, sizeof(uint32_t) * ctx->cap, ctx->stream));
  if (ctx->k2) CK(hipMemsetAsync(ctx->k2, 0, sizeof(uint32_t) * ctx->cap, ctx->stream));
  rc = alloc_sortws(ctx);
  if (!rc) rc = alloc_partials(ctx);
  CK(hipStreamSynchronize(ctx->stream));
  return rc;
}

int mpm_set_grid(mpm_ctx* ctx, const double min_bound[3], const double max_bound[3], const int32_t res[3]) {
  if (!ctx || !min_bound || !max_bound || !res) return MPM_ERR_INVALID;
  if (ctx->batch) return fail(ctx, MPM_ERR_STATE, "the context holds several scenes (mpm_set_scenes)");
  SceneHost& sc = scene0(ctx);
  int rc = set_geometry(ctx, sc, min_bound, max_bound, res);
  if (!rc) rc = layout(ctx);
  if (!rc) rc = upload_table(ctx);
  return rc;
}

int mpm_set_params(mpm_ctx* ctx, const mpm_params* p) {
  if (!ctx || !p) return MPM_ERR_INVALID;
  if (ctx->batch) return fail(ctx, MPM_ERR_STATE, "the context holds several scenes (mpm_set_scenes)");
  int rc = check_params(ctx, p);
  if (rc) return rc;
  scene0(ctx).params = *p;
  return ctx->have_grid ? upload_table(ctx) : MPM_OK;
}

int mpm_set_levelset(mpm_ctx* ctx, const mpm_levelset* prims, int32_t n) {
  if (!ctx) return MPM_ERR_INVALID;
  if (ctx->batch) return fail(ctx, MPM_ERR_STATE, "the context holds several scenes (mpm_set_scenes)");
  int rc = check_levelsets(ctx, prims, n);
  if (rc) return rc;
  scene0(ctx).ls.assign(prims, prims + n);
  return ctx->have_grid ? upload_table(ctx) : MPM_OK;
}

int mpm_set_scenes(mpm_ctx* ctx, int32_t nscenes, const mpm_scene_desc* descs) {
  if (!ctx) return MPM_ERR_INVALID;
  if (nscenes < 1 || !descs) return fail(ctx, MPM_ERR_INVALID, "nscenes >= 1 scene descriptions are required");
  std::vector<SceneHost> sc(nscenes);
  for (int k = 0; k < nscenes; ++k) {
    const mpm_scene_desc& d = descs[k];
    int rc = set_geometry(ctx, sc[k], d.min_bound, d.max_bound, d.res);
    if (!rc) rc = check_params(ctx, &d.params);
    if (!rc) rc = check_levelsets(ctx, d.levelsets, d.n_levelsets);
    if (!rc && d.n_particles < 1) rc = fail(ctx, MPM_ERR_INVALID, "every scene needs particles");
    if (rc) return rc;
    sc[k].params = d.params;
    sc[k].ls.assign(d.levelsets, d.levelsets + d.n_levelsets);
    sc[k].n = d.n_particles;
  }
  ctx->scenes = std::move(sc);
  ctx->batch = true;
  int rc = layout(ctx);
  if (!rc) rc = upload_table(ctx);
  return rc;
}

// Per-wave tables of G2P's deferred lanes (positions, masks, group sums) for the current particle
// blocks.
static int alloc_wave_tables(mpm_ctx* ctx) {
  const int64_t nw = 4 * (int64_t)ctx->nblocks_p;
  int rc = dalloc(ctx, ctx->wave_pos, (size_t)nw);
  if (!rc) rc = dalloc(ctx, ctx->cand_mask, (size_t)nw);
  if (!rc) rc = dalloc(ctx, ctx->apex_mask, (size_t)nw);
  if (!rc) rc = dalloc(ctx, ctx->list_gsum, (size_t)(2 * ((nw + 63) / 64)));
  ctx->wave_cap = rc ? 0 : nw;
  return rc;
}

// Install n particles given as host planes (NPLANES x n, caller order) as the context's state.
static int install_state(mpm_ctx* ctx, int n, const std::vector<float>& h) {
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  if (n != ctx->n) {
    int rc = dalloc(ctx, ctx->state[0], (size_t)NPLANES * n);
    if (!rc) rc = dalloc(ctx, ctx->state[1], (size_t)NPLANES * n);
    if (!rc) rc = dalloc(ctx, ctx->keys, n);
    if (!rc) rc = dalloc(ctx, ctx->pos[0], n);
    if (!rc) rc = dalloc(ctx, ctx->pos[1], n);
    if (!rc) rc = dalloc(ctx, ctx->tslot, n);
    if (!rc) rc = dalloc(ctx, ctx->k1, n);
    if (!rc) rc = dalloc(ctx, ctx->k2, n);
    if (!rc) rc = dalloc(ctx, ctx->v1, n);
    if (!rc) rc = dalloc(ctx, ctx->v2, n);
    if (!rc) rc = dalloc(ctx, ctx->src, n);
    if (!rc) rc = dalloc(ctx, ctx->apex_slot, n);
    if (!rc) rc = dalloc(ctx, ctx->apex_idx, n);
    if (!rc) rc = dalloc(ctx, ctx->apex_usv, (size_t)CAND_SVD_PLANES * n);
    if (!rc && (ctx->sur_enabled || ctx->cand)) rc = dalloc(ctx, ctx->cand, (size_t)CAND_SVD_PLANES * n);
    if (!rc && (ctx->sur_enabled || ctx->cand_slot)) rc = dalloc(ctx, ctx->cand_slot, (size_t)n);
    if (!rc && (ctx->sur_enabled || ctx->cand_idx)) rc = dalloc(ctx, ctx->cand_idx, (size_t)n);
    if (!rc && ctx->wave_cap < 4 * (int64_t)ctx->nblocks_p) rc = alloc_wave_tables(ctx);
    if (rc) return rc;
    CK(hipMemsetAsync(ctx->state[1], 0, sizeof(float) * NPLANES * (size_t)n, ctx->stream));
    CK(hipMemsetAsync(ctx->keys, 0, sizeof(uint32_t) * n, ctx->stream));
    CK(hipMemsetAsync(ctx->v1, 0, sizeof(int) * n, ctx->stream));
    CK(hipMemsetAsync(ctx->v2, 0, sizeof(int) * n, ctx->stream));
    CK(hipMemsetAsync(ctx->src, 0, sizeof(int) * n, ctx->stream));
    CK(hipMemsetAsync(ctx->pos[1], 0, sizeof(int) * n, ctx->stream));
    ctx->n = n;
    ctx->cap = n;
    rc = alloc_sortws(ctx);
    if (!rc) rc = alloc_partials(ctx);
    if (rc) return rc;
  }
  pick_sorted(ctx);
  int rc = upload_table(ctx);
  if (rc) return rc;
  CK(hipMemcpyAsync(ctx->state[0], h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  CK(hipMemsetAsync(ctx->k1, 0, sizeof(uint32_t) * n, ctx->stream));
  CK(hipMemsetAsync(ctx->k2, 0, sizeof(uint32_t) * n, ctx->stream));
  launch_iota(ctx->stream, ctx->pos[0], n);
  ctx->pos_ok = true;
  CK(hipStreamSynchronize(ctx->stream));
  ctx->stress_valid = false;
  ctx->cur = 0;
  ctx->sort_buf = 0;
  ctx->keys_valid = false;
  ctx->substeps = 0;
  ctx->next_stage = MPM_STAGE_SORT;
  return MPM_OK;
}

static void drop_mesh(mpm_ctx* ctx) {
  MeshHost& m = ctx->mesh;
  dfree(m.faces); dfree(m.rest); dfree(m.vcsr_off); dfree(m.vcsr); dfree(m.elem_fv); dfree(m.elem_stress);
  dfree(m.vert_force); dfree(m.fixed);
  m = MeshHost{};
  for (SceneHost& sc : ctx->scenes) sc.mesh_v0 = sc.mesh_e0 = sc.mesh_e1 = INT32_MAX;
}

int mpm_set_particles(mpm_ctx* ctx, int32_t n, const float* x, const float* v, const float* B, const float* F,
                      const float* plastic, const float* mass, const float* volume, const float* phi) {
  if (!ctx) return MPM_ERR_INVALID;
  if (!ctx->have_grid) return fail(ctx, MPM_ERR_STATE, "mpm_set_grid or mpm_set_scenes must precede mpm_set_particles");
  if (n <= 0 || !x || !mass) return fail(ctx, MPM_ERR_INVALID, "n > 0, x and mass are required");
  if (ctx->batch) {
    int64_t total = 0;
    for (const SceneHost& sc : ctx->scenes) total += sc.n;
    if (total != n) return fail(ctx, MPM_ERR_INVALID, "n must equal the scenes' particle counts summed");
  }
  for (const SceneHost& sc : ctx->scenes)
    if (sc.params.per_particle_phi && !phi) return fail(ctx, MPM_ERR_INVALID, "per_particle_phi needs phi");
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  drop_mesh(ctx);  // a mesh is attached after the particles (mpm_set_mesh)
  if (!ctx->batch) ctx->scenes[0].n = n;
  const int64_t cap = n;
  std::vector<float> h((size_t)NPLANES * cap, 0.0f);
  int i = 0;
  for (const SceneHost& sc : ctx->scenes) {
    const float jp0 = sc.params.material == MPM_SNOW ? 1.0f : 0.0f;
    for (int e = i + sc.n; i < e; ++i) {
      for (int a = 0; a < 3; ++a) {
        h[(PX + a) * cap + i] = x[(size_t)a * n + i];
        h[(PV + a) * cap + i] = v ? v[(size_t)a * n + i] : 0.0f;
      }
      for (int c = 0; c < 9; ++c) {
        h[(PB + c) * cap + i] = B ? B[(size_t)c * n + i] : 0.0f;
        h[(PF + c) * cap + i] = F ? F[(size_t)c * n + i] : ((c % 4 == 0) ? 1.0f : 0.0f);
      }
      h[PPL * cap + i] = plastic ? plastic[i] : jp0;
      h[PM * cap + i] = mass[i];
      h[PVOL * cap + i] = volume ? volume[i] : 0.0f;
      h[PPHI * cap + i] = phi ? phi[i] : 0.0f;
      int id = i;
      std::memcpy(&h[PID * cap + i], &id, sizeof(int));
    }
  }
  int rc = install_state(ctx, n, h);
  if (!rc) ctx->n_api = n;
  return rc;
}

int mpm_set_mesh(mpm_ctx* ctx, int32_t nv, const float* x, const float* v, int32_t nf, const int32_t* faces,
                 const uint8_t* fixed, const mpm_mesh_params* mp) {
  if (!ctx) return MPM_ERR_INVALID;
  if (ctx->batch || ctx->scenes.size() != 1) return fail(ctx, MPM_ERR_UNSUPPORTED, "a mesh needs a single-scene context");
  // a cloth after the particles, before their first substep, or a cloth alone (no particles, the
  // reference's main.cpp:77-84), which replaces any cloth the context held
  if (!ctx->have_grid) return fail(ctx, MPM_ERR_STATE, "mpm_set_grid must precede mpm_set_mesh");
  if (ctx->n_api > 0 && (ctx->n != ctx->n_api || ctx->substeps != 0 || ctx->keys_valid))
    return fail(ctx, MPM_ERR_STATE, "mpm_set_mesh follows mpm_set_particles (if any), before the first substep");
  if (nv < 3 || nf < 1 || !x || !faces || !mp) return fail(ctx, MPM_ERR_INVALID, "a mesh needs vertices, faces and parameters");
  if (!(mp->poisson_ratio < 0.5f) || !(mp->youngs_modulus > 0.0f) || !(mp->density > 0.0f) || !(mp->thickness > 0.0f))
    return fail(ctx, MPM_ERR_INVALID, "bad cloth material");
  for (int64_t k = 0; k < 3 * (int64_t)nf; ++k)
    if (faces[k] < 0 || faces[k] >= nv) return fail(ctx, MPM_ERR_INVALID, "face vertex index out of range");
  const int np = ctx->n_api, n = np + nv + nf;
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  // the particles' planes (caller order: no substep has run)
  std::vector<float> hp((size_t)NPLANES * np);
  if (np > 0) CK(hipMemcpy(hp.data(), ctx->state[ctx->cur], sizeof(float) * hp.size(), hipMemcpyDeviceToHost));
  std::vector<float> h((size_t)NPLANES * n, 0.0f);
  for (int p = 0; p < NPLANES; ++p) std::memcpy(&h[(size_t)p * n], &hp[(size_t)p * np], sizeof(float) * np);
  // LagrangianMesh::ObjMesh (LagrangianMesh.cpp:300-352): element volume 0.25 area thickness, added
  // to each vertex; masses density x volume; rest directions v2 - v1, v3 - v1, unit normal.
  std::vector<double> vvol(nv, 0.0);
  std::vector<float> rest((size_t)9 * nf, 0.0f);  // the rest R (entries 6..8) is filled on the device
  auto X = [&](int vtx, int a) { return (double)x[(size_t)a * nv + vtx]; };
  for (int e = 0; e < nf; ++e) {
    double p[3][3];
    for (int c = 0; c < 3; ++c)
      for (int a = 0; a < 3; ++a) p[c][a] = X(faces[3 * e + c], a);
    auto len = [](const double* a, const double* b) {
      return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
    };
    const double la = len(p[0], p[1]), lb = len(p[1], p[2]), lc = len(p[2], p[0]), hp2 = 0.5 * (la + lb + lc);
    const double area = std::sqrt(std::max(0.0, hp2 * (hp2 - la) * (hp2 - lb) * (hp2 - lc)));  // geometry.cpp:12-23
    const double vol = 0.25 * area * mp->thickness;
    for (int c = 0; c < 3; ++c) vvol[faces[3 * e + c]] += vol;
    double d1[3], d2[3], nr[3];
    for (int a = 0; a < 3; ++a) {
      d1[a] = p[1][a] - p[0][a];
      d2[a] = p[2][a] - p[0][a];
    }
    nr[0] = d1[1] * d2[2] - d1[2] * d2[1];
    nr[1] = d1[2] * d2[0] - d1[0] * d2[2];
    nr[2] = d1[0] * d2[1] - d1[1] * d2[0];
    const double nl = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    if (!(nl > 0.0)) return fail(ctx, MPM_ERR_INVALID, "degenerate triangle");
    for (int a = 0; a < 3; ++a) nr[a] /= nl;
    // rest QR (geometry.cpp:31-62) and inverseR of its in-plane block (:67-73)
    const double r11 = std::sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
    const double q1[3] = {d1[0] / r11, d1[1] / r11, d1[2] / r11};
    const double r12 = d2[0] * q1[0] + d2[1] * q1[1] + d2[2] * q1[2];
    double q2[3] = {d2[0] - r12 * q1[0], d2[1] - r12 * q1[1], d2[2] - r12 * q1[2]};
    const double r22 = std::sqrt(q2[0] * q2[0] + q2[1] * q2[1] + q2[2] * q2[2]);
    rest[9 * e] = float(1.0 / r11);
    rest[9 * e + 1] = float(-r12 / r11 / r22);
    rest[9 * e + 2] = float(1.0 / r22);
    // normal-force vector: the reference's (D_rest^T).col(2) or D_rest^-T e3 = (d1 x d2) / det(D_rest)
    const double Dr[3][3] = {{d1[0], d2[0], nr[0]}, {d1[1], d2[1], nr[1]}, {d1[2], d2[2], nr[2]}};
    if (mp->reference_normal_row) {
      for (int c = 0; c < 3; ++c) rest[9 * e + 3 + c] = float(Dr[2][c]);
    } else {
      const double det = Dr[0][0] * (Dr[1][1] * Dr[2][2] - Dr[1][2] * Dr[2][1]) -
                         Dr[0][1] * (Dr[1][0] * Dr[2][2] - Dr[1][2] * Dr[2][0]) +
                         Dr[0][2] * (Dr[1][0] * Dr[2][1] - Dr[1][1] * Dr[2][0]);
      const double cr[3] = {d1[1] * d2[2] - d1[2] * d2[1], d1[2] * d2[0] - d1[0] * d2[2], d1[0] * d2[1] - d1[1] * d2[0]};
      for (int c = 0; c < 3; ++c) rest[9 * e + 3 + c] = float(cr[c] / det);
    }
    // element particle
    const int i = np + nv + e;
    for (int a = 0; a < 3; ++a) {
      h[(PX + a) * (size_t)n + i] = float((p[0][a] + p[1][a] + p[2][a]) / 3.0);
      double vm = 0.0;
      if (v)
        for (int c = 0; c < 3; ++c) vm += v[(size_t)a * nv + faces[3 * e + c]];
      h[(PV + a) * (size_t)n + i] = float(vm / 3.0);
      h[(PF + 3 * a) * (size_t)n + i] = float(d1[a]);
      h[(PF + 3 * a + 1) * (size_t)n + i] = float(d2[a]);
      h[(PF + 3 * a + 2) * (size_t)n + i] = float(nr[a]);
    }
    h[PM * (size_t)n + i] = float(mp->density * vol);
    h[PVOL * (size_t)n + i] = float(vol);
    std::memcpy(&h[PID * (size_t)n + i], &i, sizeof(int));
  }
  for (int k = 0; k < nv; ++k) {
    const int i = np + k;
    for (int a = 0; a < 3; ++a) {
      h[(PX + a) * (size_t)n + i] = x[(size_t)a * nv + k];
      h[(PV + a) * (size_t)n + i] = v ? v[(size_t)a * nv + k] : 0.0f;
    }
    h[(PF + 0) * (size_t)n + i] = h[(PF + 4) * (size_t)n + i] = h[(PF + 8) * (size_t)n + i] = 1.0f;
    h[PM * (size_t)n + i] = float(mp->density * vvol[k]);
    h[PVOL * (size_t)n + i] = float(vvol[k]);
    std::memcpy(&h[PID * (size_t)n + i], &i, sizeof(int));
  }
  // vertex -> incident element corners, ascending element (deterministic force sums)
  std::vector<int> off(nv + 1, 0), ent((size_t)3 * nf), fid((size_t)3 * nf), fixed_ids;
  for (int64_t k = 0; k < 3 * (int64_t)nf; ++k) off[faces[k] + 1] += 1;
  for (int k = 0; k < nv; ++k) off[k + 1] += off[k];
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int e = 0; e < nf; ++e)
    for (int c = 0; c < 3; ++c) ent[fill[faces[3 * e + c]]++] = 3 * e + c;
  for (int64_t k = 0; k < 3 * (int64_t)nf; ++k) fid[k] = np + faces[k];
  if (fixed)
    for (int k = 0; k < nv; ++k)
      if (fixed[k]) fixed_ids.push_back(np + k);
  drop_mesh(ctx);
  MeshHost& m = ctx->mesh;
  int rc = dalloc(ctx, m.faces, (size_t)3 * nf);
  if (!rc) rc = dalloc(ctx, m.rest, (size_t)9 * nf);
  if (!rc) rc = dalloc(ctx, m.vcsr_off, (size_t)nv + 1);
  if (!rc) rc = dalloc(ctx, m.vcsr, (size_t)3 * nf);
  if (!rc) rc = dalloc(ctx, m.elem_fv, (size_t)9 * nf);
  if (!rc) rc = dalloc(ctx, m.elem_stress, (size_t)9 * nf);
  if (!rc) rc = dalloc(ctx, m.vert_force, (size_t)3 * nv);
  if (!rc && !fixed_ids.empty()) rc = dalloc(ctx, m.fixed, fixed_ids.size());
  if (rc) return rc;
  CK(hipMemcpy(m.faces, fid.data(), sizeof(int) * fid.size(), hipMemcpyHostToDevice));
  CK(hipMemcpy(m.rest, rest.data(), sizeof(float) * rest.size(), hipMemcpyHostToDevice));
  CK(hipMemcpy(m.vcsr_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
  CK(hipMemcpy(m.vcsr, ent.data(), sizeof(int) * ent.size(), hipMemcpyHostToDevice));
  if (!fixed_ids.empty()) CK(hipMemcpy(m.fixed, fixed_ids.data(), sizeof(int) * fixed_ids.size(), hipMemcpyHostToDevice));
  m.np = np;
  m.nv = nv;
  m.ne = nf;
  m.nfixed = (int)fixed_ids.size();
  SceneHost& sc = ctx->scenes[0];
  sc.n = n;
  sc.mesh_v0 = np;
  sc.mesh_e0 = np + nv;
  sc.mesh_e1 = n;
  const double E = mp->youngs_modulus, nu = mp->poisson_ratio;
  sc.c_lambda = float(E * nu / (1.0 + nu) / (1.0 - 2.0 * nu));  // LagrangianMesh.cpp:340-341
  sc.c_mu = float(E / 2.0 / (1.0 + nu));
  sc.c_gamma = mp->shear_stiffness;
  sc.c_k = mp->stiffness;
  sc.c_cf = float(std::tan(double(mp->friction_angle) * 3.14159265358979323846 / 180.0));
  rc = install_state(ctx, n, h);
  if (rc) return rc;
  launch_cloth_rest(ctx->stream, cloth_args(ctx, ctx->cur), m.rest);
  CK(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}

int mpm_get_mesh(mpm_ctx* ctx, float* vx, float* vv, float* D) {
  int rc = ready(ctx);
  if (rc) return rc;
  const MeshHost& m = ctx->mesh;
  if (m.ne == 0) return fail(ctx, MPM_ERR_STATE, "no mesh");
  CK(hipSetDevice(ctx->device));
  const int n = ctx->n;
  float* tmp = nullptr;
  rc = dalloc(ctx, tmp, (size_t)9 * n);
  if (rc) return rc;
  struct Req { float* dst; int p0; int np; int first; int count; };
  const Req reqs[3] = {{vx, PX, 3, m.np, m.nv}, {vv, PV, 3, m.np, m.nv}, {D, PF, 9, m.np + m.nv, m.ne}};
  hipError_t e = hipSuccess;
  for (const Req& r : reqs) {
    if (!r.dst) continue;
    launch_unpermute(ctx->stream, ctx->state[ctx->cur], ctx->cap, n, r.p0, r.np, tmp);
    for (int p = 0; p < r.np && e == hipSuccess; ++p)
      e = hipMemcpyAsync(r.dst + (size_t)p * r.count, tmp + (size_t)p * n + r.first, sizeof(float) * r.count,
                         hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) break;
  }
  dfree(tmp);
  if (e != hipSuccess) return fail(ctx, MPM_ERR_HIP, std::string("mpm_get_mesh: ") + hipGetErrorString(e));
  return MPM_OK;
}

int mpm_get_particles(mpm_ctx* ctx, float* x, float* v, float* B, float* F, float* plastic, float* volume) {
  int rc = ready(ctx);
  if (rc) return rc;
  CK(hipSetDevice(ctx->device));
  const int n = ctx->n;
  float* tmp = nullptr;
  rc = dalloc(ctx, tmp, (size_t)9 * n);
  if (rc) return rc;
  struct Req { float* dst; int p0; int np; };
  const Req reqs[6] = {{x, PX, 3}, {v, PV, 3}, {B, PB, 9}, {F, PF, 9}, {plastic, PPL, 1}, {volume, PVOL, 1}};
  hipError_t e = hipSuccess;
  for (const Req& r : reqs) {
    if (!r.dst) continue;
    launch_unpermute(ctx->stream, ctx->state[ctx->cur], ctx->cap, n, r.p0, r.np, tmp);
    // the particles only (a mesh's vertices and elements follow them in caller order)
    for (int p = 0; p < r.np && e == hipSuccess; ++p)
      e = hipMemcpyAsync(r.dst + (size_t)p * ctx->n_api, tmp + (size_t)p * n, sizeof(float) * ctx->n_api,
                         hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) break;
  }
  dfree(tmp);
  if (e != hipSuccess) return fail(ctx, MPM_ERR_HIP, std::string("mpm_get_particles: ") + hipGetErrorString(e));
  return MPM_OK;
}

int mpm_get_stress(mpm_ctx* ctx, float* K) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (!K) return fail(ctx, MPM_ERR_INVALID, "K is required");
  if (ctx->predicted) return fail(ctx, MPM_ERR_STATE, "no stress state with stress_at_predicted_F");
  CK(hipSetDevice(ctx->device));
  if (!ctx->stress_valid) {
    launch_stress_init(ctx->stream, ctx->T, ctx->nblocks_p, ctx->state[ctx->cur], ctx->cap, ctx->n, ctx->fallbacks);
    ctx->stress_valid = true;
  }
  float* tmp = nullptr;
  rc = dalloc(ctx, tmp, (size_t)6 * ctx->n);
  if (rc) return rc;
  launch_unpermute(ctx->stream, ctx->state[ctx->cur], ctx->cap, ctx->n, PK, 6, tmp);
  hipError_t e = hipSuccess;
  for (int p = 0; p < 6 && e == hipSuccess; ++p)
    e = hipMemcpyAsync(K + (size_t)p * ctx->n_api, tmp + (size_t)p * ctx->n, sizeof(float) * ctx->n_api,
                       hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dfree(tmp);
  if (e != hipSuccess) return fail(ctx, MPM_ERR_HIP, std::string("mpm_get_stress: ") + hipGetErrorString(e));
  return MPM_OK;
}

int mpm_set_stress(mpm_ctx* ctx, const float* K) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (!K) return fail(ctx, MPM_ERR_INVALID, "K is required");
  if (ctx->predicted) return fail(ctx, MPM_ERR_STATE, "no stress state with stress_at_predicted_F");
  if (ctx->next_stage != MPM_STAGE_SORT) return fail(ctx, MPM_ERR_STATE, "a staged substep is in progress");
  CK(hipSetDevice(ctx->device));
  float* tmp = nullptr;
  rc = dalloc(ctx, tmp, (size_t)6 * ctx->n);
  if (rc) return rc;
  hipError_t e = hipMemsetAsync(tmp, 0, sizeof(float) * 6 * ctx->n, ctx->stream);
  for (int p = 0; p < 6 && e == hipSuccess; ++p)
    e = hipMemcpyAsync(tmp + (size_t)p * ctx->n, K + (size_t)p * ctx->n_api, sizeof(float) * ctx->n_api,
                       hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    ensure_pos(ctx);
    launch_permute_in(ctx->stream, ctx->state[ctx->cur], ctx->cap, ctx->n, PK, 6, tmp, ctx->pos[ctx->cur]);
    e = hipStreamSynchronize(ctx->stream);
  }
  dfree(tmp);
  if (e != hipSuccess) return fail(ctx, MPM_ERR_HIP, std::string("mpm_set_stress: ") + hipGetErrorString(e));
  ctx->stress_valid = true;
  drop_graphs(ctx);
  return MPM_OK;
}

int mpm_estimate_volumes(mpm_ctx* ctx) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (ctx->next_stage != MPM_STAGE_SORT) return fail(ctx, MPM_ERR_STATE, "a substep is in progress");
  CK(hipSetDevice(ctx->device));
  rc = stage_sort(ctx);
  if (rc) return rc;
  P2GArgs a = p2g_args(ctx, 0.0f);
  launch_p2g(ctx->stream, a, ctx->order, P2G_MOM, ctx->p2g_grid, ctx->p2g_ranks, ctx->p2g_pass);
  launch_grid(ctx->stream, grid_args(ctx, 0.0f), GRID_NORMALIZE, ctx->grid_grid);
  launch_volumes(ctx->stream, ctx->T, ctx->order, ctx->nblocks_p, ctx->state[ctx->cur], ctx->cap, ctx->n, ctx->grid);
  ctx->stress_valid = false;
  CK(hipGetLastError());
  CK(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}

int mpm_stage(mpm_ctx* ctx, int32_t stage, float dt) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (stage != ctx->next_stage) return fail(ctx, MPM_ERR_STATE, "stages must be issued in order SORT, P2G, GRID, G2P");
  CK(hipSetDevice(ctx->device));
  rc = run_stage(ctx, stage, dt);
  if (rc) return rc;
  CK(hipGetLastError());
  ctx->next_stage = (stage + 1) % MPM_NUM_STAGES;
  return MPM_OK;
}

int mpm_substep(mpm_ctx* ctx, float dt) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (ctx->next_stage != MPM_STAGE_SORT) return fail(ctx, MPM_ERR_STATE, "a staged substep is in progress");
  CK(hipSetDevice(ctx->device));
  return one_substep(ctx, dt);
}

int mpm_run(mpm_ctx* ctx, float dt, int32_t nsteps) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (nsteps < 0) return fail(ctx, MPM_ERR_INVALID, "nsteps < 0");
  if (ctx->next_stage != MPM_STAGE_SORT) return fail(ctx, MPM_ERR_STATE, "a staged substep is in progress");
  CK(hipSetDevice(ctx->device));
  if (nsteps > 0 && (!ctx->keys_valid || (!ctx->predicted && !ctx->stress_valid))) {
    rc = one_substep(ctx, dt);
    if (rc) return rc;
    --nsteps;
  }
  const int pairs = nsteps / 2;
  if (pairs > 0) {
    const int c0 = ctx->cur;
    if (!ctx->graph[c0] || ctx->graph_dt[c0] != dt) {
      if (ctx->graph[c0]) {
        (void)hipGraphExecDestroy(ctx->graph[c0]);
        ctx->graph[c0] = nullptr;
      }
      hipGraph_t g = nullptr;
      const int64_t steps0 = ctx->substeps;
      const int sort0 = ctx->sort_buf;
      CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int r1 = one_substep(ctx, dt);
      int r2 = r1 ? r1 : one_substep(ctx, dt);
      hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
      // capture executed nothing: the state is still in the buffers it was in
      ctx->substeps = steps0;
      ctx->cur = c0;
      ctx->sort_buf = sort0;
      if (r2 || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (r2) return r2;
        return fail(ctx, MPM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
      }
      hipError_t ei = hipGraphInstantiate(&ctx->graph[c0], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) return fail(ctx, MPM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
      ctx->graph_dt[c0] = dt;
      ctx->cur = c0;  // capture did not execute anything
    }
    for (int i = 0; i < pairs; ++i) CK(hipGraphLaunch(ctx->graph[c0], ctx->stream));
    ctx->substeps += 2 * pairs;
    ctx->sort_buf = 1 - c0;
    ctx->cur = c0;
  }
  if (nsteps % 2) {
    rc = one_substep(ctx, dt);
    if (rc) return rc;
  }
  CK(hipGetLastError());
  return MPM_OK;
}

// ------------------------------------------------------------------ adaptive time step
// HybridSolver::solve (HybridSolver.cpp:827-1034): Dt = CFL / max(300, max|v_i| / h_min) from the
// grid velocities of the first P2G (:854-860), then per substep from the velocities after the
// velocity update (:878), clamped to the end of the 1/60 s frame (:880-892); the loop ends when
// t > maxt after a frame (:867). The clock lives on the device; substeps are replayed two at a time
// from a captured graph, and a substep that completes a frame makes the rest of the batch do nothing,
// so the host sees every frame boundary exactly and calls `cb` there (the reference writes its OBJ
// frames at that point, :990-1029).
int mpm_solve(mpm_ctx* ctx, double cfl, double maxt, mpm_frame_callback cb, void* user) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (!(cfl > 0.0) || !(maxt >= 0.0)) return fail(ctx, MPM_ERR_INVALID, "mpm_solve: cfl must be > 0 and maxt >= 0");
  if (ctx->scenes.size() != 1)
    return fail(ctx, MPM_ERR_UNSUPPORTED, "mpm_solve: the adaptive step couples scenes; use one scene per context");
  if (ctx->next_stage != MPM_STAGE_SORT) return fail(ctx, MPM_ERR_STATE, "a staged substep is in progress");
  CK(hipSetDevice(ctx->device));
  if (!ctx->clock) {
    rc = dalloc(ctx, ctx->clock, 1);
    if (rc) return rc;
  }
  Clock h{};
  h.cfl = cfl;
  h.maxt = maxt;
  const SceneHost& sc0 = ctx->scenes[0];
  h.hmin = std::min(sc0.h[0], std::min(sc0.h[1], sc0.h[2]));
  CK(hipMemcpyAsync(ctx->clock, &h, sizeof(Clock), hipMemcpyHostToDevice, ctx->stream));
  ctx->clk_active = ctx->clock;
  struct Off {  // fixed-step runs never see the clock
    mpm_ctx* c;
    ~Off() { c->clk_active = nullptr; }
  } off{ctx};
  // first Dt from the velocities p/m of a P2G without forces
  rc = stage_sort(ctx);
  if (rc) return rc;
  launch_p2g(ctx->stream, p2g_args(ctx, 0.0f), ctx->order, P2G_MOM, ctx->p2g_grid, ctx->p2g_ranks, ctx->p2g_pass);
  GridArgs g0 = grid_args(ctx, 0.0f);
  g0.vmax_out = &ctx->clock->vmax_bits;
  launch_grid(ctx->stream, g0, GRID_NORMALIZE, ctx->grid_grid);
  launch_clock_update(ctx->stream, ctx->clock, 1);
  CK(hipGetLastError());
  while (true) {
    int executed = 0;
    const bool eager = !ctx->keys_valid || (!ctx->predicted && !ctx->stress_valid);
    if (eager) {
      rc = one_substep(ctx, 0.0f);
      if (rc) return rc;
    } else {
      const int c0 = ctx->cur;
      if (!ctx->graph_adapt[c0]) {
        hipGraph_t gr = nullptr;
        const int64_t steps0 = ctx->substeps;
        const int sort0 = ctx->sort_buf;
        CK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int r1 = one_substep(ctx, 0.0f);
        int r2 = r1 ? r1 : one_substep(ctx, 0.0f);
        hipError_t ec = hipStreamEndCapture(ctx->stream, &gr);
        ctx->substeps = steps0;
        ctx->cur = c0;  // capture did not execute anything
        ctx->sort_buf = sort0;
        if (r2 || ec != hipSuccess) {
          if (gr) (void)hipGraphDestroy(gr);
          if (r2) return r2;
          return fail(ctx, MPM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
        }
        hipError_t ei = hipGraphInstantiate(&ctx->graph_adapt[c0], gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ei != hipSuccess) return fail(ctx, MPM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
      }
      CK(hipGraphLaunch(ctx->graph_adapt[c0], ctx->stream));
      executed = -1;  // read back below
    }
    CK(hipMemcpyAsync(&h, ctx->clock, sizeof(Clock), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (executed < 0) {
      // the replayed pair ran one or two substeps; the state is in the buffer the last one wrote
      const int c0 = ctx->cur;
      const int e = h.executed;
      ctx->substeps += e;
      ctx->cur = (e == 2) ? c0 : 1 - c0;
      ctx->sort_buf = (e == 2) ? 1 - c0 : c0;
    }
    const bool frame = h.frame_done != 0;
    const bool done = h.done != 0;
    h.frame_done = 0;
    h.skip = 0;
    h.executed = 0;
    CK(hipMemcpyAsync(ctx->clock, &h, sizeof(Clock), hipMemcpyHostToDevice, ctx->stream));
    if (frame && cb) cb(ctx, h.frame - 1, h.t, user);
    if (done) break;
  }
  CK(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}

// ------------------------------------------------------------------ autotuner
// Sort, P2G and the grid update are idempotent for a given particle state (they rebuild the bins,
// the tile partials and the node arrays from it), so candidate launch shapes are timed on the live
// state without advancing it. None of the candidates changes an order of summation: every choice
// gives the same bits.
static float time_launches(mpm_ctx* ctx, int reps, const std::function<void()>& f) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1.0f;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1.0f;
  }
  f();  // warm-up
  (void)hipEventRecord(e0, ctx->stream);
  for (int r = 0; r < reps; ++r) f();
  (void)hipEventRecord(e1, ctx->stream);
  float ms = -1.0f;
  if (hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

int mpm_autotune(mpm_ctx* ctx, float dt, int32_t reps, mpm_tuning* chosen) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (reps < 1) return fail(ctx, MPM_ERR_INVALID, "reps >= 1");
  if (ctx->next_stage != MPM_STAGE_SORT) return fail(ctx, MPM_ERR_STATE, "a staged substep is in progress");
  CK(hipSetDevice(ctx->device));
  rc = stage_sort(ctx);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  if (!ctx->predicted && !ctx->stress_valid) {
    launch_stress_init(s, ctx->T, ctx->nblocks_p, ctx->state[ctx->cur], ctx->cap, ctx->n, ctx->fallbacks);
    ctx->stress_valid = true;
  }
  const int p2g_mode = ctx->predicted ? P2G_MOM : P2G_MOM_STRESS;
  const int grid_mode = ctx->predicted ? GRID_NORMALIZE : GRID_FUSED;
  mpm_tuning best{};
  best.sort_ms = 1e30f;
  best.p2g_pass = -1;
  const int keep_mode = ctx->sort_mode;
  for (int m : {SORT_COUNTING, SORT_RADIX}) {
    ctx->sort_mode = m;
    const float ms = time_launches(ctx, reps, [&] { (void)stage_sort(ctx); });
    CK(hipGetLastError());
    if (ms >= 0.0f && ms < best.sort_ms) {
      best.sort_ms = ms;
      best.sort_mode = m;
    }
  }
  ctx->sort_mode = best.sort_ms < 1e30f ? best.sort_mode : keep_mode;
  rc = stage_sort(ctx);  // the bins of the chosen mode (identical for both) for the P2G timings
  if (rc) return rc;
  const P2GArgs a = p2g_args(ctx, dt);
  best.p2g_ms = 1e30f;
  struct P2GCand { int pass, ranks, mult; };
  const P2GCand cands[] = {{P2G_ROWS, 4, 1}, {P2G_ROWS, 4, 2}, {P2G_ROWS, 4, 4}, {P2G_ROWS, 8, 1},
                           {P2G_ROWS, 8, 2}, {P2G_ROWS, 8, 4}, {P2G_LAYERS, 2, 2}, {P2G_LAYERS, 2, 3},
                           {P2G_LAYERS, 2, 6}, {P2G_LAYERS, 4, 2}, {P2G_LAYERS, 4, 4}};
  for (const P2GCand& cd : cands) {
    const int blocks = std::max(64, cd.mult * ctx->n_cu);
    const float ms = time_launches(ctx, reps, [&] { launch_p2g(s, a, ctx->order, p2g_mode, blocks, cd.ranks, cd.pass); });
    CK(hipGetLastError());
    if (ms >= 0.0f && ms < best.p2g_ms) {
      best.p2g_ms = ms;
      best.p2g_ranks = cd.ranks;
      best.p2g_blocks = blocks;
      best.p2g_pass = cd.pass;
    }
  }
  launch_p2g(s, a, ctx->order, p2g_mode, best.p2g_blocks, best.p2g_ranks, best.p2g_pass);
  const GridArgs g = grid_args(ctx, dt);
  best.grid_ms = 1e30f;
  const int gmults[4] = {2, 4, 8, 16};
  for (int m : gmults) {
    const int blocks = std::max(64, m * ctx->n_cu);
    const float ms = time_launches(ctx, reps, [&] { launch_grid(s, g, grid_mode, blocks); });
    CK(hipGetLastError());
    if (ms >= 0.0f && ms < best.grid_ms) {
      best.grid_ms = ms;
      best.grid_blocks = blocks;
    }
  }
  CK(hipStreamSynchronize(s));
  if (best.p2g_ranks == 0 || best.grid_blocks == 0) return fail(ctx, MPM_ERR_HIP, "autotune: no timing succeeded");
  ctx->p2g_ranks = best.p2g_ranks;
  ctx->p2g_grid = best.p2g_blocks;
  ctx->p2g_pass = best.p2g_pass;
  ctx->grid_grid = best.grid_blocks;
  drop_graphs(ctx);
  if (chosen) *chosen = best;
  return MPM_OK;
}

int mpm_get_tuning(mpm_ctx* ctx, mpm_tuning* t) {
  if (!ctx || !t) return MPM_ERR_INVALID;
  *t = mpm_tuning{};
  t->p2g_ranks = ctx->p2g_ranks;
  t->p2g_blocks = ctx->p2g_grid;
  t->grid_blocks = ctx->grid_grid;
  t->sort_mode = ctx->sort_mode;
  t->p2g_pass = ctx->p2g_pass;
  return MPM_OK;
}

int mpm_set_tuning(mpm_ctx* ctx, const mpm_tuning* t) {
  if (!ctx || !t) return MPM_ERR_INVALID;
  if (t->p2g_pass != MPM_P2G_ROWS && t->p2g_pass != MPM_P2G_LAYERS)
    return fail(ctx, MPM_ERR_INVALID, "p2g_pass must be MPM_P2G_ROWS or MPM_P2G_LAYERS");
  const bool ranks_ok = t->p2g_pass == MPM_P2G_ROWS ? (t->p2g_ranks == 4 || t->p2g_ranks == 8)
                                                     : (t->p2g_ranks == 2 || t->p2g_ranks == 4);
  if (!ranks_ok || t->p2g_blocks < 1 || t->grid_blocks < 1)
    return fail(ctx, MPM_ERR_INVALID, "p2g_ranks must be 4 or 8 (rows) or 2 or 4 (layers), block counts positive");
  if (t->sort_mode != MPM_SORT_COUNTING && t->sort_mode != MPM_SORT_RADIX)
    return fail(ctx, MPM_ERR_INVALID, "sort_mode must be MPM_SORT_COUNTING or MPM_SORT_RADIX");
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  ctx->p2g_ranks = t->p2g_ranks;
  ctx->p2g_grid = t->p2g_blocks;
  ctx->grid_grid = t->grid_blocks;
  ctx->sort_mode = t->sort_mode;
  ctx->p2g_pass = t->p2g_pass;
  drop_graphs(ctx);
  return MPM_OK;
}

int mpm_get_clock(mpm_ctx* ctx, double* t, double* dt, int32_t* frames, int64_t* substeps) {
  if (!ctx) return MPM_ERR_INVALID;
  if (!ctx->clock) return fail(ctx, MPM_ERR_STATE, "mpm_solve has not been called");
  CK(hipSetDevice(ctx->device));
  Clock h{};
  CK(hipMemcpyAsync(&h, ctx->clock, sizeof(Clock), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  if (t) *t = h.t;
  if (dt) *dt = h.dt;
  if (frames) *frames = h.frame;
  if (substeps) *substeps = h.substeps;
  return MPM_OK;
}

int mpm_get_sorted(mpm_ctx* ctx, uint32_t* keys, int32_t* ids, int32_t* bins) {
  int rc = ready(ctx);
  if (rc) return rc;
  CK(hipSetDevice(ctx->device));
  const int n = ctx->n;
  CK(hipStreamSynchronize(ctx->stream));
  if (keys) CK(hipMemcpy(keys, ctx->ks_sorted, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
  if (ids) CK(hipMemcpy(ids, ctx->vs_sorted, sizeof(int) * n, hipMemcpyDeviceToHost));
  if (bins) {
    // every cell's start from the sorted keys (the scan wrote only the blocks holding particles)
    launch_fill_bins(ctx->stream, ctx->ks_sorted, n, ctx->ng, ctx->cell_start);
    CK(hipStreamSynchronize(ctx->stream));
    CK(hipMemcpy(bins, ctx->cell_start, sizeof(int) * (ctx->ng + 1), hipMemcpyDeviceToHost));
  }
  return MPM_OK;
}

int mpm_get_grid(mpm_ctx* ctx, float* m, float* mom, float* v, float* vbf) {
  if (!ctx) return MPM_ERR_INVALID;
  if (!ctx->have_grid) return fail(ctx, MPM_ERR_STATE, "no grid");
  CK(hipSetDevice(ctx->device));
  const int64_t ng = ctx->ng;
  float* tmp = nullptr;
  int rc = dalloc(ctx, tmp, (size_t)3 * ng);
  if (rc) return rc;
  struct Req { float* dst; int which; int np; };
  const Req reqs[4] = {{m, 0, 1}, {mom, 1, 3}, {v, 2, 3}, {vbf, 3, 3}};
  hipError_t e = hipSuccess;
  for (const Req& r : reqs) {
    if (!r.dst) continue;
    launch_grid_export(ctx->stream, ctx->T, ctx->grid, ng, r.which, ctx->block_slot, tmp);
    e = hipMemcpyAsync(r.dst, tmp, sizeof(float) * r.np * ng, hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) break;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dfree(tmp);
  if (e != hipSuccess) return fail(ctx, MPM_ERR_HIP, std::string("mpm_get_grid: ") + hipGetErrorString(e));
  return MPM_OK;
}

int mpm_get_stats(mpm_ctx* ctx, mpm_stats* s) {
  if (!ctx || !s) return MPM_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  std::memset(s, 0, sizeof(*s));
  s->substeps = ctx->substeps;
  s->n_particles = ctx->n;
  int cnt[CNT_NUM] = {0, 0};
  int status = 0;
  unsigned long long fb[4] = {0, 0, 0, 0};
  unsigned long long sc[4] = {0, 0, 0, 0};
  if (ctx->have_grid && ctx->n > 0) {
    CK(hipMemsetAsync(ctx->fallbacks + 2, 0, sizeof(unsigned long long), ctx->stream));
    launch_count_nodes(ctx->stream, ctx->T, ctx->grid, ctx->counters, ctx->active_blocks, ctx->fallbacks + 2, ctx->grid_grid);
  }
  CK(hipMemcpyAsync(cnt, ctx->counters, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipMemcpyAsync(&status, ctx->status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipMemcpyAsync(fb, ctx->fallbacks, sizeof(fb), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipMemcpyAsync(sc, ctx->sur_counts, sizeof(sc), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  s->status = status;
  s->n_active_tiles = cnt[CNT_TILES];
  s->n_active_blocks = cnt[CNT_BLOCKS];
  s->n_active_nodes = (int64_t)fb[2];
  s->fp64_svd_count = (int64_t)fb[0];
  s->fp64_apex_count = (int64_t)fb[1];
  s->surrogate_candidates = (int64_t)sc[0];
  s->surrogate_accepted = (int64_t)sc[1];
  s->pairs_logged = (int64_t)std::min<unsigned long long>(sc[2], (unsigned long long)std::max<int64_t>(ctx->pair_cap, 0));
  return MPM_OK;
}

int mpm_clear_status(mpm_ctx* ctx) {
  if (!ctx) return MPM_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  CK(hipMemsetAsync(ctx->status, 0, sizeof(int), ctx->stream));
  CK(hipMemsetAsync(ctx->fallbacks, 0, 2 * sizeof(unsigned long long), ctx->stream));
  CK(hipMemsetAsync(ctx->sur_counts, 0, 2 * sizeof(unsigned long long), ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}

int mpm_synchronize(mpm_ctx* ctx) {
  if (!ctx) return MPM_ERR_INVALID;
  CK(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}

// Eagerly launched substeps with HIP events around every stage, around the G2P main kernel and,
// when sur_ms is given and the surrogate is on, around its kernel (*sur_ms).
static int time_kernels(mpm_ctx* ctx, float dt, int32_t nsteps, float* total_ms, float* stage_ms, float* kernel_ms,
                        float* sur_ms) {
  int rc = ready(ctx);
  if (rc) return rc;
  if (ctx->next_stage != MPM_STAGE_SORT) return fail(ctx, MPM_ERR_STATE, "a staged substep is in progress");
  CK(hipSetDevice(ctx->device));
  hipEvent_t ev[MPM_NUM_STAGES + 1];
  for (auto& e : ev) CK(hipEventCreate(&e));
  hipEvent_t t0, t1;
  CK(hipEventCreate(&t0));
  CK(hipEventCreate(&t1));
  CK(hipEventCreate(&ctx->ev_main[0]));
  CK(hipEventCreate(&ctx->ev_main[1]));
  const bool sur = sur_ms && ctx->sur_enabled;
  if (sur)
    for (auto& e : ctx->ev_sur) CK(hipEventCreate(&e));
  struct Clear {  // the stages record the main-kernel events only inside this call
    mpm_ctx* c;
    ~Clear() {
      for (auto& e : c->ev_main) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
      }
      for (auto& e : c->ev_sur) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
      }
    }
  } clear{ctx};
  float acc[MPM_NUM_STAGES] = {0, 0, 0, 0}, g2p_main = 0.0f, sacc = 0.0f;
  CK(hipEventRecord(t0, ctx->stream));
  for (int i = 0; i < nsteps; ++i) {
    CK(hipEventRecord(ev[0], ctx->stream));
    for (int st = 0; st < MPM_NUM_STAGES; ++st) {
      rc = run_stage(ctx, st, dt);
      if (rc) return rc;
      CK(hipEventRecord(ev[st + 1], ctx->stream));
    }
    CK(hipEventSynchronize(ev[MPM_NUM_STAGES]));
    for (int st = 0; st < MPM_NUM_STAGES; ++st) {
      float ms = 0;
      CK(hipEventElapsedTime(&ms, ev[st], ev[st + 1]));
      acc[st] += ms;
    }
    float ms = 0;
    CK(hipEventElapsedTime(&ms, ctx->ev_main[0], ctx->ev_main[1]));
    g2p_main += ms;
    if (sur) {
      CK(hipEventElapsedTime(&ms, ctx->ev_sur[0], ctx->ev_sur[1]));
      sacc += ms;
    }
  }
  CK(hipEventRecord(t1, ctx->stream));
  CK(hipEventSynchronize(t1));
  float tot = 0;
  CK(hipEventElapsedTime(&tot, t0, t1));
  if (total_ms) *total_ms = tot;
  if (stage_ms)
    for (int st = 0; st < MPM_NUM_STAGES; ++st) stage_ms[st] = acc[st];
  if (kernel_ms) {
    for (int st = 0; st < MPM_NUM_STAGES; ++st) kernel_ms[st] = acc[st];
    kernel_ms[MPM_STAGE_G2P] = g2p_main;
  }
  if (sur_ms) *sur_ms = sacc;
  for (auto& e : ev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  CK(hipGetLastError());
  return MPM_OK;
}

int mpm_time_kernels(mpm_ctx* ctx, float dt, int32_t nsteps, float* total_ms, float* stage_ms, float* kernel_ms) {
  return time_kernels(ctx, dt, nsteps, total_ms, stage_ms, kernel_ms, nullptr);
}

int mpm_time_surrogate(mpm_ctx* ctx, float dt, int32_t nsteps, float* kernel_ms, int64_t* candidates) {
  if (!ctx || !kernel_ms) return MPM_ERR_INVALID;
  if (!ctx->sur_enabled) return fail(ctx, MPM_ERR_STATE, "the surrogate is not enabled");
  unsigned long long c0 = 0, c1 = 0;
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  CK(hipMemcpy(&c0, ctx->sur_counts, sizeof(c0), hipMemcpyDeviceToHost));
  float ms = 0.0f;
  int rc = time_kernels(ctx, dt, nsteps, nullptr, nullptr, nullptr, &ms);
  if (rc) return rc;
  CK(hipMemcpy(&c1, ctx->sur_counts, sizeof(c1), hipMemcpyDeviceToHost));
  *kernel_ms = ms;
  if (candidates) *candidates = (int64_t)(c1 - c0);
  return MPM_OK;
}

int mpm_time_substeps(mpm_ctx* ctx, float dt, int32_t nsteps, float* total_ms, float* stage_ms) {
  return mpm_time_kernels(ctx, dt, nsteps, total_ms, stage_ms, nullptr);
}

int mpm_bench_copy(mpm_ctx* ctx, int64_t nbytes, int32_t reps, float* ms) {
  if (!ctx || nbytes <= 0 || nbytes % 16 != 0 || reps < 1) return ctx ? fail(ctx, MPM_ERR_INVALID, "bad copy size") : MPM_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  float4 *a = nullptr, *b = nullptr;
  int rc = dalloc(ctx, a, (size_t)nbytes / 16);
  if (!rc) rc = dalloc(ctx, b, (size_t)nbytes / 16);
  if (rc) { dfree(a); dfree(b); return rc; }
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  CK(hipMemsetAsync(a, 0, (size_t)nbytes, ctx->stream));
  CK(hipEventRecord(e0, ctx->stream));
  for (int i = 0; i < reps; ++i) launch_copy(ctx->stream, a, b, nbytes / 16);
  CK(hipEventRecord(e1, ctx->stream));
  CK(hipEventSynchronize(e1));
  float t = 0.0f;
  CK(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dfree(a);
  dfree(b);
  return MPM_OK;
}

// ------------------------------------------------------------------ learned return-map surrogate
static inline uint64_t mix64(uint64_t z);

static int ensure_mlp(mpm_ctx* ctx) {
  if (ctx->mlp_master) return MPM_OK;
  int rc = dalloc(ctx, ctx->mlp_master, MLP_NPARAM);
  if (!rc) rc = dalloc(ctx, ctx->mlp_m1, MLP_NPARAM);
  if (!rc) rc = dalloc(ctx, ctx->mlp_m2, MLP_NPARAM);
  if (!rc) rc = dalloc(ctx, ctx->mlp_grad, MLP_NPARAM);
  if (!rc) rc = dalloc(ctx, ctx->mlp_wbf, MLP_BF_SIZE);
  if (!rc) rc = dalloc(ctx, ctx->mlp_wbft, MLP_T_SIZE);
  if (rc) return rc;
  CK(hipMemsetAsync(ctx->mlp_wbft, 0, sizeof(uint16_t) * MLP_T_SIZE, ctx->stream));
  mlp_set_smem_limits();
  surrogate_set_smem_limit();
  return MPM_OK;
}

static int upload_weights(mpm_ctx* ctx, const float* w) {
  CK(hipMemcpyAsync(ctx->mlp_master, w, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, ctx->stream));
  CK(hipMemsetAsync(ctx->mlp_m1, 0, sizeof(float) * MLP_NPARAM, ctx->stream));
  CK(hipMemsetAsync(ctx->mlp_m2, 0, sizeof(float) * MLP_NPARAM, ctx->stream));
  launch_mlp_to_bf16(ctx->stream, ctx->mlp_master, ctx->mlp_wbf, ctx->mlp_wbft);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(ctx->stream));
  ctx->mlp_step = 0;
  ctx->mlp_ready = true;
  return MPM_OK;
}

int32_t mpm_surrogate_nparam(void) { return MLP_NPARAM; }

int mpm_surrogate_init(mpm_ctx* ctx, uint64_t seed) {
  if (!ctx) return MPM_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  int rc = ensure_mlp(ctx);
  if (rc) return rc;
  // uniform(-a, a) with a = sqrt(3 / fan_in) (unit variance for unit inputs); output layer scaled 0.1
  std::vector<float> w(MLP_NPARAM, 0.0f);
  for (int l = 0; l < 5; ++l) {
    const int N = mlp_N(l), K = mlp_K(l), w0 = mlp_weight_offset(l);
    const int fan_in = l == 0 ? MLP_IN : K;
    const double a = std::sqrt(3.0 / fan_in) * (l == 4 ? 0.1 : 1.0);
    for (int o = 0; o < N; ++o)
      for (int k = 0; k < K; ++k) {
        const int i = w0 + o * K + k;
        if (!mlp_param_active(i)) continue;
        const double u = (double)(mix64(seed * 0x9E3779B97F4A7C15ull ^ (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0);
        w[i] = (float)((2.0 * u - 1.0) * a);
      }
  }
  return upload_weights(ctx, w.data());
}

int mpm_surrogate_set_weights(mpm_ctx* ctx, const float* w, int32_t nparam) {
  if (!ctx || !w || nparam != MLP_NPARAM) return ctx ? fail(ctx, MPM_ERR_INVALID, "nparam mismatch") : MPM_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  int rc = ensure_mlp(ctx);
  if (rc) return rc;
  return upload_weights(ctx, w);
}

int mpm_surrogate_get_weights(mpm_ctx* ctx, float* w, int32_t nparam) {
  if (!ctx || !w || nparam != MLP_NPARAM) return ctx ? fail(ctx, MPM_ERR_INVALID, "nparam mismatch") : MPM_ERR_INVALID;
  if (!ctx->mlp_ready) return fail(ctx, MPM_ERR_STATE, "surrogate weights not set");
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  CK(hipMemcpy(w, ctx->mlp_master, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
  return MPM_OK;
}

// normalised targets y = (H - eps) / CORR_SCALE
static void targets_of(const float* in, const float* H, int64_t n, std::vector<float>& y) {
  y.resize((size_t)3 * n);
  for (int64_t r = 0; r < n; ++r)
    for (int c = 0; c < 3; ++c) y[3 * r + c] = (H[3 * r + c] - in[5 * r + c]) / MLP_CORR_SCALE;
}

// workgroups of the forward kernels: one per CU at most (each stages the weights once and its 8
// waves take 64 rows at a time)
static int forward_grid(const mpm_ctx* ctx, int64_t rows) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(ctx->n_cu, (rows + MLP_FWD_THREADS - 1) / MLP_FWD_THREADS));
}

int mpm_surrogate_predict(mpm_ctx* ctx, const float* in, float* H, int32_t n) {
  if (!ctx || !in || !H || n < 0) return MPM_ERR_INVALID;
  if (!ctx->mlp_ready) return fail(ctx, MPM_ERR_STATE, "surrogate weights not set");
  if (n == 0) return MPM_OK;
  CK(hipSetDevice(ctx->device));
  float *din = nullptr, *dout = nullptr;
  int rc = dalloc(ctx, din, (size_t)5 * n);
  if (!rc) rc = dalloc(ctx, dout, (size_t)3 * n);
  if (rc) { dfree(din); dfree(dout); return rc; }
  hipError_t e = hipMemcpyAsync(din, in, sizeof(float) * 5 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    launch_mlp_forward(ctx->stream, ctx->mlp_wbf, ctx->mlp_master, din, dout, n, nullptr, forward_grid(ctx, n));
    e = hipMemcpyAsync(H, dout, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dfree(din);
  dfree(dout);
  if (e != hipSuccess) return fail(ctx, MPM_ERR_HIP, std::string("mpm_surrogate_predict: ") + hipGetErrorString(e));
  for (int64_t r = 0; r < n; ++r)
    for (int c = 0; c < 3; ++c) H[3 * r + c] = in[5 * r + c] + MLP_CORR_SCALE * H[3 * r + c];
  return MPM_OK;
}

// one optimisation pass machinery: forward (saving pre-activations), backward, ordered reduction
struct TrainBufs {
  float* in = nullptr;
  float* tgt = nullptr;
  float* y = nullptr;
  float* z = nullptr;
  float* gpart = nullptr;
  float* lpart = nullptr;
  float* lacc = nullptr;
};

static void free_train(TrainBufs& b) {
  dfree(b.in); dfree(b.tgt); dfree(b.y); dfree(b.z); dfree(b.gpart); dfree(b.lpart); dfree(b.lacc);
}

// workgroups of the backward pass (each owns a partial-gradient slice of MLP_NPARAM floats)
static int backward_grid(int rows) { return std::min((rows + 63) / 64, 1024); }

static int alloc_train(mpm_ctx* ctx, TrainBufs& b, int64_t n, int batch, int nacc) {
  const int groups = backward_grid(batch);
  int rc = dalloc(ctx, b.in, (size_t)5 * n);
  if (!rc) rc = dalloc(ctx, b.tgt, (size_t)3 * n);
  if (!rc) rc = dalloc(ctx, b.y, (size_t)3 * batch);
  if (!rc) rc = dalloc(ctx, b.z, (size_t)((batch + 15) / 16) * 4 * 16 * MLP_HIDDEN);
  if (!rc) rc = dalloc(ctx, b.gpart, (size_t)groups * MLP_NPARAM);
  if (!rc) rc = dalloc(ctx, b.lpart, (size_t)groups);
  if (!rc) rc = dalloc(ctx, b.lacc, (size_t)nacc);
  return rc;
}

static void train_batch(mpm_ctx* ctx, TrainBufs& b, int64_t r0, int nb, float* lacc) {
  const int groups = backward_grid(nb);
  launch_mlp_forward(ctx->stream, ctx->mlp_wbf, ctx->mlp_master, b.in + 5 * r0, b.y, nb, b.z, forward_grid(ctx, nb));
  launch_mlp_backward(ctx->stream, ctx->mlp_wbft, b.in + 5 * r0, b.y, b.tgt + 3 * r0, b.z, nb, b.gpart, b.lpart, groups);
  launch_mlp_reduce(ctx->stream, b.gpart, groups, ctx->mlp_grad, b.lpart, lacc);
}

int mpm_surrogate_gradient(mpm_ctx* ctx, const float* in, const float* H, int32_t n, float* grad, float* loss) {
  if (!ctx || !in || !H || n <= 0) return MPM_ERR_INVALID;
  if (!ctx->mlp_ready) return fail(ctx, MPM_ERR_STATE, "surrogate weights not set");
  CK(hipSetDevice(ctx->device));
  TrainBufs b;
  int rc = alloc_train(ctx, b, n, n, 1);
  if (rc) { free_train(b); return rc; }
  std::vector<float> y;
  targets_of(in, H, n, y);
  hipError_t e = hipMemcpyAsync(b.in, in, sizeof(float) * 5 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(b.tgt, y.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(b.lacc, 0, sizeof(float), ctx->stream);
  if (e == hipSuccess) {
    train_batch(ctx, b, 0, n, b.lacc);
    e = hipGetLastError();
  }
  float lsum = 0.0f;
  if (e == hipSuccess && grad) e = hipMemcpyAsync(grad, ctx->mlp_grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&lsum, b.lacc, sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  free_train(b);
  if (e != hipSuccess) return fail(ctx, MPM_ERR_HIP, std::string("mpm_surrogate_gradient: ") + hipGetErrorString(e));
  if (loss) *loss = lsum / (float)n;
  return MPM_OK;
}

int mpm_surrogate_gradient_noalloc(mpm_ctx* ctx, const float* in, const float* H, int32_t n, float* grad, float* loss) {
  if (!ctx) return MPM_ERR_INVALID;
  if (!in || !H || !grad || !loss) return fail(ctx, MPM_ERR_INVALID, "mpm_surrogate_gradient_noalloc: null pointer");
  if (n < 1 || n > NOALLOC_MAX_ROWS)
    return fail(ctx, MPM_ERR_INVALID, "mpm_surrogate_gradient_noalloc: n must be in [1, " +
                                          std::to_string(NOALLOC_MAX_ROWS) + "], got " + std::to_string(n));
  if (!ctx->mlp_ready) return fail(ctx, MPM_ERR_STATE, "surrogate weights not set");
  if (!ctx->scratch) return fail(ctx, MPM_ERR_STATE, "no scratch address for this device");
  CK(hipSetDevice(ctx->device));
  // the same kernels and launch shapes as mpm_surrogate_gradient; every scratch region a kernel
  // reads (rows < n, the first `groups` partials) is written earlier in the same call
  float* s = ctx->scratch;
  const int groups = backward_grid(n);
  launch_mlp_targets(ctx->stream, in, H, n, s + NOALLOC_TGT);
  launch_mlp_forward(ctx->stream, ctx->mlp_wbf, ctx->mlp_master, in, s + NOALLOC_Y, n, s + NOALLOC_Z, forward_grid(ctx, n));
  launch_mlp_backward(ctx->stream, ctx->mlp_wbft, in, s + NOALLOC_Y, s + NOALLOC_TGT, s + NOALLOC_Z, n, s + NOALLOC_GPART,
                      s + NOALLOC_LPART, groups);
  launch_mlp_reduce(ctx->stream, s + NOALLOC_GPART, groups, grad, s + NOALLOC_LPART, nullptr);
  launch_mlp_loss_mean(ctx->stream, s + NOALLOC_LPART, groups, n, loss);
  CK(hipGetLastError());
  return MPM_OK;
}

int mpm_surrogate_train(mpm_ctx* ctx, const float* in, const float* H, int32_t n, int32_t epochs, int32_t batch,
                        float lr, uint64_t seed, float* epoch_loss) {
  if (!ctx || !in || !H || n <= 0 || epochs < 1 || batch < 16) return ctx ? fail(ctx, MPM_ERR_INVALID, "bad training arguments") : MPM_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  int rc = ensure_mlp(ctx);
  if (rc) return rc;
  if (!ctx->mlp_ready) {
    rc = mpm_surrogate_init(ctx, seed);
    if (rc) return rc;
  }
  // seeded shuffle (Fisher-Yates with a splitmix64 stream)
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  for (int i = n - 1; i > 0; --i) {
    const int j = (int)(mix64(seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(i + 1))) % (uint64_t)(i + 1));
    std::swap(order[i], order[j]);
  }
  std::vector<float> ins((size_t)5 * n), hs((size_t)3 * n), y;
  for (int r = 0; r < n; ++r) {
    for (int c = 0; c < 5; ++c) ins[5 * (size_t)r + c] = in[5 * (size_t)order[r] + c];
    for (int c = 0; c < 3; ++c) hs[3 * (size_t)r + c] = H[3 * (size_t)order[r] + c];
  }
  targets_of(ins.data(), hs.data(), n, y);
  if (batch > n) batch = n;
  TrainBufs b;
  rc = alloc_train(ctx, b, n, batch, epochs);
  if (rc) { free_train(b); return rc; }
  hipError_t e = hipMemcpyAsync(b.in, ins.data(), sizeof(float) * 5 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(b.tgt, y.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(b.lacc, 0, sizeof(float) * epochs, ctx->stream);
  for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
    for (int64_t r0 = 0; r0 < n; r0 += batch) {
      const int nb = (int)std::min<int64_t>(batch, n - r0);
      train_batch(ctx, b, r0, nb, b.lacc + ep);
      launch_mlp_adam(ctx->stream, ctx->mlp_master, ctx->mlp_m1, ctx->mlp_m2, ctx->mlp_grad, ctx->mlp_wbf, ctx->mlp_wbft,
                      lr, ++ctx->mlp_step);
    }
    e = hipGetLastError();
  }
  std::vector<float> l(epochs, 0.0f);
  if (e == hipSuccess) e = hipMemcpyAsync(l.data(), b.lacc, sizeof(float) * epochs, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  free_train(b);
  if (e != hipSuccess) return fail(ctx, MPM_ERR_HIP, std::string("mpm_surrogate_train: ") + hipGetErrorString(e));
  if (epoch_loss)
    for (int ep = 0; ep < epochs; ++ep) epoch_loss[ep] = l[ep] / (float)n;
  return MPM_OK;
}

int mpm_set_pair_logging(mpm_ctx* ctx, int64_t capacity) {
  if (!ctx || capacity < 0) return MPM_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  int rc = dalloc(ctx, ctx->pair_log, (size_t)8 * capacity);
  if (rc) return rc;
  ctx->pair_cap = capacity;
  CK(hipMemsetAsync(ctx->sur_counts + 2, 0, sizeof(unsigned long long), ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}

int mpm_get_pairs(mpm_ctx* ctx, float* pairs, int64_t max_pairs, int64_t* count) {
  if (!ctx || max_pairs < 0) return MPM_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  unsigned long long c = 0;
  CK(hipMemcpy(&c, ctx->sur_counts + 2, sizeof(c), hipMemcpyDeviceToHost));
  int64_t m = std::min<int64_t>((int64_t)c, ctx->pair_cap);
  m = std::min<int64_t>(m, max_pairs);
  if (count) *count = m;
  if (pairs && m > 0) CK(hipMemcpy(pairs, ctx->pair_log, sizeof(float) * 8 * m, hipMemcpyDeviceToHost));
  return MPM_OK;
}

int mpm_set_surrogate_log(mpm_ctx* ctx, int64_t capacity) {
  if (!ctx || capacity < 0) return MPM_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  int rc = dalloc(ctx, ctx->dec_log, (size_t)10 * capacity);
  if (rc) return rc;
  ctx->dec_cap = capacity;
  return MPM_OK;
}

int mpm_get_surrogate_log(mpm_ctx* ctx, float* records, int64_t max_records, int64_t* count) {
  if (!ctx || max_records < 0) return MPM_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  int c = 0;
  if (ctx->counters) CK(hipMemcpy(&c, ctx->counters + CNT_CAND, sizeof(c), hipMemcpyDeviceToHost));
  int64_t m = ctx->sur_enabled ? std::min<int64_t>(std::max(0, std::min(c, ctx->n)), ctx->dec_cap) : 0;
  m = std::min<int64_t>(m, max_records);
  if (count) *count = m;
  if (records && m > 0) CK(hipMemcpy(records, ctx->dec_log, sizeof(float) * 10 * m, hipMemcpyDeviceToHost));
  return MPM_OK;
}

int mpm_set_surrogate(mpm_ctx* ctx, int32_t enable, float gate_tol) {
  if (!ctx || !(gate_tol >= 0.0f)) return MPM_ERR_INVALID;
  if (enable && !ctx->mlp_ready) return fail(ctx, MPM_ERR_STATE, "surrogate weights not set");
  CK(hipSetDevice(ctx->device));
  CK(hipStreamSynchronize(ctx->stream));
  drop_graphs(ctx);
  if (enable && ctx->n > 0 && !ctx->cand) {
    int rc = dalloc(ctx, ctx->cand, (size_t)CAND_SVD_PLANES * ctx->n);
    if (!rc) rc = dalloc(ctx, ctx->cand_slot, (size_t)ctx->n);
    if (!rc) rc = dalloc(ctx, ctx->cand_idx, (size_t)ctx->n);
    if (rc) return rc;
  }
  if (ctx->n > 0 && ctx->wave_cap < 4 * (int64_t)ctx->nblocks_p) {
    int rc = alloc_wave_tables(ctx);
    if (rc) return rc;
  }
  ctx->sur_enabled = enable ? 1 : 0;
  ctx->gate_tol = gate_tol;
  return MPM_OK;
}

// ------------------------------------------------------------------ generator (host only)
static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int mpm_generate_block(uint64_t seed, const double lo[3], const double hi[3], const int32_t cells[3], int32_t ppa,
                       int64_t n, float* x) {
  if (!lo || !hi || !cells || !x || ppa < 1) return MPM_ERR_INVALID;
  const int64_t pp = (int64_t)ppa * ppa * ppa;
  const int64_t expect = (int64_t)cells[0] * cells[1] * cells[2] * pp;
  if (cells[0] < 1 || cells[1] < 1 || cells[2] < 1 || n != expect) return MPM_ERR_INVALID;
  const uint64_t sk = seed * 0x632BE59BD9B4E019ull;
  for (int k = 0; k < cells[2]; ++k)
    for (int j = 0; j < cells[1]; ++j)
      for (int i = 0; i < cells[0]; ++i) {
        const int64_t cell = ((int64_t)k * cells[1] + j) * cells[0] + i;
        const int c[3] = {i, j, k};
        for (int sz = 0; sz < ppa; ++sz)
          for (int sy = 0; sy < ppa; ++sy)
            for (int sx = 0; sx < ppa; ++sx) {
              const int64_t id = cell * pp + ((int64_t)sz * ppa + sy) * ppa + sx;
              const int s3[3] = {sx, sy, sz};
              for (int a = 0; a < 3; ++a) {
                const uint64_t r = splitmix64(sk ^ (uint64_t)(4 * id + a));
                const double u = (double)(r >> 40) * (1.0 / 16777216.0);
                const double t = (c[a] + (s3[a] + u) / ppa) / cells[a];
                x[(size_t)a * n + id] = (float)(lo[a] + (hi[a] - lo[a]) * t);
              }
            }
      }
  return MPM_OK;
}

int mpm_generate_shape(uint64_t seed, int32_t shape, const double* a, int64_t n, float* x) {
  if (!a || !x || n < 1) return MPM_ERR_INVALID;
  if (shape == MPM_SHAPE_BALL && !(a[3] > 0.0)) return MPM_ERR_INVALID;
  if (shape == MPM_SHAPE_BLOCK &&
      (!(a[3] > a[0]) || !(a[4] > a[1]) || !(a[5] > a[2]) || !(a[6] >= 0.0) || !(2.0 * a[6] < a[5] - a[2])))
    return MPM_ERR_INVALID;
  if (shape == MPM_SHAPE_CYLINDER && (!(a[3] > 0.0) || !(a[4] > 0.0) || (a[5] != 0.0 && a[5] != 1.0 && a[5] != 2.0)))
    return MPM_ERR_INVALID;
  if (shape < MPM_SHAPE_BALL || shape > MPM_SHAPE_CYLINDER) return MPM_ERR_INVALID;
  const uint64_t s = splitmix64((seed * 0x632BE59BD9B4E019ull) ^ (uint64_t)shape);
  const int64_t limit = 64 * n + 4096;
  int64_t i = 0;
  for (int64_t k = 0; k < limit && i < n; ++k) {
    double u[3];
    for (int c = 0; c < 3; ++c) u[c] = (double)(splitmix64(s ^ (uint64_t)(4 * k + c)) >> 40) * (1.0 / 16777216.0);
    double p[3];
    if (shape == MPM_SHAPE_BALL) {
      const double q[3] = {2.0 * u[0] - 1.0, 2.0 * u[1] - 1.0, 2.0 * u[2] - 1.0};
      if (q[0] * q[0] + q[1] * q[1] + q[2] * q[2] > 1.0) continue;
      for (int c = 0; c < 3; ++c) p[c] = a[c] + a[3] * q[c];
    } else if (shape == MPM_SHAPE_BLOCK) {
      for (int c = 0; c < 3; ++c) p[c] = a[c] + (a[3 + c] - a[c]) * u[c];
      const double hc[3] = {a[0], a[1], 0.5 * (a[5] + a[2])};
      const double d2 = (p[0] - hc[0]) * (p[0] - hc[0]) + (p[1] - hc[1]) * (p[1] - hc[1]) + (p[2] - hc[2]) * (p[2] - hc[2]);
      if (std::sqrt(d2) < a[6]) continue;
    } else {
      const double d[2] = {2.0 * u[0] - 1.0, 2.0 * u[1] - 1.0};
      if (d[0] * d[0] + d[1] * d[1] > 1.0) continue;
      const int ax = (int)a[5];
      for (int c = 0, j = 0; c < 3; ++c) p[c] = c == ax ? a[c] + a[4] * u[2] : a[c] + a[3] * d[j++];
    }
    for (int c = 0; c < 3; ++c) x[(size_t)c * n + i] = (float)p[c];
    ++i;
  }
  return i == n ? MPM_OK : MPM_ERR_INVALID;
}

}  // extern "C"
