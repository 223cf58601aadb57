// mpm_kernels.h -- layout constants, kernel argument blocks and host launchers of the gfx950 MPM
// substep (internal to libmpm_hip.so).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_device.hpp"

namespace mpm {

// Particle state planes (each `cap` floats; PID holds int32 bits).
enum Plane : int {
  PX = 0,     // x0 x1 x2
  PV = 3,     // v0 v1 v2
  PB = 6,     // B row-major (rows = affineMomenta_1..3)
  PF = 15,    // F_E row-major
  PPL = 24,   // J_P (snow) or plasticAmount q (sand)
  PM = 25,    // mass
  PVOL = 26,  // volume
  PPHI = 27,  // per-particle friction angle (degrees)
  PID = 28,   // caller index (int bits)
  PK = 29,    // volume-weighted Kirchhoff stress V tau: xx yy zz xy xz yz (stress at F^n only)
  NPLANES = 35
};

// Grid planes: m (1), momentum (3), then per node two float4 (v0 v1 v2 vbf0 | vbf1 vbf2 0 0).
constexpr int GRID_M = 0, GRID_MOM = 1, GRID_VEL = 4, GRID_FLOATS = 12;

constexpr int SORT_IPT = 16;
constexpr int SORT_TILE = 256 * SORT_IPT;  // keys per sort pass block
constexpr int HIST_IPT = 16;
constexpr int HIST_TILE = 256 * HIST_IPT;  // keys per histogram block
constexpr int SCAN_IPT = 16;
constexpr int SCAN_TILE = 256 * SCAN_IPT;  // cells per cell-scan block
constexpr int TILE = 4;                    // cells per tile edge
constexpr int FOOT = TILE + 3;             // footprint nodes per edge (cell window c-1 .. c+2)
constexpr int FOOT3 = FOOT * FOOT * FOOT;
// Per tile slot: m, p(3) as one float4 per footprint node, then f(3) as three planes; padded to a
// multiple of 4 floats so that every slot's float4 part is 16-byte aligned.
constexpr int PART_STRIDE = 7 * FOOT3 + 3;

enum Counter : int { CNT_TILES = 0, CNT_BLOCKS = 1, CNT_CAND = 2, CNT_APEX = 3, CNT_NUM = 4 };
enum P2GMode : int { P2G_MOM_STRESS = 0, P2G_MOM = 1, P2G_FORCE = 2 };
enum GridMode : int { GRID_FUSED = 0, GRID_NORMALIZE = 1, GRID_FORCE_UPDATE = 2 };

// Scenes of one context. A context simulates S >= 1 independent scenes; their particles, cells,
// tiles and nodes are concatenated scene-major, so that one radix sort, one P2G launch, one grid
// launch and one G2P launch cover all of them. Each table has S + 1 entries (prefix offsets).
// Particle-indexed launches run over 256-particle blocks that never straddle two scenes, so the
// scene of a workgroup is uniform.
struct SceneTable {
  int nscenes;
  const DevParams* params;   // [S] per-scene parameters, grid geometry and level sets
  const int* part_off;       // [S+1] particles (caller order and sorted order are both scene-major)
  const int* pblock_off;     // [S+1] 256-particle blocks
  const int* tile_off;       // [S+1] 4x4x4-cell tiles (= 4x4x4-node blocks)
  const int64_t* node_off;   // [S+1] nodes (= cells): global cell key = node_off[s] + local key
  const int* thr_off;        // [S] per-axis cell thresholds of the scene
};

// Largest s in [0, S) with off[s] <= x (off ascending).
template <class T>
__device__ __forceinline__ int scene_of(const T* off, int S, T x) {
  int lo = 0, hi = S - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Workgroups are dispatched to the chip's 8 XCDs round robin (workgroup b runs on XCD b % 8), and
// every XCD has its own L2. A persistent workgroup's share of `count` work items (tiles, node
// blocks): the items of XCD x are the contiguous range [first, end) of about count / 8 items, and
// the workgroups of that XCD take every stride-th item of it from first, so that neighbouring items
// (whose partials and nodes overlap) meet in one L2. With fewer than 8 workgroups, the plain grid
// stride. (Measured with 2^22 sand particles on 256^3: k_grid 58 -> 47 us, P2G unchanged; G2P, its
// particle blocks renumbered the same way, was 5% slower and keeps the plain order.)
constexpr int N_XCD = 8;
struct XcdRange {
  int first, end, stride;
};
__device__ __forceinline__ XcdRange xcd_range(int count) {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  if (nb < N_XCD) return {b, count, nb};
  const int x = b % N_XCD;
  const int lo = (int)((int64_t)count * x / N_XCD), hi = (int)((int64_t)count * (x + 1) / N_XCD);
  return {lo + b / N_XCD, hi, nb / N_XCD + (x < nb % N_XCD ? 1 : 0)};
}

// Scene and particle index of thread threadIdx.x of particle block pblk (valid: j < end).
struct PBlock {
  int s, j, end;
};
__device__ __forceinline__ PBlock particle_block(const SceneTable& T, int pblk) {
  PBlock b;
  b.s = __builtin_amdgcn_readfirstlane(scene_of(T.pblock_off, T.nscenes, pblk));
  b.j = T.part_off[b.s] + (pblk - T.pblock_off[b.s]) * 256 + (int)threadIdx.x;
  b.end = T.part_off[b.s + 1];
  return b;
}
__device__ __forceinline__ PBlock particle_block(const SceneTable& T) { return particle_block(T, (int)blockIdx.x); }

// Device clock of the adaptive time step (HybridSolver::solve, HybridSolver.cpp:860,878-892). The
// substep kernels take a pointer to it (null for fixed-step runs): they read Dt from it and do
// nothing once `skip` is set, which the last kernel of a substep does when the substep completed a
// frame or the run, so that a replayed batch of substeps stops exactly there.
struct Clock {
  double t;          // simulated time, advanced by 1/60 at each frame (the reference's t)
  double inner;      // time inside the current frame (innerFramet)
  double dt;         // Dt of the current substep
  double cfl, maxt, hmin;
  float dt_f;        // Dt as the kernels use it
  unsigned int vmax_bits;  // max |v| over nodes after the velocity update (float bits, atomicMax)
  int frame;         // frames completed
  int frame_now;     // the current substep completes a frame
  int frame_done;    // a frame completed since the host last cleared it
  int done;          // t > maxt
  int skip;          // the rest of the replayed batch does nothing
  int executed;      // substeps completed in the current batch
  long long substeps;  // substeps completed in total
};

__device__ __forceinline__ bool clock_skip(const Clock* c) { return c != nullptr && c->skip != 0; }
__device__ __forceinline__ float clock_dt(const Clock* c, float dt) { return c != nullptr ? c->dt_f : dt; }

struct ResetArgs {
  uint32_t* zero_words;
  int64_t n_zero;
  const int* counters;
  const int* active_tiles;
  int* tile_slot;
  const int* active_blocks;
  int* block_slot;
  int64_t max_tiles;
  const Clock* clk;
};

struct SortBufs {
  uint32_t *k1, *k2;
  int *v1, *v2;
  uint32_t* hist;     // [4][256]
  uint32_t* tickets;  // [4] digit passes
  uint32_t* status;   // [passes][nblocks][256] digit-pass look-back words
  uint32_t* scan_agg;   // [scan blocks] particles counted in each scan block's cells
  int* cell_count;    // [cells] zero between substeps (the bin pass clears the occupied cells)
  int* cell_start;    // [cells + 1] exclusive prefix of the counts; cell c holds [start[c], start[c+1])
                      // (written for the scan blocks with particles, see occ)
  uint32_t* occ;        // [scan blocks] 1 where a cell of the block holds a particle (set by the count)
  uint32_t* occ_clear;  // [scan blocks] the other substep's flags, cleared by the count
  int64_t ncells;
  SceneTable T;       // the counting sort also writes src and flags the occupied tiles
  int* tile_slot;
  const int* pid;     // caller id at each physical slot (the state's id plane)
  int* tslot;         // [n] scratch: the physical slot of each binned caller id
  int* pos;           // caller id -> physical slot: written by the radix mode's key conversion
  int* src;
  int* counters;      // zeroed by the cell scan (after the counting kernel has read them)
  ResetArgs reset;    // the counting sort's reset work, done by its counting kernel
  const Clock* clk;
};


struct P2GArgs {
  SceneTable T;
  const float* state;
  int64_t cap;
  int n;
  const int* src;   // sorted position -> physical index in `state`
  const int* cell_start;
  const int* cell_end;
  const uint32_t* occ;  // per scan block of cells: 1 if it holds particles (else its cells are empty)
  const int* counters;
  const int* active_tiles;
  float* partials;
  const float4* gvel;  // node velocities (force pass)
  float dt;
  unsigned long long* fallback_svd;
  const Clock* clk;
  const float* vert_force;   // cloth: force per mesh vertex (3), or null without a mesh
  const float* elem_stress;  // cloth: V dF3 d3^T per mesh element (3x3 row-major)
};

struct GridArgs {
  SceneTable T;
  const int* counters;
  const int* active_blocks;
  const int* tile_slot;
  const float* partials;
  float* gm;
  float* gmom;
  float4* gvel;
  int64_t ng;
  float dt;
  const Clock* clk;
  unsigned int* vmax_out;  // receives the max node speed after the velocity update (atomicMax of float bits)
};

struct G2PArgs {
  SceneTable T;
  const float* sin;
  float* sout;
  int64_t cap;
  int n;
  const int* perm;     // caller ids in sorted order
  const int* src;      // sorted position -> physical index in sin
  const uint32_t* ks;  // cell key of each sorted position (the last sort's; used only to place node boxes)
  int* pos_out;        // caller id -> physical index in sout (null: not maintained, no cloth)
  const float4* gvel;
  float dt;
  int write_stress;    // store V tau of the new state (stress at F^n)
  const float* thr;
  uint32_t* keys;      // next cell key, indexed by physical slot in sout
  int* status;
  unsigned long long* fallback_svd;
  unsigned long long* fallback_apex;
  // surrogate hooks: 0 = exact, 1 = exact + pair logging, 2 = defer cone-branch particles
  int sur_mode;
  float* pair_log;
  int64_t pair_cap;
  unsigned long long* pair_count;
  float* cand;      // candidate records: planes sigma 3, U 9, V 9 (of n floats), at the record index
  uint64_t* cand_mask;  // per wave (4 blockIdx.x + wave): which lanes are candidates
  int* wave_pos;        // per wave: the sorted position of its lane 0
  uint64_t* apex_mask;  // per wave: which lanes deferred their FP64 apex projection (null: none defer)
  float* apex_usv;      // their trial SVD: planes sigma 3, U 9, V 9 (of n floats), at the record index
  const Clock* clk;
};

// The deferred FP64 apex projections (after G2P): the rest of the return map from the stored SVD.
struct ApexArgs {
  SceneTable T;
  float* st;   // the G2P output buffer
  int64_t cap;
  int n;
  const int* apex_slot;   // the deferred particles' sorted slots, ascending
  const float* apex_usv;  // their trial SVDs
  const int* apex_idx;    // the record index of each listed particle in apex_usv
  const int* apex_count;
  int write_stress;
  int log_pairs;
  float* pair_log;
  int64_t pair_cap;
  unsigned long long* pair_count;
  unsigned long long* fallback_svd;
  unsigned long long* fallback_apex;
  const Clock* clk;
};
void launch_apex_finish(hipStream_t s, const ApexArgs& a, int grid);

// Planes of a deferred particle's trial SVD (surrogate candidates and FP64 apex projections). k_g2p
// writes the records of a wave's deferred lanes next to each other: the record index of lane L is
// the sorted position of the wave's lane 0 plus the number of deferred lanes below L (so it is at
// most L's own position, below n), and the position lists carry it with every entry.
constexpr int CAND_SIGMA = 0, CAND_U = 3, CAND_V = 12, CAND_SVD_PLANES = 21;

struct SurrogateArgs {
  SceneTable T;
  const uint16_t* wbf;   // BF16 forward weights
  const float* master;   // FP32 master weights (biases)
  float* st;             // particle state written by k_g2p (sorted slots)
  int64_t cap;
  int n;
  const float* cand;  // candidate SVDs (written by k_g2p)
  const int* cand_slot;
  const int* cand_idx;  // the record index of each listed candidate in cand
  const int* cand_count;
  float gate_tol;
  int write_stress;
  unsigned long long* accepted;
  unsigned long long* candidates;  // cumulative candidate count (block 0 adds this substep's)
  float* dec_log;                  // optional per-candidate gate record (10 floats per row)
  int64_t dec_cap;
  const Clock* clk;
};

// Particle-indexed launchers take the number of 256-particle blocks (SceneTable::pblock_off[S]).
// pos[id] = physical slot of caller id id, from the id plane pid of a state
void launch_pos_of_ids(hipStream_t s, const int* pid, int n, int* pos);
void launch_keys(hipStream_t s, const SceneTable& T, int nblocks, const float* st, int64_t cap, int n,
                 const float* thr, uint32_t* keys, int* status);
void launch_stress_init(hipStream_t s, const SceneTable& T, int nblocks, float* st, int64_t cap, int n,
                        unsigned long long* fallback);
void launch_reset(hipStream_t s, const ResetArgs& r);
// cell_start of the scan blocks without particles (only the getter of the bins needs them).
void launch_fill_bins(hipStream_t s, const uint32_t* ks, int n, int64_t ncells, int* cell_start);
// Stable sort of the cell keys by caller id. Both modes count the particles of every cell and scan
// the counts into cell_start. SORT_COUNTING then places each caller id at its stable rank inside
// its cell; SORT_RADIX runs `passes` 8-bit LSD digit passes (1..4). Both give the same order. The
// result is in k1/v1 for an odd `passes`, k2/v2 for an even one.
enum SortMode : int { SORT_COUNTING = 0, SORT_RADIX = 1 };
void launch_sort(hipStream_t s, const SortBufs& b, const uint32_t* keys_in, int n, int passes, int mode);
// Active tiles and node blocks from the sorted keys; clears the counts of the occupied cells and
// writes src[j] = pos[perm[j]]: the physical index of the particle at sorted position j.
void launch_bounds(hipStream_t s, const SceneTable& T, int nblocks, const uint32_t* ks, int n, int64_t ntiles,
                   int* cell_count, int* tile_slot, int* active_tiles, int* block_slot, int* active_blocks,
                   int* counters, const int* perm, const int* pos, int* src, const Clock* clk, int mode);
// order: B-spline order of every scene (2 or 3); ranks: particles per cell staged per chunk (4 or 8)
// pass P2G_ROWS (8-wave workgroups, ranks 4 or 8) or P2G_LAYERS (4-wave workgroups, ranks 2 or 4):
// two schedules of the same sums in the same order, so every choice gives the same bits.
enum P2GPass : int { P2G_ROWS = 0, P2G_LAYERS = 1 };
void launch_p2g(hipStream_t s, const P2GArgs& a, int order, int mode, int grid, int ranks, int pass);
void launch_grid(hipStream_t s, const GridArgs& a, int mode, int grid);
// With ev_start set, the launch records the kernel's own start and end into ev_start / ev_stop.
void launch_g2p(hipStream_t s, const G2PArgs& a, int order, int nblocks, hipEvent_t ev_start = nullptr,
                hipEvent_t ev_stop = nullptr);
void launch_volumes(hipStream_t s, const SceneTable& T, int order, int nblocks, float* st, int64_t cap, int n,
                    const float* gm);
void launch_unpermute(hipStream_t s, const float* st, int64_t cap, int n, int p0, int np, float* out);
void launch_permute_in(hipStream_t s, float* st, int64_t cap, int n, int p0, int np, const float* in, const int* pos);
void launch_grid_export(hipStream_t s, const SceneTable& T, const float* grid, int64_t ng, int which, const int* block_slot,
                        float* out);
void launch_iota(hipStream_t s, int* p, int n);
void launch_count_nodes(hipStream_t s, const SceneTable& T, const float* gm, const int* counters, const int* active_blocks,
                        unsigned long long* out, int grid);
void launch_fill_int(hipStream_t s, int* p, int64_t n, int v);
void launch_copy(hipStream_t s, const float4* src, float4* dst, int64_t n4);
// Dt from the max node speed (init: the first Dt, no frame bookkeeping), then the frame clamp.
void launch_clock_update(hipStream_t s, Clock* c, int init);
// End of a substep: count it, close a completed frame, stop the batch after a frame or the run.
void launch_clock_end(hipStream_t s, Clock* c);
// Cloth (mesh vertices and elements among the particles of scene 0; caller ids from the DevParams).
struct ClothArgs {
  DevParams P;
  float* st;               // particle state (state[cur] before P2G, sout after G2P)
  int64_t cap;
  int n;
  const int* pos;          // caller id -> physical index in st
  int nv, ne;
  const int* faces;        // [e][3] vertex caller ids
  const float* rest;       // [e][9] inverse rest in-plane R (i11 i12 i22), normal-force vector, rest R (R11 R12 R22)
  const int* vcsr_off;     // [nv+1] incident element corners of each vertex ...
  const int* vcsr;         // ... as 3 e + corner, ascending e
  float* elem_fv;          // [e][9] vertex forces f1 f2 f3 of each element
  float* elem_stress;      // [e][9]
  float* vert_force;       // [v][3]
  const int* fixed;        // caller ids of the fixed vertices
  int nfixed;
  float4* gvel;
  uint32_t* keys;
  const float* thr;
  int* status;
  unsigned long long* fallback;
  const Clock* clk;
};
void launch_cloth_rest(hipStream_t s, const ClothArgs& a, float* rest);  // rest in-plane R from the directions
void launch_cloth_forces(hipStream_t s, const ClothArgs& a);   // element forces, then the per-vertex sums
void launch_cloth_fixed(hipStream_t s, const ClothArgs& a);    // after the grid update
void launch_cloth_elements(hipStream_t s, const ClothArgs& a); // after G2P: positions, velocities, D, return map
void surrogate_set_smem_limit();
// Up to two lists of the sorted positions G2P flagged (mask[l][w] of wave w, whose lane 0 sits at
// wpos[w]), in ascending order and built without atomics; list l's length goes to *count[l].
struct ListArgs {
  const uint64_t* mask[2];
  int* list[2];
  int* idx[2];  // each entry's record index (see CAND_SIGMA)
  int* count[2];
  int* gsum;  // [2][(nwaves + 63) / 64]
  const int* wpos;
  int nwaves, nlists;
};
void launch_position_lists(hipStream_t s, const ListArgs& a);
// k_surrogate: MFMA network, gate and state update of the listed candidates (ev: optional start
// and stop events of the dispatch)
void launch_surrogate(hipStream_t s, const SurrogateArgs& a, int n_cu, hipEvent_t* ev);

}  // namespace mpm
