// mpm_kernels.hip -- gfx950 kernels of the explicit MPM substep and their host launchers.
//
// Data layout in HBM (a context holds one scene, or several scenes side by side):
//   particles   two ping-pong buffers of NPLANES float planes of `cap` floats each (SoA):
//               x(3) v(3) B(9) F(9) plastic m V phi id(int bits) V*tau(6). After the first substep
//               the planes are kept physically sorted by cell key, so every particle pass streams
//               them. A cloth's vertices and elements are particles of the same buffers.
//   keys        uint32 cell key per particle in the current physical order (written by G2P).
//   sort        counting sort: per-cell counts (zero between substeps), cell_start over all cells
//               plus one, scan status words and tickets; LSD radix: K1/K2 keys and V1/V2 index
//               ping-pong buffers and per-pass look-back status words.
//   grid        dense planes m, mom(3) and interleaved float4 pairs (v, v_BF) over the nodes; only
//               the 4x4x4 node blocks around active tiles are written each substep.
//   tiles       4x4x4-cell tiles: slot per tile (-1 = inactive), active list, per-slot P2G partials
//               over the tile's 7x7x7 node footprint.
//
// Substep (stress at F^n): reset -> cell count -> cell scan -> cell scatter -> cell rank (bins,
// tile flags) -> claim tiles -> claim node blocks -> p2g -> grid -> g2p -> apex finish.
// Reference-compatible substep (stress at predicted F^): p2g(momentum) -> grid(normalise) ->
// p2g(force, gathers grad v) -> grid(force update) replace the p2g -> grid pair. A cloth adds its
// element and vertex forces before P2G, the fixed-vertex node stops after the grid update, and the
// element update (directions, QR return map) after G2P.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "mpm_device.hpp"
#include "mpm_kernels.h"
#include "mpm_mlp_device.hpp"

namespace mpm {

// ================================================================== small helpers
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Indices read from device buffers are kept inside [0, n) so that an inconsistent buffer can only
// produce wrong values, never an out-of-bounds access.
__device__ __forceinline__ int clamp_index(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// wave-uniform float held in a scalar register (the readfirstlane builtin works on 32-bit integers)
__device__ __forceinline__ float uniform_f(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}

// The scene parameters of a workgroup-uniform scene, read once and held in scalar registers for the
// whole kernel (the table lives in global memory, which the compiler reads with a vector load at
// every use otherwise). The level-set primitives are not copied: kernels that use this copy do not
// evaluate level sets.
__device__ __forceinline__ DevParams uniform_params(const DevParams& G) {
  DevParams P;
  const auto ui = [](int x) { return __builtin_amdgcn_readfirstlane(x); };
  P.order = ui(G.order);
  P.material = ui(G.material);
  P.predicted = ui(G.predicted);
  P.friction_model = ui(G.friction_model);
  P.lambda0 = uniform_f(G.lambda0);
  P.mu0 = uniform_f(G.mu0);
  P.theta_c = uniform_f(G.theta_c);
  P.theta_s = uniform_f(G.theta_s);
  P.xi = uniform_f(G.xi);
  P.h0 = uniform_f(G.h0);
  P.h1 = uniform_f(G.h1);
  P.h2 = uniform_f(G.h2);
  P.h3 = uniform_f(G.h3);
  P.per_particle_phi = ui(G.per_particle_phi);
  P.alpha_f = uniform_f(G.alpha_f);
  const long long ad = __builtin_bit_cast(long long, G.alpha_d);
  P.alpha_d = __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)ui((int)(ad >> 32)) << 32) |
                                                    (unsigned)ui((int)ad)));
  P.friction = uniform_f(G.friction);
  P.damping = uniform_f(G.damping);
  P.svd_ratio = uniform_f(G.svd_ratio);
  P.apex_radius = uniform_f(G.apex_radius);
  P.dinv = uniform_f(G.dinv);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    P.gravity[d] = uniform_f(G.gravity[d]);
    P.gmin[d] = uniform_f(G.gmin[d]);
    P.h[d] = uniform_f(G.h[d]);
    P.inv_h[d] = uniform_f(G.inv_h[d]);
    P.res[d] = ui(G.res[d]);
    P.tiles[d] = ui(G.tiles[d]);
  }
  P.nls = 0;
  P.mesh_v0 = ui(G.mesh_v0);
  P.mesh_e0 = ui(G.mesh_e0);
  P.mesh_e1 = ui(G.mesh_e1);
  P.c_mu = uniform_f(G.c_mu);
  P.c_lambda = uniform_f(G.c_lambda);
  P.c_gamma = uniform_f(G.c_gamma);
  P.c_k = uniform_f(G.c_k);
  P.c_cf = uniform_f(G.c_cf);
  return P;
}

// Exact cell index along one axis: the number of thresholds thr[c] <= x, where thr[c] is the
// smallest float f with (double(f) - min)/h >= c. This equals static_cast<int>((x - min)/h) of the
// reference evaluated in double (HybridSolver.cpp:33-36) for x >= min, using FP32 compares only.
__device__ __forceinline__ int axis_cell(const float* thr, int res, float x, float gmin, float inv_h, int& oob) {
  if (!(x >= thr[0])) { oob = 1; return 0; }
  if (x >= thr[res]) { oob = 1; return res - 1; }
  int c = (int)floorf((x - gmin) * inv_h);
  c = c < 0 ? 0 : (c > res - 1 ? res - 1 : c);
  while (c > 0 && x < thr[c]) --c;
  while (c < res - 1 && x >= thr[c + 1]) ++c;
  return c;
}

// Lowest stencil node along one axis of a particle at grid coordinate g: floor(g - 0.5) for the
// quadratic kernel, floor(g) - 1 for the cubic one. The floor is clamped to [-8, res + 8] before the
// conversion, so that a non-finite or far out-of-grid position (an unstable run) gives a stencil whose
// nodes all lie outside the grid instead of an integer overflow: every index formed from it as
// base + o (o <= 3) stays far inside the int range, and every node is then dropped as out of grid.
template <int ORDER>
__device__ __forceinline__ int stencil_base(float g, int res) {
  float b = ORDER == 2 ? floorf(g - 0.5f) : floorf(g) - 1.0f;
  b = fminf(fmaxf(b, -8.0f), (float)res + 8.0f);  // NaN gives -8
  return (int)b;
}

__device__ __forceinline__ uint32_t cell_key(const DevParams& P, const float* thr, float x0, float x1, float x2, int& oob) {
  int c0 = axis_cell(thr, P.res[0], x0, P.gmin[0], P.inv_h[0], oob);
  int c1 = axis_cell(thr + (P.res[0] + 1), P.res[1], x1, P.gmin[1], P.inv_h[1], oob);
  int c2 = axis_cell(thr + (P.res[0] + 1) + (P.res[1] + 1), P.res[2], x2, P.gmin[2], P.inv_h[2], oob);
  return (uint32_t)(c2 * P.res[0] * P.res[1] + c1 * P.res[0] + c0);
}

// ================================================================== keys of the caller order
// keys[id] = cell key of caller particle id (physical index pos[id]); the radix sort consumes the
// keys in caller order, so that equal keys keep caller order as in the reference.
__global__ __launch_bounds__(256) void k_keys(SceneTable T, const float* __restrict__ st, int64_t cap, int n,
                                              const float* __restrict__ thr, uint32_t* __restrict__ keys,
                                              int* __restrict__ status) {
  const PBlock b = particle_block(T);  // physical slots are scene-major like the caller ids
  const DevParams& P = T.params[b.s];
  const int j = b.j;
  int oob = 0;
  if (j < b.end)
    keys[j] = (uint32_t)T.node_off[b.s] +
              cell_key(P, thr + T.thr_off[b.s], st[PX * cap + j], st[(PX + 1) * cap + j], st[(PX + 2) * cap + j], oob);
  if (__ballot(oob) && lane_id() == 0) atomicOr(status, MPM_STATUS_OUT_OF_DOMAIN);
}

// Volume-weighted Kirchhoff stress of the current F (after the caller sets particles or parameters).
__global__ __launch_bounds__(256) void k_stress_init(SceneTable T, float* __restrict__ st, int64_t cap, int n,
                                                     unsigned long long* __restrict__ fallback) {
  const PBlock b = particle_block(T);
  const DevParams& P = T.params[b.s];
  const int i = b.j;
  if (i >= b.end) return;
  float F[9], K[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int c = 0; c < 9; ++c) F[c] = st[(PF + c) * cap + i];
  int fb = 0;
  // cloth vertices and elements carry no particle stress (their F planes hold D for elements)
  const int pid = reinterpret_cast<const int*>(st + (size_t)PID * cap)[i];
  if (particle_kind(P, pid) == KIND_PARTICLE) kirchhoff_of_F(P, F, st[PPL * cap + i], st[PVOL * cap + i], K, fb);
#pragma unroll
  for (int e = 0; e < 6; ++e) st[(PK + e) * cap + i] = K[e];
  if (fb) atomicAdd(fallback, (unsigned long long)fb);
}

// ================================================================== radix sort (stable LSD, 8-bit digits)
// Per substep: one histogram launch for all three digit positions, then one launch per digit.
// A pass block ranks its 4096 keys stably with wave ballots (per-wave, per-digit running counts in
// LDS), publishes its per-digit counts and finds its global offsets by decoupled look-back over the
// preceding blocks (blocks take tickets in start order, so a block only waits on running blocks).
constexpr uint32_t FLAG_AGG = 1u << 30, FLAG_PREFIX = 2u << 30, VALUE_MASK = (1u << 30) - 1u;

__global__ __launch_bounds__(256) void k_reset(uint32_t* __restrict__ zero_words, int64_t n_zero,
                                               const int* __restrict__ counters, const int* __restrict__ active_tiles,
                                               int* __restrict__ tile_slot, const int* __restrict__ active_blocks,
                                               int* __restrict__ block_slot, const Clock* clk) {
  if (clock_skip(clk)) return;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = t0; i < n_zero; i += stride) zero_words[i] = 0u;
  int nt = counters[CNT_TILES], nb = counters[CNT_BLOCKS];
  for (int64_t j = t0; j < nt; j += stride) tile_slot[active_tiles[j]] = -1;
  for (int64_t j = t0; j < nb; j += stride) block_slot[active_blocks[j]] = -1;
}

// Lanes of a wave that share a digit are grouped with 8 ballots; one lane per group adds the count.
__device__ __forceinline__ uint64_t digit_peers(bool valid, uint32_t d) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int bit = 0; bit < 8; ++bit) {
    bool set = (d >> bit) & 1u;
    uint64_t bb = __ballot(valid && set);
    peers &= set ? bb : ~bb;
  }
  return peers;
}

__global__ __launch_bounds__(256) void k_sort_hist(const uint32_t* __restrict__ keys, int n, int passes,
                                                   uint32_t* __restrict__ hist, const Clock* clk) {
  if (clock_skip(clk)) return;
  __shared__ uint32_t h[4][256];
  for (int i = threadIdx.x; i < 1024; i += 256) (&h[0][0])[i] = 0u;
  __syncthreads();
  const int base = blockIdx.x * HIST_TILE;
  for (int k = 0; k < HIST_IPT; ++k) {
    const int i = base + k * 256 + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = valid ? keys[i] : 0u;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0ull) continue;
    const int first = __ffsll((long long)vmask) - 1;
    for (int p = 0; p < passes; ++p) {
      const uint32_t d = (key >> (8 * p)) & 255u;
      const uint32_t d0 = __shfl(d, first);
      if (__ballot(valid && d == d0) == vmask) {  // one digit across the wave: one add
        if ((int)lane_id() == first) atomicAdd(&h[p][d0], (uint32_t)__popcll(vmask));
      } else if (valid) {
        atomicAdd(&h[p][d], 1u);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256 * passes; i += 256) {
    uint32_t c = (&h[0][0])[i];
    if (c) atomicAdd(&hist[i], c);
  }
}

constexpr int LOOKBACK = 8;  // predecessor status words read per look-back step

__global__ __launch_bounds__(256) void k_sort_pass(const uint32_t* __restrict__ kin, const int* __restrict__ vin,
                                                   uint32_t* __restrict__ kout, int* __restrict__ vout, int n, int pass,
                                                   const uint32_t* __restrict__ hist, uint32_t* __restrict__ tickets,
                                                   uint32_t* __restrict__ status, const Clock* clk) {
  if (clock_skip(clk)) return;
  __shared__ uint32_t s_k[SORT_TILE];
  __shared__ int s_v[SORT_TILE];
  __shared__ uint32_t s_wh[4][256];
  __shared__ uint32_t s_off[256];
  __shared__ uint32_t s_loc[256];
  __shared__ uint32_t s_scan[256];
  __shared__ int s_tile;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  if (tid == 0) s_tile = (int)atomicAdd(&tickets[pass], 1u);
  for (int i = tid; i < 1024; i += 256) (&s_wh[0][0])[i] = 0u;
  __syncthreads();
  const int b = s_tile;
  if (b >= (int)gridDim.x) return;  // tickets are zeroed per substep; uniform across the block
  const int base = b * SORT_TILE + w * (64 * SORT_IPT);
  const int shift = 8 * pass;
  uint32_t kr[SORT_IPT];
  int vr[SORT_IPT];
  uint32_t rr[SORT_IPT];
  const uint64_t lt = (1ull << lane) - 1ull;
  // stable rank of every key among the equal digits before it in the block (waves in order,
  // 64-key chunks in order, lanes in order)
#pragma unroll
  for (int i = 0; i < SORT_IPT; ++i) {
    int pos = base + i * 64 + lane;
    bool valid = pos < n;
    uint32_t k = valid ? kin[pos] : 0u;
    int v = valid ? (vin ? vin[pos] : pos) : 0;
    uint32_t d = (k >> shift) & 255u;
    const uint64_t peers = digit_peers(valid, d);
    uint32_t prev = s_wh[w][d];
    if (valid && (peers >> lane) == 1ull) s_wh[w][d] = prev + (uint32_t)__popcll(peers);
    kr[i] = k;
    vr[i] = v;
    rr[i] = prev + (uint32_t)__popcll(peers & lt);
  }
  __syncthreads();
  {
    const int d = tid;
    uint32_t t0 = s_wh[0][d], t1 = s_wh[1][d], t2 = s_wh[2][d], t3 = s_wh[3][d];
    s_wh[0][d] = 0u;
    s_wh[1][d] = t0;
    s_wh[2][d] = t0 + t1;
    s_wh[3][d] = t0 + t1 + t2;
    const uint32_t total = t0 + t1 + t2 + t3;
    uint32_t* my = status + (size_t)b * 256 + d;
    uint32_t excl = 0u;
    if (b == 0) {
      __hip_atomic_store(my, FLAG_PREFIX | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(my, FLAG_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // look back over the predecessors LOOKBACK at a time (independent loads), summing published
      // aggregates nearest first until an inclusive prefix is found; an unpublished entry is
      // polled again. Every block publishes its aggregate before it looks back, so the wait ends.
      int bb = b - 1;
      while (true) {
        uint32_t sv[LOOKBACK];
#pragma unroll
        for (int j = 0; j < LOOKBACK; ++j)
          sv[j] = bb - j >= 0 ? __hip_atomic_load(status + (size_t)(bb - j) * 256 + d, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : FLAG_PREFIX;
        int used = 0;
        bool found = false;
#pragma unroll
        for (int j = 0; j < LOOKBACK; ++j) {
          if (found || used < j) continue;           // stop at the first prefix or unpublished entry
          const uint32_t fl = sv[j] & ~VALUE_MASK;
          if (fl == 0u) continue;
          excl += sv[j] & VALUE_MASK;
          used = j + 1;
          if (fl == FLAG_PREFIX) found = true;
        }
        if (found) break;
        bb -= used;
      }
      __hip_atomic_store(my, FLAG_PREFIX | (excl + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // exclusive scans over the digits: global histogram of this pass, and this block's totals
    const uint32_t gh = hist[pass * 256 + d];
    s_scan[d] = gh;
    s_loc[d] = total;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      uint32_t a1 = d >= off ? s_scan[d - off] : 0u;
      uint32_t a2 = d >= off ? s_loc[d - off] : 0u;
      __syncthreads();
      s_scan[d] += a1;
      s_loc[d] += a2;
      __syncthreads();
    }
    const uint32_t loc = s_loc[d] - total;
    __syncthreads();
    s_loc[d] = loc;
    s_off[d] = (s_scan[d] - gh) + excl - loc;  // global destination = s_off[d] + block-local position
  }
  __syncthreads();
  // place the block's keys in LDS in output order, then write contiguous runs
#pragma unroll
  for (int i = 0; i < SORT_IPT; ++i) {
    int pos = base + i * 64 + lane;
    if (pos < n) {
      uint32_t d = (kr[i] >> shift) & 255u;
      uint32_t lp = s_loc[d] + s_wh[w][d] + rr[i];
      if (lp < (uint32_t)SORT_TILE) {
        s_k[lp] = kr[i];
        s_v[lp] = vr[i];
      }
    }
  }
  __syncthreads();
  const int cnt = min(SORT_TILE, n - b * SORT_TILE);
#pragma unroll
  for (int i = 0; i < SORT_IPT; ++i) {
    const int t = i * 256 + tid;
    if (t < cnt) {
      const uint32_t k = s_k[t];
      const uint32_t dst = s_off[(k >> shift) & 255u] + (uint32_t)t;
      if (dst < (uint32_t)n) {  // always true for a consistent histogram; keeps a bad one in bounds
        kout[dst] = k;
        vout[dst] = s_v[t];
      }
    }
  }
}

// ------------------------------------------------------------------ one-digit counting sort
// The cell key itself is the single digit: count the particles of every cell (atomic ranks, in
// no particular order), scan the counts into cell offsets, scatter the caller ids into their
// cells, then order each cell by caller id (the rank of id i in its cell is the number of smaller
// ids there). The result equals the LSD passes' stable sort bit for bit. The count and the scatter
// run over physical slots, where the previous sort left the particles in cell order: the keys are
// read in order, lanes of a wave share few distinct keys, and the scatter's writes land close to
// where they are read; the scatter keeps each id's slot beside it, so that the rank pass finds the
// particle's key and slot without gathering by caller id.
//
// Counting: the lanes of a wave that share a key are grouped first (leader = the lowest such lane);
// then every leader adds its group's size with one atomic, all leaders in the same instruction, so
// a wave waits for one atomic round trip.
//
// With `reset` set (the counting sort, which has no k_reset launch) the kernel also does k_reset's
// work: it returns the previous substep's active tiles and node blocks to -1. The counters those
// lists are read with are zeroed by k_cell_scan.
__global__ __launch_bounds__(256) void k_cell_count(const uint32_t* __restrict__ keys, int n, int* __restrict__ cnt,
                                                    uint32_t* __restrict__ rank, ResetArgs r, int reset,
                                                    uint32_t* __restrict__ occ, uint32_t* __restrict__ occ_clear,
                                                    int nocc, const Clock* clk) {
  if (clock_skip(clk)) return;
  // the key is loaded first, so that its round trip overlaps the reset's
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool valid = i < n;
  const uint32_t key = valid ? keys[i] : 0u;
  // the previous substep's block flags (its P2G has run) start clean for the next one
  for (int q = i; q < nocc; q += gridDim.x * 256) occ_clear[q] = 0u;
  if (reset) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t q = t0; q < r.n_zero; q += stride) r.zero_words[q] = 0u;
    const int nt = r.counters[CNT_TILES], nb = r.counters[CNT_BLOCKS];
    for (int64_t q = t0; q < nt; q += stride) r.tile_slot[r.active_tiles[q]] = -1;
    for (int64_t q = t0; q < nb; q += stride) r.block_slot[r.active_blocks[q]] = -1;
  }
  const int lane = (int)lane_id();
  const uint64_t lt = (1ull << lane) - 1ull;
  uint64_t todo = __ballot(valid);
  int leader = lane, group = 0, within = 0;
  while (todo) {
    const int first = __ffsll((long long)todo) - 1;
    const uint32_t k0 = (uint32_t)__shfl((int)key, first);
    const uint64_t peers = __ballot(valid && key == k0);
    if (valid && key == k0) {
      leader = first;
      group = __popcll(peers);
      within = __popcll(peers & lt);
    }
    todo &= ~peers;
  }
  int base = 0;
  if (valid && leader == lane) {
    base = atomicAdd(&cnt[key], group);
    occ[key / SCAN_TILE] = 1u;
  }
  base = __shfl(base, leader);
  if (valid) rank[i] = (uint32_t)(base + within);
}

// Total of each scan block's SCAN_TILE cell counts (the blocks are independent).
__global__ __launch_bounds__(256) void k_scan_sums(const int* __restrict__ cnt, int64_t ncells, uint32_t* __restrict__ agg,
                                                   const uint32_t* __restrict__ occ, const Clock* clk) {
  if (clock_skip(clk)) return;
  __shared__ uint32_t s_wsum[4];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  if (!occ[blockIdx.x]) {  // no particle in the block's cells: nothing to read
    if (tid == 0) agg[blockIdx.x] = 0u;
    return;
  }
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)tid * SCAN_IPT;
  uint32_t t = 0;
  if (base + SCAN_IPT <= ncells) {
    const int4* p4 = reinterpret_cast<const int4*>(cnt + base);
#pragma unroll
    for (int q = 0; q < SCAN_IPT / 4; ++q) {
      const int4 v = p4[q];
      t += (uint32_t)v.x + (uint32_t)v.y + (uint32_t)v.z + (uint32_t)v.w;
    }
  } else {
    for (int q = 0; q < SCAN_IPT; ++q)
      if (base + q < ncells) t += (uint32_t)cnt[base + q];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
  if (lane == 0) s_wsum[w] = t;
  __syncthreads();
  if (tid == 0) agg[blockIdx.x] = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
}

// Exclusive scan of the cell counts into cell_start[0 .. ncells]: block b starts at the sum of the
// totals of the blocks before it (read and summed by the block itself) and scans its SCAN_TILE
// cells; the blocks are independent. A block without particles writes nothing: its cells are
// empty, which every reader learns from occ (the getter of the bins fills them, k_fill_bins). A
// block with particles also writes the start of the next block's first cell (its own end), which
// the next block, if it has particles, writes with the same value. The last block writes
// cell_start[ncells].
// Block 0 also zeroes the substep's counters (active tiles and node blocks, candidates, apex
// projections), which k_cell_count has read.
__global__ __launch_bounds__(256) void k_cell_scan(const int* __restrict__ cnt, int64_t ncells, int* __restrict__ cell_start,
                                                   const uint32_t* __restrict__ agg, const uint32_t* __restrict__ occ,
                                                   int* __restrict__ counters, const Clock* clk) {
  if (clock_skip(clk)) return;
  if (blockIdx.x == 0 && threadIdx.x < CNT_NUM) counters[threadIdx.x] = 0;
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_boff[4];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int b = blockIdx.x;
  const bool last = (int64_t)(b + 1) * SCAN_TILE >= ncells;
  if (!occ[b] && !last) return;
  {
    uint32_t p = 0;
    for (int q = tid; q < b; q += 256) p += agg[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off);
    if (lane == 0) s_boff[w] = p;
  }
  const int64_t base = (int64_t)b * SCAN_TILE + (int64_t)tid * SCAN_IPT;
  int c[SCAN_IPT];
  if (base + SCAN_IPT <= ncells) {
    const int4* p4 = reinterpret_cast<const int4*>(cnt + base);
#pragma unroll
    for (int q = 0; q < SCAN_IPT / 4; ++q) {
      const int4 v = p4[q];
      c[4 * q] = v.x; c[4 * q + 1] = v.y; c[4 * q + 2] = v.z; c[4 * q + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < SCAN_IPT; ++q) c[q] = base + q < ncells ? cnt[base + q] : 0;
  }
  uint32_t tsum = 0;
#pragma unroll
  for (int q = 0; q < SCAN_IPT; ++q) tsum += (uint32_t)c[q];
  uint32_t x = tsum;  // inclusive scan over the wave's threads
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_wsum[w] = x;
  __syncthreads();
  const uint32_t boff = s_boff[0] + s_boff[1] + s_boff[2] + s_boff[3];
  uint32_t wpre = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) wpre += k < w ? s_wsum[k] : 0u;
  uint32_t run = boff + wpre + (x - tsum);
  if (base + SCAN_IPT <= ncells) {
    int4* ps = reinterpret_cast<int4*>(cell_start + base);
#pragma unroll
    for (int q = 0; q < SCAN_IPT / 4; ++q) {
      int4 vs;
      vs.x = (int)run; run += (uint32_t)c[4 * q];
      vs.y = (int)run; run += (uint32_t)c[4 * q + 1];
      vs.z = (int)run; run += (uint32_t)c[4 * q + 2];
      vs.w = (int)run; run += (uint32_t)c[4 * q + 3];
      ps[q] = vs;
    }
  } else {
#pragma unroll
    for (int q = 0; q < SCAN_IPT; ++q) {
      if (base + q < ncells) {
        cell_start[base + q] = (int)run;
        run += (uint32_t)c[q];
      }
    }
  }
  if (base < ncells && base + SCAN_IPT >= ncells) cell_start[ncells] = (int)run;  // the total
  if (tid == 255 && !last) cell_start[(int64_t)(b + 1) * SCAN_TILE] = (int)run;  // the next block's first cell
}

// The bins of the last sort for the host: cell_start[c] = the first sorted position whose key is at
// least c (a binary search of the sorted keys), for every cell c in [0, ncells]. The sort's own scan
// writes only the scan blocks that hold particles, which is all the substep reads; the keys, unlike
// the block flags, belong to the last sort that ran, also when the device clock skipped a later
// captured substep.
__global__ __launch_bounds__(256) void k_fill_bins(const uint32_t* __restrict__ ks, int n, int64_t ncells,
                                                   int* __restrict__ cell_start) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c > ncells) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if ((int64_t)ks[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  cell_start[c] = lo;
}

// The caller id -> physical slot map from the state's id plane and, when kc is given, the keys by
// caller id from the keys by physical slot (the input of the radix passes).
__global__ __launch_bounds__(256) void k_caller_keys(const uint32_t* __restrict__ keys, const int* __restrict__ pid, int n,
                                                     uint32_t* __restrict__ kc, int* __restrict__ pos, const Clock* clk) {
  if (clock_skip(clk)) return;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int i = pid[j];
  if ((unsigned)i < (unsigned)n) {  // always true for a consistent id plane
    if (kc) kc[i] = keys[j];
    pos[i] = j;
  }
}

__global__ __launch_bounds__(256) void k_cell_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ rank,
                                                      int n, const int* __restrict__ cell_start,
                                                      const int* __restrict__ pid, int* __restrict__ tmp,
                                                      int* __restrict__ tslot, const Clock* clk) {
  if (clock_skip(clk)) return;
  const int j = blockIdx.x * 256 + threadIdx.x;  // physical slot
  if (j >= n) return;
  const int p = cell_start[keys[j]] + (int)rank[j];
  if ((unsigned)p < (unsigned)n) {  // always true for consistent counts
    tmp[p] = pid[j];
    tslot[p] = j;
  }
}

// Thread per sorted position: the caller id placed there by the scatter moves to its stable rank
// within the cell (a cell of c particles costs c comparisons per particle). It also does the bin
// pass's work (k_bounds) for its new position: src = the particle's physical slot, and the cell's
// first particle clears the cell's count and flags the cell's tile.
// The workgroup first copies the ids of its 256 positions and RANK_HALO on either side into LDS (one
// coalesced round trip): a cell of at most RANK_HALO particles around any of its positions lies in
// that window, and its ids are compared there; a larger cell reads them from global memory.
constexpr int RANK_HALO = 64;
__global__ __launch_bounds__(256) void k_cell_rank(SceneTable T, const uint32_t* __restrict__ keys,
                                                   const int* __restrict__ tmp, const int* __restrict__ tslot, int n,
                                                   const int* __restrict__ cell_start, uint32_t* __restrict__ ks,
                                                   int* __restrict__ vs, int* __restrict__ cnt,
                                                   int* __restrict__ tile_slot, int* __restrict__ src, const Clock* clk) {
  if (clock_skip(clk)) return;
  constexpr int WIN = 256 + 2 * RANK_HALO;
  __shared__ int s_ids[WIN];
  const int wlo = (int)blockIdx.x * 256 - RANK_HALO;  // the window holds positions wlo .. wlo + WIN - 1
  for (int q = (int)threadIdx.x; q < WIN; q += 256) {
    const int p = wlo + q;
    s_ids[q] = (p >= 0 && p < n) ? tmp[p] : 0;
  }
  __syncthreads();
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int i = clamp_index(s_ids[j - wlo], n);
  // the particle's physical slot (the scatter kept it beside the id) and, through it, its key: the
  // slots of a cell's particles are nearly consecutive, so these reads are close to coalesced
  const int pi = clamp_index(tslot[j], n);
  const uint32_t k = keys[pi];
  const int s0 = cell_start[k], e0 = min(cell_start[k + 1], n);
  int r = 0;
  if (s0 >= wlo && e0 <= wlo + WIN) {
    for (int m = s0 - wlo; m < e0 - wlo; ++m) r += s_ids[m] < i ? 1 : 0;
  } else {
    // 8 ids per step, loaded together (past the cell's end the loads read its last entry, in
    // bounds, and are not counted)
    for (int m0 = s0; m0 < e0; m0 += 8) {
      int t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = tmp[min(m0 + u, e0 - 1)];
#pragma unroll
      for (int u = 0; u < 8; ++u) r += (m0 + u < e0 && t[u] < i) ? 1 : 0;
    }
  }
  const int d = s0 + r;
  if ((unsigned)d >= (unsigned)n) return;  // never, for consistent counts
  vs[d] = i;
  ks[d] = k;
  src[d] = pi;
  if (r != 0) return;
  cnt[k] = 0;
  const int sc = scene_of(T.part_off, T.nscenes, i);
  const DevParams& P = T.params[sc];
  const uint32_t lk = k - (uint32_t)T.node_off[sc];
  const int r01 = P.res[0] * P.res[1];
  const int c2 = (int)(lk / (uint32_t)r01), rem = (int)(lk % (uint32_t)r01);
  const int c1 = rem / P.res[0], c0 = rem % P.res[0];
  tile_slot[T.tile_off[sc] + ((c2 >> 2) * P.tiles[1] + (c1 >> 2)) * P.tiles[0] + (c0 >> 2)] = -2;
}

// ================================================================== bins and tiles
// Thread per sorted position: the first position of each run of equal keys clears the cell's count
// for the next substep and flags the cell's tile (plain stores of one value).
__global__ __launch_bounds__(256) void k_bounds(SceneTable T, const uint32_t* __restrict__ ks, int n,
                                                int* __restrict__ cnt, int* __restrict__ tile_slot,
                                                const int* __restrict__ perm, const int* __restrict__ pos,
                                                int* __restrict__ src, const Clock* clk) {
  if (clock_skip(clk)) return;
  const PBlock b = particle_block(T);  // sorted positions are scene-major as well
  const int j = b.j;
  if (j >= b.end) return;
  const DevParams& P = T.params[b.s];
  src[j] = clamp_index(pos[clamp_index(perm[j], n)], n);
  const uint32_t k = ks[j];
  if (j != 0 && ks[j - 1] == k) return;
  cnt[k] = 0;
  const uint32_t lk = k - (uint32_t)T.node_off[b.s];
  const int r01 = P.res[0] * P.res[1];
  const int c2 = (int)(lk / (uint32_t)r01), rem = (int)(lk % (uint32_t)r01);
  const int c1 = rem / P.res[0], c0 = rem % P.res[0];
  tile_slot[T.tile_off[b.s] + ((c2 >> 2) * P.tiles[1] + (c1 >> 2)) * P.tiles[0] + (c0 >> 2)] = -2;
}

// Thread per tile t (tiles and node blocks share one index space). A flagged tile (-2) receives the
// next slot of the active tile list. Node block t is active when any of its 27 neighbour tiles is
// (the 7x7x7-node P2G footprint of those tiles covers it) and receives the next slot of the active
// block list. A workgroup takes one atomic per list. While one thread reads a neighbour's entry,
// another workgroup may be turning its -2 into a slot: both values mean "active", so the result does
// not depend on the timing. Slot numbers depend on the order in which waves run; no result depends
// on them.
__global__ __launch_bounds__(256) void k_claim(SceneTable T, int64_t count, int* __restrict__ tile_slot,
                                               int* __restrict__ active_tiles, int* __restrict__ block_slot,
                                               int* __restrict__ active_blocks, int* __restrict__ counters,
                                               const Clock* clk) {
  if (clock_skip(clk)) return;
  __shared__ int s_cnt[2][4];
  __shared__ int s_base[2];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool tile_on = false, block_on = false;
  if (t < count) {
    const int sc = scene_of(T.tile_off, T.nscenes, (int)t);
    const int* tl = T.params[sc].tiles;
    const int off = T.tile_off[sc], lt = (int)t - off;
    const int t0 = lt % tl[0], t1 = (lt / tl[0]) % tl[1], t2 = lt / (tl[0] * tl[1]);
    int v[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) {  // all 27 loads issued together
      const int b0 = t0 + q % 3 - 1, b1 = t1 + (q / 3) % 3 - 1, b2 = t2 + q / 9 - 1;
      const bool in = b0 >= 0 && b1 >= 0 && b2 >= 0 && b0 < tl[0] && b1 < tl[1] && b2 < tl[2];
      v[q] = in ? tile_slot[off + (b2 * tl[1] + b1) * tl[0] + b0] : -1;
    }
    tile_on = v[13] == -2;  // only this thread writes tile t
#pragma unroll
    for (int q = 0; q < 27; ++q) block_on |= v[q] != -1;
  }
  const uint64_t mt = __ballot(tile_on), mb = __ballot(block_on);
  const int lane = (int)lane_id(), w = (int)threadIdx.x >> 6;
  if (lane == 0) {
    s_cnt[0][w] = __popcll(mt);
    s_cnt[1][w] = __popcll(mb);
  }
  __syncthreads();
  // one atomic per list and workgroup (a single counter takes many device-scope atomics slowly across XCDs)
  if (threadIdx.x < 2) {
    const int k = (int)threadIdx.x;
    const int total = s_cnt[k][0] + s_cnt[k][1] + s_cnt[k][2] + s_cnt[k][3];
    s_base[k] = total ? atomicAdd(&counters[k == 0 ? CNT_TILES : CNT_BLOCKS], total) : 0;
  }
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
  if (tile_on) {
    int base = s_base[0];
    for (int k = 0; k < w; ++k) base += s_cnt[0][k];
    const int slot = base + __popcll(mt & below);
    tile_slot[t] = slot;
    active_tiles[slot] = (int)t;
  }
  if (block_on) {
    int base = s_base[1];
    for (int k = 0; k < w; ++k) base += s_cnt[1][k];
    const int slot = base + __popcll(mb & below);
    block_slot[t] = slot;
    active_blocks[slot] = (int)t;
  }
}

// ================================================================== P2G
// One workgroup (8 waves) per active 4x4x4-cell tile; wave w owns the cell layer z = w & 3 and walks
// rows (y) 2 (w >> 2) and 2 (w >> 2) + 1 one after another, with no workgroup barrier until the
// tile is done. (Two row halves per layer halve a tile's latency, which evens out the last round
// when the active tiles do not divide evenly over the resident workgroups.)
//
// Row pass. The four cells of a row (x = 0..3) are contiguous in the sorted order. In chunk k,
// lane (cx, r), r < RANKS, stages the particle of rank RANKS k + r of cell cx: it loads the
// particle (one chunk ahead) and writes a 48-float payload (B-spline weights of the window's 4 nodes
// per axis, mass, affine momentum terms, stress impulse, and the x-column terms every lane of the
// cell needs) into the wave's LDS slots. Then lane (cx, zo, oy) adds the chunk's particles of cell
// cx, in ascending rank, to its 4 register sums: nodes (cx-1+ox, cy-1+oy, z-1+zo), ox = 0..3, of the
// cell's 4x4x4 node window c-1 .. c+2 (a superset of the quadratic stencil and exactly the cubic
// one). At the end of the row the sums are added into the wave's own LDS node slab (7 x 7 x 4 nodes)
// in four phases ox = 0..3; within a phase every lane has a different node, so the slab receives
// the row sums in a fixed order.
//
// When all waves are done, node (X, Y, Z) of the tile's 7x7x7 footprint is the sum of the slabs of
// the waves of layers l with 0 <= Z - l <= 3, in ascending l and row half, and is stored as the
// tile's partial. k_grid sums the covering tiles' partials in ascending tile order: every node value
// is a sum in a fixed order, without float atomics on global memory, and reruns are bitwise identical.
constexpr int PAY = 48;                 // payload floats per staged particle (12 x float4)
// LDS floats between consecutive ranks' payloads: 52 = 20 (mod 32), so the 128-bit payload writes of
// the eight staging lanes of a cell fall on distinct groups of the 32 banks (48 would give 4-way
// bank conflicts).
constexpr int PAY_STRIDE = 52;
// RANKS (2, 4 or 8, chosen by the autotuner) = ranks per cell staged per chunk; one cell's staged
// payloads take RANKS * PAY_STRIDE + 4 floats (the 4 shift the cells of a row onto different banks).
// The chunk size does not change the order in which a lane adds its cell's particles, so results do
// not depend on it.
template <int RANKS>
constexpr int pay_cell() { return RANKS * PAY_STRIDE + 4; }
constexpr int SLAB = FOOT * FOOT * 4;   // nodes of a wave's slab
constexpr int P2G_WAVES = 8;            // waves per P2G workgroup: 4 cell layers x 2 row halves
// payload layout (floats): the x weights of window nodes c-1 .. c+2, the y and z weights and
// derivatives (per cell width), the x-column terms g[ox][r] = ox wx[ox] Q[r][0] - dwx[ox] S[r][0]
// that every lane of the cell adds with weight wy wz, the mass, q0 (a cloth vertex's force in the
// force pass), and the y and z columns of Q and of the scaled stress matrix S.
constexpr int PY_WX = 0, PY_WY = 4, PY_DWY = 8, PY_WZ = 12, PY_DWZ = 16, PY_G = 20;
constexpr int PY_M = 32, PY_Q0 = 33, PY_Q12 = 36, PY_S12 = 42;

// Weights and derivatives (per cell width) of the window nodes c-1 .. c+2 of a particle at fraction
// f of cell c. Quadratic: the three nodes from base = floor(c + f - 0.5) (c - 1 for f < 0.5, else c),
// fx = c + f - base in [0.5, 1.5): N = 0.5 (1.5 - fx)^2, 0.75 - (fx - 1)^2, 0.5 (fx - 0.5)^2, in
// closed form and without branches (HybridSolver.cpp:18-97, interpolation.cpp:9-33).
template <int ORDER>
__device__ __forceinline__ void window_weights(float f, float* w, float* dw) {
  if constexpr (ORDER == 2) {
    const bool hi = f >= 0.5f;
    const float fx = hi ? f : f + 1.0f;
    const float a = 1.5f - fx, b = fx - 1.0f, c = fx - 0.5f;
    const float w0 = 0.5f * a * a, w1 = 0.75f - b * b, w2 = 0.5f * c * c;
    const float d0 = -a, d1 = -2.0f * b, d2 = c;
    w[0] = hi ? 0.0f : w0;
    w[1] = hi ? w0 : w1;
    w[2] = hi ? w1 : w2;
    w[3] = hi ? w2 : 0.0f;
    dw[0] = hi ? 0.0f : d0;
    dw[1] = hi ? d0 : d1;
    dw[2] = hi ? d1 : d2;
    dw[3] = hi ? d2 : 0.0f;
  } else {
#pragma unroll
    for (int o = 0; o < 4; ++o) spline<ORDER>(f + 1.0f - (float)o, w[o], dw[o]);
  }
}

// Scene constants of the tile being transferred, held in registers for the whole tile.
struct P2GScene {
  float gmin[3], inv_h[3], h[3], dinv;
  int res[3];
  int mesh_v0, mesh_e0, mesh_e1;
};


__device__ __forceinline__ P2GScene p2g_scene(const DevParams& P) {
  P2GScene S;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    S.gmin[d] = uniform_f(P.gmin[d]);
    S.inv_h[d] = uniform_f(P.inv_h[d]);
    S.h[d] = uniform_f(P.h[d]);
    S.res[d] = __builtin_amdgcn_readfirstlane(P.res[d]);
  }
  S.dinv = uniform_f(P.dinv);
  S.mesh_v0 = __builtin_amdgcn_readfirstlane(P.mesh_v0);
  S.mesh_e0 = __builtin_amdgcn_readfirstlane(P.mesh_e0);
  S.mesh_e1 = __builtin_amdgcn_readfirstlane(P.mesh_e1);
  return S;
}

// Particle data a staging lane needs, loaded one chunk ahead of its use (the loads of the next
// chunk are in flight while the wave accumulates the current one). The force pass of the
// predicted-stress mode loads inside p2g_stage instead.
struct StageIn {
  int pi, pid;  // pid: the particle's id when the scene has a cloth mesh (else -1)
  float x[3], m, v[3], B[9], K[6];
};

// The physical slot of sorted position idx, loaded a chunk before the particle's own data so that
// the two dependent loads do not stall the wave back to back; it is clamped where it is used.
__device__ __forceinline__ int p2g_slot(const P2GArgs& a, int idx) { return a.src[idx]; }

template <int MODE>
__device__ __forceinline__ void p2g_load_at(const P2GArgs& a, const P2GScene& S, int slot, StageIn& in) {
  const int pi = clamp_index(slot, a.n);
  const float* st = a.state;
  const int64_t cap = a.cap;
  in.pi = pi;
  // a scene without a mesh loads word 0 of the id plane in every lane (one cache line) instead of a
  // branch around the load: a branch here makes the wave wait for this chunk's loads before the
  // accumulation of the previous one
  in.pid = reinterpret_cast<const int*>(st + (size_t)PID * cap)[S.mesh_e1 > S.mesh_v0 ? pi : 0];
#pragma unroll
  for (int d = 0; d < 3; ++d) in.x[d] = st[(PX + d) * cap + pi];
  if constexpr (MODE != P2G_FORCE) {
    in.m = st[PM * cap + pi];
#pragma unroll
    for (int d = 0; d < 3; ++d) in.v[d] = st[(PV + d) * cap + pi];
#pragma unroll
    for (int i = 0; i < 9; ++i) in.B[i] = st[(PB + i) * cap + pi];
  }
  if constexpr (MODE == P2G_MOM_STRESS) {
#pragma unroll
    for (int i = 0; i < 6; ++i) in.K[i] = st[(PK + i) * cap + pi];
  }
}

template <int MODE>
__device__ __forceinline__ void p2g_load(const P2GArgs& a, const P2GScene& S, int idx, StageIn& in) {
  p2g_load_at<MODE>(a, S, p2g_slot(a, idx), in);
}

template <int ORDER, int MODE>
__device__ __forceinline__ void p2g_stage(const P2GArgs& a, const P2GScene& S, const DevParams& P, int64_t noff,
                                          float dt, const StageIn& in, int g0, int g1, int g2, float4* dst) {
  // No FMA contraction: the row and the layer pass inline this function into different kernels, and
  // the payload must have the same bits in both (contraction is decided per kernel otherwise).
#pragma clang fp contract(off)
  float pay[PAY];
#pragma unroll
  for (int i = 0; i < PAY; ++i) pay[i] = 0.0f;
  // kind of the point from its id (loaded with the rest of its data, used only here)
  const int pi = in.pi;
  const int kind = (S.mesh_e1 <= S.mesh_v0 || in.pid < S.mesh_v0) ? KIND_PARTICLE
                                                                    : (in.pid < S.mesh_e0 ? KIND_VERTEX : KIND_ELEMENT);
  const int mid = in.pid - (kind == KIND_VERTEX ? S.mesh_v0 : S.mesh_e0);
  const float* st = a.state;
  const int64_t cap = a.cap;
  const int cc[3] = {g0, g1, g2};
  float f[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) f[d] = (in.x[d] - S.gmin[d]) * S.inv_h[d] - (float)cc[d];
  float wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
  window_weights<ORDER>(f[0], wx, dwx);
  window_weights<ORDER>(f[1], wy, dwy);
  window_weights<ORDER>(f[2], wz, dwz);
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    pay[PY_WX + o] = wx[o];
    pay[PY_WY + o] = wy[o];
    pay[PY_DWY + o] = dwy[o];
    pay[PY_WZ + o] = wz[o];
    pay[PY_DWZ + o] = dwz[o];
  }
  float Q0[3] = {0.0f, 0.0f, 0.0f}, S0[3] = {0.0f, 0.0f, 0.0f};  // x columns of Q and of S
  if constexpr (MODE != P2G_FORCE) {
    // momentum w*(q0 + Q o): Q = m D^-1 B diag(h), q0 = m v - Q (1 + f)  (HybridSolver.cpp:134,169-170)
    const float mp = in.m;
    float Q[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) Q[3 * r + c] = mp * S.dinv * in.B[3 * r + c] * S.h[c];
    pay[PY_M] = mp;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      float q0 = mp * in.v[r] - (Q[3 * r] * (1.0f + f[0]) + Q[3 * r + 1] * (1.0f + f[1]) + Q[3 * r + 2] * (1.0f + f[2]));
      if (MODE == P2G_MOM_STRESS && kind == KIND_VERTEX) q0 += dt * a.vert_force[3 * mid + r];  // in-plane impulse (:378)
      pay[PY_Q0 + r] = q0;
      Q0[r] = Q[3 * r];
      pay[PY_Q12 + 2 * r] = Q[3 * r + 1];
      pay[PY_Q12 + 2 * r + 1] = Q[3 * r + 2];
    }
  }
  float Sm[9];  // stress matrix, columns scaled by 1/h (and by dt when folded into the momentum)
#pragma unroll
  for (int i = 0; i < 9; ++i) Sm[i] = 0.0f;
  if constexpr (MODE == P2G_MOM_STRESS) {
    // stress at F^n: the Kirchhoff stress V P(F) F^T written by the previous G2P (or by k_stress_init)
    const int R[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};  // symmetric 3x3 from (xx yy zz xy xz yz)
    if (kind == KIND_PARTICLE) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Sm[3 * r + c] = dt * in.K[R[3 * r + c]] * S.inv_h[c];
    } else if (kind == KIND_ELEMENT) {  // element stress of the normal/shear response (:434-454)
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Sm[3 * r + c] = dt * a.elem_stress[9 * mid + 3 * r + c] * S.inv_h[c];
    }
  }
  if constexpr (MODE == P2G_FORCE) {
    if (kind == KIND_VERTEX) {
#pragma unroll
      for (int r = 0; r < 3; ++r) pay[PY_Q0 + r] = a.vert_force[3 * mid + r];
    } else if (kind == KIND_ELEMENT) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Sm[3 * r + c] = a.elem_stress[9 * mid + 3 * r + c] * S.inv_h[c];
    } else {
      // stress matrix V P(F^) F^n^T at the predicted F^ = F + dt (sum_i v_i grad w_ip^T) F from the
      // post-P2G grid velocity (HybridSolver.cpp:269-339)
      float Fn[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) Fn[i] = st[(PF + i) * cap + pi];
      float G[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) G[i] = 0.0f;
      for (int oz = 0; oz < 4; ++oz) {
        int n2 = g2 - 1 + oz;
        if (n2 < 0 || n2 >= S.res[2]) continue;
        for (int oy = 0; oy < 4; ++oy) {
          int n1 = g1 - 1 + oy;
          if (n1 < 0 || n1 >= S.res[1]) continue;
          for (int ox = 0; ox < 4; ++ox) {
            int n0 = g0 - 1 + ox;
            if (n0 < 0 || n0 >= S.res[0]) continue;
            float gw[3] = {dwx[ox] * wy[oy] * wz[oz] * S.inv_h[0], wx[ox] * dwy[oy] * wz[oz] * S.inv_h[1],
                           wx[ox] * wy[oy] * dwz[oz] * S.inv_h[2]};
            int64_t node = noff + ((int64_t)n2 * S.res[1] + n1) * S.res[0] + n0;
            const float4 gv4 = a.gvel[2 * node];
            const float vr[3] = {gv4.x, gv4.y, gv4.z};
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
              for (int c = 0; c < 3; ++c) G[3 * r + c] += vr[r] * gw[c];
          }
        }
      }
      float GF[9], Fh[9];
      mat_mul(G, Fn, GF);
#pragma unroll
      for (int i = 0; i < 9; ++i) Fh[i] = Fn[i] + dt * GF[i];
      float Jp = st[PPL * cap + pi];
      float vol = st[PVOL * cap + pi];
      float Sp[9];
      int fb = 0;
      stress_matrix(P, Fh, Fn, Jp, vol, Sp, fb);
      if (fb) atomicAdd(a.fallback_svd, (unsigned long long)fb);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Sm[3 * r + c] = Sp[3 * r + c] * S.inv_h[c];
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    S0[r] = Sm[3 * r];
    pay[PY_S12 + 2 * r] = Sm[3 * r + 1];
    pay[PY_S12 + 2 * r + 1] = Sm[3 * r + 2];
  }
#pragma unroll
  for (int ox = 0; ox < 4; ++ox)
#pragma unroll
    for (int r = 0; r < 3; ++r) pay[PY_G + 3 * ox + r] = (float)ox * wx[ox] * Q0[r] - dwx[ox] * S0[r];
#pragma unroll
  for (int i = 0; i < PAY / 4; ++i) dst[i] = make_float4(pay[4 * i], pay[4 * i + 1], pay[4 * i + 2], pay[4 * i + 3]);
}

template <int ORDER, int MODE, int RANKS>
__global__ __launch_bounds__(64 * P2G_WAVES) void k_p2g(P2GArgs a) {
  if (clock_skip(a.clk)) return;
  const float dt = clock_dt(a.clk, a.dt);
  constexpr int P2G_RANKS = RANKS;
  constexpr int PAY_CELL = pay_cell<RANKS>();
  __shared__ float4 s_pay[P2G_WAVES][4 * PAY_CELL / 4];
  __shared__ float s_slab[P2G_WAVES][4][SLAB];  // [wave][quantity][node]
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int layer = w & 3, half = w >> 2;  // wave w owns cell layer z = w & 3, rows 2 (w >> 2) and 2 (w >> 2) + 1
  const int ntile = a.counters[CNT_TILES];
  constexpr int NQ = (MODE == P2G_FORCE) ? 3 : 4;
  // staging role: lane (cx, sr), sr < RANKS; accumulation role: lane (cx, zo, oy)
  const int cx = lane >> 4, sr = lane & 15;
  const int zo = (lane >> 2) & 3, oy = lane & 3;
  const float foy = (float)oy, fzo = (float)zo;
  float* slab = &s_slab[w][0][0];
  for (int i = lane; i < 4 * SLAB; i += 64) slab[i] = 0.0f;
  // a context with one scene reads its constants once; the next tile's index is loaded a tile ahead
  const bool one_scene = a.T.nscenes == 1;
  P2GScene S1{};
  int64_t noff1 = 0;
  int toff1 = 0, nt0_1 = 1, nt1_1 = 1;
  if (one_scene) {
    S1 = p2g_scene(a.T.params[0]);
    noff1 = a.T.node_off[0];
    toff1 = a.T.tile_off[0];
    nt0_1 = __builtin_amdgcn_readfirstlane(a.T.params[0].tiles[0]);
    nt1_1 = __builtin_amdgcn_readfirstlane(a.T.params[0].tiles[1]);
  }
  // the tiles of one XCD are a contiguous range of the (ascending) active tiles
  const XcdRange tr = xcd_range(ntile);
  int gtile_next = tr.first < tr.end ? a.active_tiles[tr.first] : 0;
  for (int slot = tr.first; slot < tr.end; slot += tr.stride) {
    const int gtile = __builtin_amdgcn_readfirstlane(gtile_next);
    if (slot + tr.stride < tr.end) gtile_next = a.active_tiles[slot + tr.stride];
    const int sc = one_scene ? 0 : __builtin_amdgcn_readfirstlane(scene_of(a.T.tile_off, a.T.nscenes, gtile));
    const DevParams& P = a.T.params[sc];
    const P2GScene S = one_scene ? S1 : p2g_scene(P);
    const int64_t noff = one_scene ? noff1 : a.T.node_off[sc];
    const int tile = gtile - (one_scene ? toff1 : a.T.tile_off[sc]);
    const int nt0 = one_scene ? nt0_1 : __builtin_amdgcn_readfirstlane(P.tiles[0]);
    const int nt1 = one_scene ? nt1_1 : __builtin_amdgcn_readfirstlane(P.tiles[1]);
    const int t0 = tile % nt0, t1 = (tile / nt0) % nt1, t2 = tile / (nt0 * nt1);
    const int g2 = 4 * t2 + layer;
    const int g0 = 4 * t0 + cx;  // this lane's cell of each row (both roles)
    // the bins of this lane's cell in the wave's two rows, and each row's largest count (uniform)
    int rstart0 = 0, rcnt0 = 0, rstart1 = 0, rcnt1 = 0;
    {
      const int g1a = 4 * t1 + 2 * half, g1b = g1a + 1;
      if (g0 < S.res[0] && g2 < S.res[2]) {
        if (g1a < S.res[1]) {
          const int64_t key = noff + (g2 * S.res[1] + g1a) * S.res[0] + g0;
          if (a.occ[key / SCAN_TILE]) {
            rstart0 = a.cell_start[key];
            rcnt0 = a.cell_end[key] - rstart0;
          }
        }
        if (g1b < S.res[1]) {
          const int64_t key = noff + (g2 * S.res[1] + g1b) * S.res[0] + g0;
          if (a.occ[key / SCAN_TILE]) {
            rstart1 = a.cell_start[key];
            rcnt1 = a.cell_end[key] - rstart1;
          }
        }
      }
    }
    const int rmax0 = max(max(__builtin_amdgcn_readlane(rcnt0, 0), __builtin_amdgcn_readlane(rcnt0, 16)),
                          max(__builtin_amdgcn_readlane(rcnt0, 32), __builtin_amdgcn_readlane(rcnt0, 48)));
    const int rmax1 = max(max(__builtin_amdgcn_readlane(rcnt1, 0), __builtin_amdgcn_readlane(rcnt1, 16)),
                          max(__builtin_amdgcn_readlane(rcnt1, 32), __builtin_amdgcn_readlane(rcnt1, 48)));
    auto rmax_of = [&](int r) { return r == 0 ? rmax0 : rmax1; };
    auto rcnt_of = [&](int r) { return r == 0 ? rcnt0 : rcnt1; };
    auto rstart_of = [&](int r) { return r == 0 ? rstart0 : rstart1; };
    // walk the (row, chunk) steps in order (row 2 = done), loading the particles ahead
    auto next_row = [&](int r) { return (r == 0 && rmax0 > 0) ? 0 : ((r <= 1 && rmax1 > 0) ? 1 : 2); };
    auto advance = [&](int& r, int& kk) {
      if (r >= 2) return;
      if (P2G_RANKS * (kk + 1) >= rmax_of(r)) {
        r = next_row(r + 1);
        kk = 0;
      } else {
        ++kk;
      }
    };
    int cr = next_row(0), k = 0;
    StageIn in;
    // slot of this lane's particle two steps ahead (step after the next)
    int cr2 = cr, k2 = k;
    advance(cr2, k2);
    int pi2 = 0;
    if (MODE != P2G_FORCE && cr < 2 && sr < P2G_RANKS && sr < rcnt_of(cr)) p2g_load<MODE>(a, S, rstart_of(cr) + sr, in);
    if (MODE != P2G_FORCE && cr2 < 2 && sr < P2G_RANKS && P2G_RANKS * k2 + sr < rcnt_of(cr2))
      pi2 = p2g_slot(a, rstart_of(cr2) + P2G_RANKS * k2 + sr);
    float acc[4][NQ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int q = 0; q < NQ; ++q) acc[i][q] = 0.0f;
    while (cr < 2) {
      const int cy = 2 * half + cr;
      const int g1 = 4 * t1 + cy;
      const int cnt = rcnt_of(cr), rstart = rstart_of(cr);
      // ---- stage ranks RANKS k .. RANKS k + RANKS - 1 of each cell of the row
      if (sr < P2G_RANKS) {
        const int rank = P2G_RANKS * k + sr;
        if (rank < cnt) {
          if constexpr (MODE == P2G_FORCE) p2g_load<MODE>(a, S, rstart + rank, in);
          p2g_stage<ORDER, MODE>(a, S, P, noff, dt, in, g0, g1, g2, &s_pay[w][(cx * PAY_CELL + sr * PAY_STRIDE) / 4]);
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      // ---- the next step: its particles' data (their slots arrived during this step), and the
      // slots of the step after it
      const int ncr = cr2, nk = k2;
      if (MODE != P2G_FORCE && ncr < 2 && sr < P2G_RANKS && P2G_RANKS * nk + sr < rcnt_of(ncr))
        p2g_load_at<MODE>(a, S, pi2, in);
      advance(cr2, k2);
      if (MODE != P2G_FORCE && cr2 < 2 && sr < P2G_RANKS && P2G_RANKS * k2 + sr < rcnt_of(cr2))
        pi2 = p2g_slot(a, rstart_of(cr2) + P2G_RANKS * k2 + sr);
      // ---- accumulate this lane's cell, ascending rank
      const int nr = min(P2G_RANKS, cnt - P2G_RANKS * k);
      for (int r = 0; r < nr; ++r) {
        const float4* src = &s_pay[w][(cx * PAY_CELL + r * PAY_STRIDE) / 4];
        const float* pf = reinterpret_cast<const float*>(src);
        const float4 WX = src[0];
        const float wxs[4] = {WX.x, WX.y, WX.z, WX.w};
        const float wy = pf[PY_WY + oy], dwy = pf[PY_DWY + oy], wz = pf[PY_WZ + zo], dwz = pf[PY_DWZ + zo];
        float g[12];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const float4 t = src[PY_G / 4 + i];
          g[4 * i] = t.x; g[4 * i + 1] = t.y; g[4 * i + 2] = t.z; g[4 * i + 3] = t.w;
        }
        float qs[16];  // m, q0(3), Q12(6), S12(6)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float4 t = src[PY_M / 4 + i];
          qs[4 * i] = t.x; qs[4 * i + 1] = t.y; qs[4 * i + 2] = t.z; qs[4 * i + 3] = t.w;
        }
        const float wyz = wy * wz, dyz = dwy * wz, ydz = wy * dwz;
        float e[3];
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) {
          float qrow = qs[1 + rr];
          if constexpr (MODE != P2G_FORCE)
            qrow = __builtin_fmaf(qs[4 + 2 * rr + 1], fzo, __builtin_fmaf(qs[4 + 2 * rr], foy, qrow));
          e[rr] = wyz * qrow;
          if constexpr (MODE != P2G_MOM)
            e[rr] = __builtin_fmaf(-qs[10 + 2 * rr + 1], ydz, __builtin_fmaf(-qs[10 + 2 * rr], dyz, e[rr]));
        }
        const float mw = qs[0] * wyz;
#pragma unroll
        for (int ox = 0; ox < 4; ++ox) {
          const float wx = wxs[ox];
          constexpr int q1 = (MODE != P2G_FORCE) ? 1 : 0;
          if constexpr (MODE != P2G_FORCE) acc[ox][0] = __builtin_fmaf(wx, mw, acc[ox][0]);
#pragma unroll
          for (int rr = 0; rr < 3; ++rr)
            acc[ox][q1 + rr] = __builtin_fmaf(wyz, g[3 * ox + rr], __builtin_fmaf(wx, e[rr], acc[ox][q1 + rr]));
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (ncr != cr) {
        // ---- row sums into the wave's slab: phase ox, node (cx + ox, cy + oy, zo)
#pragma unroll
        for (int ox = 0; ox < 4; ++ox) {
          const int ln = (cx + ox) + FOOT * (cy + oy) + FOOT * FOOT * zo;
#pragma unroll
          for (int q = 0; q < NQ; ++q) {  // the wave's own slab, a different node per lane: no atomics
            slab[q * SLAB + ln] += acc[ox][q];
            acc[ox][q] = 0.0f;
          }
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
      }
      cr = ncr;
      k = nk;
    }
    __syncthreads();
    // ---- tile partial: node (X, Y, Z) = sum over layers l, 0 <= Z - l <= 3, ascending, and within a
    // layer over the row halves in order; slabs cleared
    float* part = a.partials + (size_t)slot * PART_STRIDE + (MODE == P2G_FORCE ? 4 * FOOT3 : 0);
    for (int i = tid; i < FOOT3; i += 64 * P2G_WAVES) {
      const int X = i % FOOT, Y = (i / FOOT) % FOOT, Z = i / (FOOT * FOOT);
      float s[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) s[q] = 0.0f;
#pragma unroll
      for (int ww = 0; ww < P2G_WAVES; ++ww) {
        const int l = ww >> 1, hh = ww & 1, wv = l + 4 * hh;
        const int z = Z - l;
        if (z < 0 || z > 3) continue;
        const int ln = X + FOOT * Y + FOOT * FOOT * z;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          s[q] += s_slab[wv][q][ln];
          s_slab[wv][q][ln] = 0.0f;
        }
      }
      if constexpr (NQ == 4) {
        reinterpret_cast<float4*>(part)[i] = make_float4(s[0], s[1], s[2], s[3]);
      } else {
#pragma unroll
        for (int q = 0; q < NQ; ++q) part[q * FOOT3 + i] = s[q];
      }
    }
    __syncthreads();
  }
}

// Layer pass (the alternative P2G schedule, chosen by the autotuner). One workgroup (4 waves) per
// active tile; wave w owns cell layer z = w and all 16 cells of it at once. Staging lane (c, r),
// c = lane >> 2 (cell x = c & 3, y = c >> 2), r < RANKS, stages rank RANKS k + r of cell c; then
// lane (c, zo) adds the chunk's particles of cell c, in ascending rank, to 16 x 4 register sums:
// nodes (x-1+ox, y-1+oy, z-1+zo) of the cell's 4x4x4 window. Four lanes read each payload, instead
// of the row pass's sixteen. The arithmetic per particle and node, the split of a layer into two
// row halves with a slab each, the order in which a slab node receives the cell sums (rows
// ascending, then ox ascending: phases oy = 3..0, ox = 0..3) and the order of the slab sums are
// those of the row pass, so both passes give the same bits.
template <int ORDER, int MODE, int RANKS>
__global__ __launch_bounds__(256) void k_p2g_layer(P2GArgs a) {
  if (clock_skip(a.clk)) return;
  const float dt = clock_dt(a.clk, a.dt);
  constexpr int PAY_CELL = pay_cell<RANKS>();
  __shared__ float4 s_pay[4][16 * PAY_CELL / 4];
  __shared__ float s_slab[8][4][SLAB];  // [layer + 4 row half][quantity][node], as the row pass's waves
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int ntile = a.counters[CNT_TILES];
  constexpr int NQ = (MODE == P2G_FORCE) ? 3 : 4;
  const int c = lane >> 2, sr = lane & 3, zo = lane & 3;  // staging (c, sr); accumulation (c, zo)
  const int cx = c & 3, cy = c >> 2;
  const float fzo = (float)zo;
  for (int i = tid; i < 8 * 4 * SLAB; i += 256) (&s_slab[0][0][0])[i] = 0.0f;
  __syncthreads();
  float* slab = &s_slab[w + 4 * (cy >> 1)][0][0];
  const XcdRange tr = xcd_range(ntile);
  for (int slot = tr.first; slot < tr.end; slot += tr.stride) {
    const int gtile = a.active_tiles[slot];
    const int sc = __builtin_amdgcn_readfirstlane(scene_of(a.T.tile_off, a.T.nscenes, gtile));
    const DevParams& P = a.T.params[sc];
    const P2GScene S = p2g_scene(P);
    const int64_t noff = a.T.node_off[sc];
    const int tile = gtile - a.T.tile_off[sc];
    const int nt0 = __builtin_amdgcn_readfirstlane(P.tiles[0]), nt1 = __builtin_amdgcn_readfirstlane(P.tiles[1]);
    const int t0 = tile % nt0, t1 = (tile / nt0) % nt1, t2 = tile / (nt0 * nt1);
    const int g0 = 4 * t0 + cx, g1 = 4 * t1 + cy, g2 = 4 * t2 + w;
    int start = 0, cnt = 0;
    if (g0 < S.res[0] && g1 < S.res[1] && g2 < S.res[2]) {
      const int64_t key = noff + (g2 * S.res[1] + g1) * S.res[0] + g0;
      if (a.occ[key / SCAN_TILE]) {
        start = a.cell_start[key];
        cnt = a.cell_end[key] - start;
      }
    }
    int cmax = cnt;
#pragma unroll
    for (int off = 4; off < 64; off <<= 1) cmax = max(cmax, __shfl_xor(cmax, off));
    cmax = __builtin_amdgcn_readfirstlane(cmax);
    if (cmax > 0) {
      float acc[4][4][NQ];  // [oy][ox][quantity]
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int q = 0; q < NQ; ++q) acc[i][j][q] = 0.0f;
      StageIn in;
      if (MODE != P2G_FORCE && sr < RANKS && sr < cnt) p2g_load<MODE>(a, S, start + sr, in);
      int pi2 = 0;  // slot of this lane's particle in the chunk after the next
      if (MODE != P2G_FORCE && sr < RANKS && RANKS + sr < cnt) pi2 = p2g_slot(a, start + RANKS + sr);
      for (int k = 0; k * RANKS < cmax; ++k) {
        if (sr < RANKS) {
          const int rank = RANKS * k + sr;
          if (rank < cnt) {
            if constexpr (MODE == P2G_FORCE) p2g_load<MODE>(a, S, start + rank, in);
            p2g_stage<ORDER, MODE>(a, S, P, noff, dt, in, g0, g1, g2, &s_pay[w][(c * PAY_CELL + sr * PAY_STRIDE) / 4]);
          }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (MODE != P2G_FORCE && sr < RANKS) {  // the next chunk's data and the slots of the one after
          if (RANKS * (k + 1) + sr < cnt) p2g_load_at<MODE>(a, S, pi2, in);
          if (RANKS * (k + 2) + sr < cnt) pi2 = p2g_slot(a, start + RANKS * (k + 2) + sr);
        }
        const int nr = min(RANKS, cnt - RANKS * k);
        for (int r = 0; r < nr; ++r) {
          const float4* src = &s_pay[w][(c * PAY_CELL + r * PAY_STRIDE) / 4];
          const float* pf = reinterpret_cast<const float*>(src);
          const float4 WX = src[PY_WX / 4], WY = src[PY_WY / 4], DWY = src[PY_DWY / 4];
          const float wxs[4] = {WX.x, WX.y, WX.z, WX.w};
          const float wys[4] = {WY.x, WY.y, WY.z, WY.w};
          const float dwys[4] = {DWY.x, DWY.y, DWY.z, DWY.w};
          const float wz = pf[PY_WZ + zo], dwz = pf[PY_DWZ + zo];
          float g[12];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const float4 t = src[PY_G / 4 + i];
            g[4 * i] = t.x; g[4 * i + 1] = t.y; g[4 * i + 2] = t.z; g[4 * i + 3] = t.w;
          }
          float qs[16];  // m, q0(3), Q12(6), S12(6)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float4 t = src[PY_M / 4 + i];
            qs[4 * i] = t.x; qs[4 * i + 1] = t.y; qs[4 * i + 2] = t.z; qs[4 * i + 3] = t.w;
          }
#pragma unroll
          for (int oy = 0; oy < 4; ++oy) {
            const float wy = wys[oy], dwy = dwys[oy];
            const float wyz = wy * wz, dyz = dwy * wz, ydz = wy * dwz;
            float e[3];
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
              float qrow = qs[1 + rr];
              if constexpr (MODE != P2G_FORCE)
                qrow = __builtin_fmaf(qs[4 + 2 * rr + 1], fzo, __builtin_fmaf(qs[4 + 2 * rr], (float)oy, qrow));
              e[rr] = wyz * qrow;
              if constexpr (MODE != P2G_MOM)
                e[rr] = __builtin_fmaf(-qs[10 + 2 * rr + 1], ydz, __builtin_fmaf(-qs[10 + 2 * rr], dyz, e[rr]));
            }
            const float mw = qs[0] * wyz;
#pragma unroll
            for (int ox = 0; ox < 4; ++ox) {
              const float wx = wxs[ox];
              constexpr int q1 = (MODE != P2G_FORCE) ? 1 : 0;
              if constexpr (MODE != P2G_FORCE) acc[oy][ox][0] = __builtin_fmaf(wx, mw, acc[oy][ox][0]);
#pragma unroll
              for (int rr = 0; rr < 3; ++rr)
                acc[oy][ox][q1 + rr] =
                    __builtin_fmaf(wyz, g[3 * ox + rr], __builtin_fmaf(wx, e[rr], acc[oy][ox][q1 + rr]));
            }
          }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
      // ---- the cells' sums into the slab of their row half: phase (oy descending, ox ascending),
      // node (cx + ox, cy + oy, zo); a node receives its cells in ascending row, then ascending ox
#pragma unroll
      for (int oy = 3; oy >= 0; --oy)
#pragma unroll
        for (int ox = 0; ox < 4; ++ox) {
          const int ln = (cx + ox) + FOOT * (cy + oy) + FOOT * FOOT * zo;
#pragma unroll
          for (int q = 0; q < NQ; ++q) slab[q * SLAB + ln] += acc[oy][ox][q];  // own slab, a node per lane
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
    }
    __syncthreads();
    // ---- tile partial: as the row pass (layers l ascending, row halves in order); slabs cleared
    float* part = a.partials + (size_t)slot * PART_STRIDE + (MODE == P2G_FORCE ? 4 * FOOT3 : 0);
    for (int i = tid; i < FOOT3; i += 256) {
      const int X = i % FOOT, Y = (i / FOOT) % FOOT, Z = i / (FOOT * FOOT);
      float sum[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) sum[q] = 0.0f;
#pragma unroll
      for (int ww = 0; ww < 8; ++ww) {
        const int l = ww >> 1, hh = ww & 1, wv = l + 4 * hh;
        const int z = Z - l;
        if (z < 0 || z > 3) continue;
        const int ln = X + FOOT * Y + FOOT * FOOT * z;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          sum[q] += s_slab[wv][q][ln];
          s_slab[wv][q][ln] = 0.0f;
        }
      }
      if constexpr (NQ == 4) {
        reinterpret_cast<float4*>(part)[i] = make_float4(sum[0], sum[1], sum[2], sum[3]);
      } else {
#pragma unroll
        for (int q = 0; q < NQ; ++q) part[q * FOOT3 + i] = sum[q];
      }
    }
    __syncthreads();
  }
}

// ================================================================== grid update
// One 64-thread workgroup per active 4x4x4 node block; each node sums the partials of the (at most
// 2 per axis) tiles whose footprint covers it, in ascending tile order, then applies
// v = p/m (HybridSolver.cpp:233-240), v += dt (f/m + g) (:457, :725-737) and collision (:460-511).
template <int MODE>
__global__ __launch_bounds__(64) void k_grid(GridArgs a) {
  if (clock_skip(a.clk)) return;
  const float dt = clock_dt(a.clk, a.dt);
  const int nblk = a.counters[CNT_BLOCKS];
  float vmax = 0.0f;  // max |v| over this lane's nodes after the velocity update
  const int lx = threadIdx.x & 3, ly = (threadIdx.x >> 2) & 3, lz = threadIdx.x >> 4;
  const XcdRange br = xcd_range(nblk);
  for (int s = br.first; s < br.end; s += br.stride) {
    const int gblk = a.active_blocks[s];
    const int sc = __builtin_amdgcn_readfirstlane(scene_of(a.T.tile_off, a.T.nscenes, gblk));
    const DevParams& P = a.T.params[sc];
    const int toff = a.T.tile_off[sc];
    const int blk = gblk - toff;
    const int nt0 = __builtin_amdgcn_readfirstlane(P.tiles[0]), nt1 = __builtin_amdgcn_readfirstlane(P.tiles[1]),
              nt2 = __builtin_amdgcn_readfirstlane(P.tiles[2]);
    const int b0 = blk % nt0, b1 = (blk / nt0) % nt1, b2 = blk / (nt0 * nt1);
    // the slots of the 27 tiles around the block (b - 1 .. b + 1 per axis), one lane each; every
    // node of the block is covered by tiles among them
    int slot27 = -1;
    {
      const int q = (int)threadIdx.x;
      const int u0 = b0 - 1 + q % 3, u1 = b1 - 1 + (q / 3) % 3, u2 = b2 - 1 + q / 9;
      if (q < 27 && u0 >= 0 && u1 >= 0 && u2 >= 0 && u0 < nt0 && u1 < nt1 && u2 < nt2)
        slot27 = a.tile_slot[toff + (u2 * nt1 + u1) * nt0 + u0];
    }
    const int r0 = __builtin_amdgcn_readfirstlane(P.res[0]), r1 = __builtin_amdgcn_readfirstlane(P.res[1]),
              r2 = __builtin_amdgcn_readfirstlane(P.res[2]);
    const int n0 = 4 * b0 + lx, n1 = 4 * b1 + ly, n2 = 4 * b2 + lz;
    const bool in_grid = n0 < r0 && n1 < r1 && n2 < r2;
    const int64_t node = a.T.node_off[sc] + ((int64_t)n2 * r1 + n1) * r0 + n0;
    const int nn[3] = {n0, n1, n2}, ntl[3] = {nt0, nt1, nt2};
    // covering tiles of this node along each axis: tlo and tlo + 1 (if <= thi), both in b - 1 .. b + 1
    int tlo[3], thi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      int lo = (nn[d] - 5 + 3) / 4;  // ceil((n-5)/4) for n >= 2; clamped below
      if (nn[d] - 5 < 0) lo = 0;
      int hi = (nn[d] + 1) / 4;
      thi[d] = hi < ntl[d] - 1 ? hi : ntl[d] - 1;
      tlo[d] = lo;
    }
    const int qoff = (MODE == GRID_FORCE_UPDATE) ? 4 : 0;
    constexpr int NQ = (MODE == GRID_FORCE_UPDATE) ? 3 : 4;
    // all partial loads of the node first (at most 2 x 2 x 2 tiles), then the sums in ascending tile
    // order (t2, t1, t0), as a loop over the covering tiles would add them
    float val[8][NQ];
    bool use[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int t0 = tlo[0] + (k & 1), t1 = tlo[1] + ((k >> 1) & 1), t2 = tlo[2] + (k >> 2);
      const int q27 = (t2 - (b2 - 1)) * 9 + (t1 - (b1 - 1)) * 3 + (t0 - (b0 - 1));
      const int sl = __shfl(slot27, q27);  // q27 is in [0, 27): lanes 0..26 hold the slots
      use[k] = in_grid && t0 <= thi[0] && t1 <= thi[1] && t2 <= thi[2] && q27 >= 0 && q27 < 27 && sl >= 0;
      const int ln = (n0 - (4 * t0 - 1)) + FOOT * (n1 - (4 * t1 - 1)) + FOOT * FOOT * (n2 - (4 * t2 - 1));
      const float* part = a.partials + (size_t)(use[k] ? sl : 0) * PART_STRIDE + qoff * FOOT3;
      const int pn = use[k] ? ln : 0;
      // unconditional loads (an unused tile reads slot 0, node 0: in bounds) keep them all in flight
      if constexpr (NQ == 4) {  // m, p of a node: one float4
        const float4 p4 = reinterpret_cast<const float4*>(part)[pn];
        val[k][0] = p4.x;
        val[k][1] = p4.y;
        val[k][2] = p4.z;
        val[k][3] = p4.w;
      } else {
#pragma unroll
        for (int q = 0; q < NQ; ++q) val[k][q] = part[q * FOOT3 + pn];
      }
    }
    float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int q = 0; q < NQ; ++q) sum[q] = use[k] ? sum[q] + val[k][q] : sum[q];
    if (!in_grid) continue;
    float v[3], vbf[3];
    if constexpr (MODE == GRID_FORCE_UPDATE) {
      const float m = a.gm[node];
      const float4 g4 = a.gvel[2 * node];
      v[0] = g4.x; v[1] = g4.y; v[2] = g4.z;
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (m > 0.0f) v[d] += dt * (sum[d] / m + P.gravity[d]);
    } else {
      const float m = sum[0];
      a.gm[node] = m;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        a.gmom[d * a.ng + node] = sum[1 + d];
        v[d] = m > 0.0f ? sum[1 + d] / m : 0.0f;
        if (MODE == GRID_FUSED && m > 0.0f) v[d] += dt * P.gravity[d];
      }
    }
    vmax = fmaxf(vmax, sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    if constexpr (MODE == GRID_NORMALIZE) {
      vbf[0] = v[0]; vbf[1] = v[1]; vbf[2] = v[2];
    } else {
      const float xn[3] = {P.gmin[0] + (float)n0 * P.h[0], P.gmin[1] + (float)n1 * P.h[1], P.gmin[2] + (float)n2 * P.h[2]};
      collide(P, xn, v, vbf);
    }
    a.gvel[2 * node] = make_float4(v[0], v[1], v[2], vbf[0]);
    a.gvel[2 * node + 1] = make_float4(vbf[1], vbf[2], 0.0f, 0.0f);
  }
  if (a.vmax_out != nullptr) {
    // max over the wave, one atomic per wave (non-negative floats order as their bit patterns)
    for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
    if (threadIdx.x == 0 && vmax > 0.0f) atomicMax(a.vmax_out, __float_as_uint(vmax));
  }
}

// ================================================================== adaptive time step
// Dt = CFL / max(300, max|v| / h_min) (HybridSolver.cpp:860,878; RegularGrid.cpp:188-200,
// RegularGrid.h:60), evaluated in double as in the reference from the max node speed after the
// velocity update; then the frame clamp to 1/60 s (:880-892): a substep that reaches the frame end
// is shortened to it, t advances by 1/60 and the frame time restarts.
__global__ void k_clock_update(Clock* c, int init) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || c->skip) return;
  const double vmax = (double)__uint_as_float(c->vmax_bits);
  c->vmax_bits = 0u;
  const double rate = vmax / c->hmin;
  double Dt = c->cfl / (rate > 300.0 ? rate : 300.0);
  if (!init) {
    if (c->inner + Dt >= 1.0 / 60.0) {
      Dt = 1.0 / 60.0 - c->inner;
      c->t += 1.0 / 60.0;
      c->inner = 0.0;
      c->frame_now = 1;
    } else {
      c->inner += Dt;
    }
  }
  c->dt = Dt;
  c->dt_f = (float)Dt;
}

// End of a substep: the reference closes a frame after the substep that completed it (output at
// HybridSolver.cpp:990-1029) and leaves its loop when t > maxt (:867).
__global__ void k_clock_end(Clock* c) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || c->skip) return;
  c->executed += 1;
  c->substeps += 1;
  if (c->frame_now) {
    c->frame_now = 0;
    c->frame += 1;
    c->frame_done = 1;
    c->skip = 1;
  }
  if (c->t > c->maxt) {
    c->done = 1;
    c->skip = 1;
  }
}

// ================================================================== G2P + F update + return map
constexpr int G2P_BX = 12;               // x extent of a wave's node box (nodes)
constexpr int G2P_BOX = G2P_BX * 4 * 4;  // nodes of the box: G2P_BX x 4 (y) x 4 (z)

// A node box's nodes, G2P_BOX / 64 per lane (slot i = it * 64 + lane: x = i % G2P_BX, then y, then
// z), in the grid's record layout; slots outside the box read the box's corner (in bounds).
__device__ __forceinline__ void g2p_box_load(const float4* gvel, int lane, int x0, int y0, int z0, int nx, int ny,
                                             int nz, int64_t res0, int64_t res01, float4 (&t)[2 * G2P_BOX / 64]) {
#pragma unroll
  for (int it = 0; it < G2P_BOX / 64; ++it) {
    const int i = it * 64 + lane;
    const int ix = i % G2P_BX, row = i / G2P_BX, iy = row & 3, iz = row >> 2;
    const bool in = ix < nx && iy < ny && iz < nz;
    const int64_t node = (z0 + (in ? iz : 0)) * res01 + (y0 + (in ? iy : 0)) * res0 + (x0 + (in ? ix : 0));
    t[2 * it] = gvel[2 * node];
    t[2 * it + 1] = gvel[2 * node + 1];
  }
}
__device__ __forceinline__ void g2p_box_store(float4* box, int lane, const float4 (&t)[2 * G2P_BOX / 64]) {
#pragma unroll
  for (int it = 0; it < G2P_BOX / 64; ++it) {
    const int i = it * 64 + lane;
    box[2 * i] = t[2 * it];
    box[2 * i + 1] = t[2 * it + 1];
  }
}

// floor(a / d) for 0 <= a < 2^31 with the multiplier m and shift l of d (integer arithmetic only;
// l = 0 means d = 1): t = floor(a m / 2^32), then (t + ((a - t) >> 1)) >> (l - 1).
__device__ __forceinline__ int floor_div(int a, uint32_t m, int l) {
  if (l == 0) return a;
  const uint32_t n = (uint32_t)a, t = __umulhi(n, m);
  return (int)((t + ((n - t) >> 1)) >> (l - 1));
}

// Thread per sorted position j (caller id pid = perm[j]): gathers the stencil of grid v
// (post-collision) and v_BF (pre-friction), computes v_p, APIC B (HybridSolver.cpp:739-825),
// x += dt sum w v_BF (:942-950), F^ = (I + dt sum v_BF grad w^T) F (:553-577) and the return map
// (:612-681), and writes the particle to slot j of the other buffer, so that the buffers stay
// physically sorted by cell. It also records where pid now lives and pid's next cell key.
template <int ORDER>
__global__ __launch_bounds__(256) void k_g2p(G2PArgs a) {
  if (clock_skip(a.clk)) return;
  const float dt = clock_dt(a.clk, a.dt);
  const int pblk = (int)blockIdx.x;
  const PBlock pb = particle_block(a.T, pblk);
  const DevParams P = uniform_params(a.T.params[pb.s]);
  const int64_t noff = a.T.node_off[pb.s];
  const float4* gvel = a.gvel + 2 * noff;
  const int j = pb.j;
  int oob = 0, bad = 0, cfl = 0, fb_svd = 0, fb_apex = 0;
  bool cand = false, apex = false;
  const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int64_t res0 = P.res[0], res01 = (int64_t)P.res[0] * P.res[1];
  // The wave's node boxes. When the wave's positions lie in one row of cells, or in two consecutive
  // rows of one z plane (the sorted cell keys of its positions tell), the nodes c-1 .. c+2 around
  // each row's cells (at most G2P_BX x 4 x 4 per row) are copied into the wave's LDS area, in the
  // grid's own record layout, and the stencil gathers read them there: one round trip to HBM, made
  // together with the particle loads, instead of one per z plane. The keys only place the boxes: a
  // wave uses them only if every lane's stencil lies inside its row's box, and they are clamped to
  // the grid whatever the keys hold.
  __shared__ float4 s_box[4][2][2 * G2P_BOX];
  const uint32_t dm0 = a.T.params[pb.s].div_m[0], dm1 = a.T.params[pb.s].div_m[1];
  const int ds0 = a.T.params[pb.s].div_s[0], ds1 = a.T.params[pb.s].div_s[1];
  // The particle's state and its sorted key, loaded by every lane (lanes past the end read the
  // last position), so that the box below is placed while they are in flight.
  const int jc = j < pb.end ? j : pb.end - 1;
  const uint32_t kj = a.ks[jc];
  const int perm_j = a.perm[jc], src_j = a.src[jc];
  asm volatile("" ::: "memory");  // the three coalesced loads are issued together, first
  const int pid = clamp_index(perm_j, a.n);
  const int pi = clamp_index(src_j, a.n);
  const float* st = a.sin;
  const int64_t cap = a.cap;
  float xp[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) xp[d] = st[(PX + d) * cap + pi];
  float Fn[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fn[i] = st[(PF + i) * cap + pi];
  float plastic = st[PPL * cap + pi];
  const float phi_p = st[PPHI * cap + pi];
  const float vol = st[PVOL * cap + pi];
  const float mass = st[PM * cap + pi];
  // Box q covers x bx0q .. +bnxq, y by0q .. +bnyq, z bz0 .. +bnz (uniform). One box when the wave's
  // cells are in one row; two when they are in two consecutive rows of one z plane (lanes below
  // `split` in the first). nbox = 0: the wave gathers from the grid.
  int bx00 = 0, by00 = 0, bnx0 = 0, bny0 = 0, bx01 = 0, by01 = 0, bnx1 = 0, bny1 = 0;
  int bz0 = 0, bnz = 0, nbox = 0, split = 64;
  const int jw = __builtin_amdgcn_readfirstlane(j - lane);
  if (jw < pb.end) {
    const int64_t ncell = (int64_t)P.res[0] * P.res[1] * P.res[2];
    const int64_t lf = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int)kj, 0) - noff;
    const int64_t ll = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int)kj, 63) - noff;
    if (lf >= 0 && ll >= lf && ll < ncell && ll <= INT32_MAX) {
      const int rf = floor_div((int)lf, dm0, ds0);
      const int rl = floor_div((int)ll, dm0, ds0);
      const int zc = floor_div(rf, dm1, ds1), yc = rf - zc * P.res[1];
      const int xf = (int)lf - rf * P.res[0], xl = (int)ll - rl * P.res[0];
      int xa = xl, xb = xf;  // last x of the first row, first x of the second
      if (rf == rl) {
        nbox = 1;
      } else if (rl == rf + 1 && yc + 1 < P.res[1]) {
        // keys are sorted: the lanes whose cell precedes row rl's first cell are the first row's
        const int64_t r0 = P.res[0], row_end = (int64_t)rl * r0;
        split = __popcll(__ballot((int64_t)(uint32_t)kj - noff < row_end));
        if (split >= 1 && split <= 63) {
          // (clamped to the row: only keys that are not sorted could fall outside it)
          const int64_t ka = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int)kj, split - 1) - noff - (int64_t)rf * r0;
          const int64_t kb = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int)kj, split) - noff - row_end;
          xa = (int)(ka < 0 ? 0 : (ka >= r0 ? r0 - 1 : ka));
          xb = (int)(kb < 0 ? 0 : (kb >= r0 ? r0 - 1 : kb));
          nbox = 2;
        }
      }
      if (nbox == 1) split = 64;
      bz0 = max(zc - 1, 0);
      bnz = min(zc + 2, P.res[2] - 1) - bz0 + 1;
      bx00 = max(xf - 1, 0);
      bnx0 = min((nbox == 2 ? xa : xl) + 2, P.res[0] - 1) - bx00 + 1;
      by00 = max(yc - 1, 0);
      bny0 = min(yc + 2, P.res[1] - 1) - by00 + 1;
      bx01 = max(xb - 1, 0);
      bnx1 = min(xl + 2, P.res[0] - 1) - bx01 + 1;
      by01 = max(yc, 0);
      bny1 = min(yc + 3, P.res[1] - 1) - by01 + 1;
      bool ok = bnz >= 1 && bnx0 >= 1 && bnx0 <= G2P_BX && bny0 >= 1;
      if (nbox == 2) ok = ok && bnx1 >= 1 && bnx1 <= G2P_BX && bny1 >= 1;
      if (!ok) nbox = 0;
    }
  }
  if (nbox > 0) {
    // the second row's box after the first (one more round trip for those waves, but no more live
    // registers than one box needs)
    {
      float4 t[2 * G2P_BOX / 64];
      g2p_box_load(gvel, lane, bx00, by00, bz0, bnx0, bny0, bnz, res0, res01, t);
      g2p_box_store(&s_box[wv][0][0], lane, t);
    }
    if (nbox == 2) {
      float4 t[2 * G2P_BOX / 64];
      g2p_box_load(gvel, lane, bx01, by01, bz0, bnx1, bny1, bnz, res0, res01, t);
      g2p_box_store(&s_box[wv][1][0], lane, t);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  if (j < pb.end) {
    float g[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) g[d] = (xp[d] - P.gmin[d]) * P.inv_h[d];
    int base[3];
    float w[3][4], dw[3][4];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      base[d] = stencil_base<ORDER>(g[d], P.res[d]);
      if constexpr (ORDER == 2) {
        quad_weights(g[d] - (float)base[d], w[d], dw[d]);
      } else {
#pragma unroll
        for (int o = 0; o < ORDER + 1; ++o) spline<ORDER>(g[d] - (float)(base[d] + o), w[d][o], dw[d][o]);
      }
    }
    // Out-of-grid nodes are dropped (HybridSolver.cpp:44-47): they get zero weight and an in-range
    // (clamped) index, so every load stays inside the stencil's active node blocks.
    int nid[3][ORDER + 1];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int o = 0; o < ORDER + 1; ++o) {
        const int nn = base[d] + o;
        const bool in = nn >= 0 && nn < P.res[d];
        w[d][o] = in ? w[d][o] : 0.0f;
        dw[d][o] = in ? dw[d][o] : 0.0f;
        nid[d][o] = nn < 0 ? 0 : (nn >= P.res[d] ? P.res[d] - 1 : nn);
      }
    // The stencil sums factor over the axes: along each x row of nodes (fixed y, z) the row sums
    //   A = sum wx v, Ax = sum wx (x_i - x_p) v, C = sum wx v_BF, Cd = sum wx' v_BF
    // are formed first, then weighted by the row's y and z factors:
    //   v_p += wyz A, B[:,0] += wyz Ax, B[:,1] += wyz dy A, B[:,2] += wyz dz A,
    //   sum w v_BF += wyz C, grad v_BF[:,0] += wyz/h Cd, [:,1] += wy' wz/h C, [:,2] += wy wz'/h C.
    float wdx[ORDER + 1];
#pragma unroll
    for (int ox = 0; ox < ORDER + 1; ++ox) wdx[ox] = w[0][ox] * (((float)(base[0] + ox) - g[0]) * P.h[0]);
    float vp[3] = {0.0f, 0.0f, 0.0f}, vb[3] = {0.0f, 0.0f, 0.0f};
    float Bn[9], Gv[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) { Bn[i] = 0.0f; Gv[i] = 0.0f; }
    // node addressing: the wave's box when every lane's stencil lies inside it, else the grid
    const int mb = lane >= split ? 1 : 0;  // this lane's box
    const int lx0 = mb ? bx01 : bx00, ly0 = mb ? by01 : by00, lnx = mb ? bnx1 : bnx0, lny = mb ? bny1 : bny0;
    const bool inside = nid[0][0] >= lx0 && nid[0][ORDER] < lx0 + lnx && nid[1][0] >= ly0 &&
                        nid[1][ORDER] < ly0 + lny && nid[2][0] >= bz0 && nid[2][ORDER] < bz0 + bnz;
    const bool use_box = nbox > 0 && __ballot(!inside) == 0;
    // LDS offsets (in float4, from the wave's first box) of the stencil's x and y nodes in its box
    int bxo[ORDER + 1], byo[ORDER + 1];
#pragma unroll
    for (int o = 0; o < ORDER + 1; ++o) {
      bxo[o] = 2 * (nid[0][o] - lx0);
      byo[o] = 2 * G2P_BOX * mb + 2 * G2P_BX * (nid[1][o] - ly0);
    }
    // one z plane of rows per iteration (not unrolled): its loads are in flight together without
    // holding the whole stencil in registers
#pragma unroll 1
    for (int oz = 0; oz < ORDER + 1; ++oz) {
      float wz = w[2][0], dwz = dw[2][0];
      int nz = nid[2][0];
#pragma unroll
      for (int o = 1; o < ORDER + 1; ++o) {
        wz = oz == o ? w[2][o] : wz;
        dwz = oz == o ? dw[2][o] : dwz;
        nz = oz == o ? nid[2][o] : nz;
      }
      const float dz = ((float)(base[2] + oz) - g[2]) * P.h[2];
      // the plane's (ORDER+1)^2 nodes are loaded together and waited for once (v and v_BF of a
      // node are its two float4, of which the second holds v_BF y, z in x, y), from the wave's box
      // in LDS or from the grid (the branch is uniform)
      float4 qa[ORDER + 1][ORDER + 1];
      float2 qb[ORDER + 1][ORDER + 1];
      if (use_box) {
        const float4* pp = &s_box[wv][0][0] + 2 * (4 * G2P_BX) * (nz - bz0);
#pragma unroll
        for (int oy = 0; oy < ORDER + 1; ++oy)
#pragma unroll
          for (int ox = 0; ox < ORDER + 1; ++ox) {
            qa[oy][ox] = pp[byo[oy] + bxo[ox]];
            qb[oy][ox] = *reinterpret_cast<const float2*>(pp + byo[oy] + bxo[ox] + 1);
          }
      } else {
        const float4* pp = gvel + 2 * (nz * res01);
#pragma unroll
        for (int oy = 0; oy < ORDER + 1; ++oy) {
          const float4* rp = pp + 2 * (nid[1][oy] * res0);
#pragma unroll
          for (int ox = 0; ox < ORDER + 1; ++ox) {
            qa[oy][ox] = rp[2 * nid[0][ox]];
            qb[oy][ox] = *reinterpret_cast<const float2*>(rp + 2 * nid[0][ox] + 1);
          }
        }
      }
#pragma unroll
      for (int oy = 0; oy < ORDER + 1; ++oy) {
        const float dy = ((float)(base[1] + oy) - g[1]) * P.h[1];
        float A[3], Ax[3], C[3], Cd[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) { A[r] = 0.0f; Ax[r] = 0.0f; C[r] = 0.0f; Cd[r] = 0.0f; }
#pragma unroll
        for (int ox = 0; ox < ORDER + 1; ++ox) {
          const float4 q0 = qa[oy][ox];
          const float2 q1 = qb[oy][ox];
          const float vi[3] = {q0.x, q0.y, q0.z};
          const float bi[3] = {q0.w, q1.x, q1.y};
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            A[r] += w[0][ox] * vi[r];
            Ax[r] += wdx[ox] * vi[r];
            C[r] += w[0][ox] * bi[r];
            Cd[r] += dw[0][ox] * bi[r];
          }
        }
        const float wyz = w[1][oy] * wz;
        const float wyzy = wyz * dy, wyzz = wyz * dz;
        const float gx = wyz * P.inv_h[0];
        const float gy = dw[1][oy] * wz * P.inv_h[1];
        const float gz = w[1][oy] * dwz * P.inv_h[2];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          vp[r] += wyz * A[r];
          Bn[3 * r] += wyz * Ax[r];
          Bn[3 * r + 1] += wyzy * A[r];
          Bn[3 * r + 2] += wyzz * A[r];
          vb[r] += wyz * C[r];
          Gv[3 * r] += gx * Cd[r];
          Gv[3 * r + 1] += gy * C[r];
          Gv[3 * r + 2] += gz * C[r];
        }
      }
    }
    // cloth vertices and elements keep only the skew part of B (damping 1, HybridSolver.cpp:926-932)
    const int kind = particle_kind(P, pid);
    const float damping = kind == KIND_PARTICLE ? P.damping : 1.0f;
    if (damping != 0.0f) {
      float C[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) C[i] = Bn[i];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          float sym = 0.5f * (C[3 * r + c] + C[3 * c + r]);
          Bn[3 * r + c] = (C[3 * r + c] - sym) + (1.0f - damping) * sym;
        }
    }
    float Fh[9], GF[9];
    mat_mul(Gv, Fn, GF);
#pragma unroll
    for (int i = 0; i < 9; ++i) Fh[i] = Fn[i] + dt * GF[i];
    float K[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    SurrogateIO io;
    io.mode = a.sur_mode;
    io.pair_log = a.pair_log;
    io.pair_cap = a.pair_cap;
    io.pair_count = a.pair_count;
    io.deferred = false;
    io.defer_apex = a.apex_mask != nullptr;
    io.apex_deferred = false;
    // particles: return mapping. A cloth element's F planes hold D, of which F^ updates d3
    // (HybridSolver.cpp:599-603); its d1, d2 and return mapping follow in k_cloth_elements.
    if (kind == KIND_PARTICLE) return_map(P, Fh, plastic, phi_p, vol, K, fb_svd, fb_apex, io);
    // A deferred particle's trial SVD goes to its record: the wave's deferred lanes write next to
    // each other from the position of lane 0 (coalesced even where few lanes defer); the wave's mask
    // below lists the particle, and the list carries the record index.
    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t dmask = __ballot(io.deferred), amask = __ballot(io.apex_deferred);
    if (io.deferred) {  // a candidate of the surrogate
      cand = true;
      const int64_t n = a.n, r = (j - lane) + __popcll(dmask & below);
#pragma unroll
      for (int k = 0; k < 3; ++k) a.cand[(CAND_SIGMA + k) * n + r] = io.sigma[k];
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        a.cand[(CAND_U + k) * n + r] = io.U[k];
        a.cand[(CAND_V + k) * n + r] = io.V[k];
      }
    }
    if (io.apex_deferred) {  // an FP64 apex projection for k_apex_finish
      apex = true;
      const int64_t n = a.n, r = (j - lane) + __popcll(amask & below);
#pragma unroll
      for (int k = 0; k < 3; ++k) a.apex_usv[(CAND_SIGMA + k) * n + r] = io.sigma[k];
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        a.apex_usv[(CAND_U + k) * n + r] = io.U[k];
        a.apex_usv[(CAND_V + k) * n + r] = io.V[k];
      }
    }
    float xn[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xn[d] = xp[d] + dt * vb[d];
      if (fabsf(xn[d] - xp[d]) > P.h[d]) cfl = 1;
    }
    float* so = a.sout;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      so[(PX + d) * cap + j] = xn[d];
      so[(PV + d) * cap + j] = vp[d];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      so[(PB + i) * cap + j] = Bn[i];
      so[(PF + i) * cap + j] = Fh[i];
    }
    if (a.write_stress) {
#pragma unroll
      for (int e = 0; e < 6; ++e) so[(PK + e) * cap + j] = K[e];
    }
    so[PPL * cap + j] = plastic;
    so[PM * cap + j] = mass;
    so[PVOL * cap + j] = vol;
    so[PPHI * cap + j] = phi_p;
    reinterpret_cast<int*>(so + (size_t)PID * cap)[j] = pid;
    if (a.pos_out) a.pos_out[pid] = j;
    a.keys[j] = (uint32_t)noff + cell_key(P, a.thr + a.T.thr_off[pb.s], xn[0], xn[1], xn[2], oob);
    bad = !(isfinite(xn[0]) && isfinite(xn[1]) && isfinite(xn[2]) && isfinite(vp[0]) && isfinite(vp[1]) &&
            isfinite(vp[2]) && isfinite(Fh[0]) && isfinite(Fh[4]) && isfinite(Fh[8]) && isfinite(plastic));
  }
  // every wave records its deferred lanes (none past the end of the scene) for the position lists
  if (a.apex_mask || a.sur_mode == 2) {
    const uint64_t am = __ballot(apex), dm = __ballot(cand);
    if (lane == 0) {
      a.wave_pos[4 * pblk + wv] = j;
      if (a.apex_mask) a.apex_mask[4 * pblk + wv] = am;
      if (a.sur_mode == 2) a.cand_mask[4 * pblk + wv] = dm;
    }
  }
  const int flags = (oob ? MPM_STATUS_OUT_OF_DOMAIN : 0) | (bad ? MPM_STATUS_NONFINITE : 0) | (cfl ? MPM_STATUS_CFL : 0);
  if (__ballot(flags != 0)) {
    if (flags) atomicOr(a.status, flags);
  }
  // fallback counts: one atomic per wave
  if (__ballot(fb_svd != 0)) {
    unsigned long long c = (unsigned long long)fb_svd;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane_id() == 0) atomicAdd(a.fallback_svd, c);
  }
  if (__ballot(fb_apex != 0)) {
    unsigned long long c = (unsigned long long)fb_apex;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane_id() == 0) atomicAdd(a.fallback_apex, c);
  }
}

// The FP64 apex projections deferred by k_g2p, gathered into full waves: from the trial SVD that
// k_g2p stored and the particle's q, phi and V, the FP64 projection and F_E and V tau (the same
// function return_map calls, on the same operands), stored over the trial F^ left in the output
// buffer. The result equals computing it inside k_g2p.
__global__ __launch_bounds__(256) void k_apex_finish(ApexArgs a) {
  if (clock_skip(a.clk)) return;
  const int count = min(*a.apex_count, a.n);
  int fb_apex = 0;
  for (int c0 = blockIdx.x * 256; c0 < count; c0 += gridDim.x * 256) {
    const int c = c0 + (int)threadIdx.x;
    if (c < count) {
      const int j = clamp_index(a.apex_slot[c], a.n);
      const int r = clamp_index(a.apex_idx[c], a.n);
      const int sc = scene_of(a.T.part_off, a.T.nscenes, j);
      const DevParams& P = a.T.params[sc];
      float* st = a.st;
      const int64_t cap = a.cap, n = a.n;
      float S[3], U[9], V[9];
#pragma unroll
      for (int k = 0; k < 3; ++k) S[k] = a.apex_usv[(CAND_SIGMA + k) * n + r];
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        U[k] = a.apex_usv[(CAND_U + k) * n + r];
        V[k] = a.apex_usv[(CAND_V + k) * n + r];
      }
      float plastic = st[PPL * cap + j];
      const float phi_p = st[PPHI * cap + j];
      const float vol = st[PVOL * cap + j];
      SurrogateIO io;
      io.mode = a.log_pairs ? 1 : 0;
      io.pair_log = a.pair_log;
      io.pair_cap = a.pair_cap;
      io.pair_count = a.pair_count;
      // the FP32 trial strains, as the return map computed them (used by the pair log only)
      float l[3] = {0.0f, 0.0f, 0.0f};
      if (io.mode == 1) {
        l[0] = logf(S[0]);
        l[1] = logf(S[1]);
        l[2] = logf(S[2]);
      }
      float Fe[9], K[6];
      sand_apex_finish(P, U, S, V, l, P.per_particle_phi ? phi_p : P.h0, phi_p, plastic, vol, Fe, K, fb_apex, io);
#pragma unroll
      for (int i = 0; i < 9; ++i) st[(PF + i) * cap + j] = Fe[i];
      st[PPL * cap + j] = plastic;
      if (a.write_stress) {
#pragma unroll
        for (int e = 0; e < 6; ++e) st[(PK + e) * cap + j] = K[e];
      }
    }
  }
  // one atomic per wave
  unsigned long long c = (unsigned long long)fb_apex;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(a.fallback_apex, c);
}

// ================================================================== position lists
// Lists of the sorted positions that G2P flagged, in ascending order and without atomics: the
// deferred FP64 apex projections and the surrogate's candidates (two lists, one set of launches).
// Every G2P wave w stores the ballots of its flagged lanes (mask[l][w]) and the position of its
// lane 0 (wpos[w]). Waves form groups of 64: k_list_sums totals each group; k_list_write (one
// workgroup per group) adds up the totals of the groups before its own, places each wave's
// positions after that offset and the waves before it in the group, and the last group's
// workgroup writes the list lengths.
__global__ __launch_bounds__(256) void k_list_sums(ListArgs a) {
  const int ng = (a.nwaves + 63) / 64;
  const int g = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (g >= ng) return;
  const int w = g * 64 + lane;
  for (int l = 0; l < a.nlists; ++l) {
    int c = w < a.nwaves ? __popcll(a.mask[l][w]) : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) a.gsum[l * ng + g] = c;
  }
}

__global__ __launch_bounds__(256) void k_list_write(ListArgs a) {
  __shared__ uint64_t s_m[2][64];
  __shared__ int s_off[2][64], s_p[64], s_red[2][4];
  const int ng = (a.nwaves + 63) / 64;
  const int g = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int w = g * 64 + lane;
  // each list's offset of this group: the totals of the groups before it, summed by the workgroup
  // itself (integer sums, so the order does not matter)
  for (int l = 0; l < a.nlists; ++l) {
    int sum = 0;
    for (int i = t; i < g; i += 256) sum += a.gsum[l * ng + i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (lane == 0) s_red[l][wv] = sum;
  }
  __syncthreads();
  if (wv < a.nlists) {  // wave l: the offsets of list l's 64 masks of the group
    const int pre = s_red[wv][0] + s_red[wv][1] + s_red[wv][2] + s_red[wv][3];
    const uint64_t m = w < a.nwaves ? a.mask[wv][w] : 0ull;
    const int c = __popcll(m);
    int x = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    s_m[wv][lane] = m;
    s_off[wv][lane] = pre + x - c;
    // the last group's workgroup writes the list's length
    if (g == ng - 1 && lane == 63) *a.count[wv] = pre + x;
  } else if (wv == 3) {
    s_p[lane] = w < a.nwaves ? a.wpos[w] : 0;
  }
  __syncthreads();
  // wave wv writes the positions of masks 16 wv .. 16 wv + 15 of each list, one mask at a time
  for (int l = 0; l < a.nlists; ++l)
    for (int i = 16 * wv; i < 16 * wv + 16; ++i) {
      const uint64_t mi = s_m[l][i];
      if ((mi >> lane) & 1ull) {
        const int rank = __popcll(mi & ((1ull << lane) - 1ull));
        a.list[l][s_off[l][i] + rank] = s_p[i] + lane;
        a.idx[l][s_off[l][i] + rank] = s_p[i] + rank;
      }
    }
}

// ================================================================== surrogate return map
// Cone-branch particles deferred by k_g2p (sur_mode 2) are finished by k_surrogate. Each wave takes
// 64 listed candidates, one per lane, evaluates the MLP for them with BF16 MFMA (weights staged once
// per workgroup in LDS, activations in registers) and then gates each prediction H = eps + 1e-3 y
// of the projected Hencky strains:
//   yield equality   | |dev H| + c alpha tr H | <= tol       (c = (3 lambda + 2 mu) / (2 mu))
//   volume           | tr H - tr eps | <= tol                (the cone return is deviatoric)
//   flow direction   the deviatoric correction eps - H is parallel to dev eps (orthogonal part <= tol)
//                    and points along it
// An accepted prediction gives Sigma' = exp(H), dq = |eps - H|; a rejected one runs the exact
// projection (HybridSolver.cpp:646-673) through the same code as k_g2p, so rejecting every
// prediction reproduces the exact path bit for bit. With a decision log, candidate row r < dec_cap
// records (slot bits, eps, phi, K/mu, H, accepted) at dec_log[10 r ..].
// The gate and the update of a wave's candidates follow its network evaluation in the same kernel:
// their loads and stores (the trial SVD, the new F, q and stress) overlap the other waves' network
// arithmetic, which is bound by the SiLU's transcendentals, instead of taking a pass of their own.

// Gate and state update of candidate `row` at sorted slot j (S, U, V: its trial SVD, eps = log S,
// y: the network's outputs, vol: its volume). Returns whether the prediction was accepted.
__device__ __forceinline__ bool surrogate_update(const SurrogateArgs& a, int row, int j, const DevParams& P,
                                                 const float S[3], const float eps[3], const float y[3], float q,
                                                 float phi_p, float phi, float kmu, const float U[9], const float V[9],
                                                 float vol) {
#pragma clang fp contract(off)
  const int64_t cap = a.cap;
  const float c_yield = (3.0f * P.lambda0 + 2.0f * P.mu0) / (2.0f * P.mu0);
  float H[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) H[k] = eps[k] + MLP_CORR_SCALE * y[k];
  const float fa = phi * 3.14159265358979323846f / 180.0f;
  const float sf = sinf(fa);
  const float alpha = sqrtf(2.0f / 3.0f) * 2.0f * sf / (3.0f - sf);
  const float trH = H[0] + H[1] + H[2], trE = eps[0] + eps[1] + eps[2];
  float dH[3], dE[3], d[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dH[k] = H[k] - trH / 3.0f;
    dE[k] = eps[k] - trE / 3.0f;
  }
  const float nH = sqrtf(dH[0] * dH[0] + dH[1] * dH[1] + dH[2] * dH[2]);
  const float nE = sqrtf(dE[0] * dE[0] + dE[1] * dE[1] + dE[2] * dE[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) d[k] = dE[k] - dH[k];  // deviatoric part of eps - H
  const float along = (d[0] * dE[0] + d[1] * dE[1] + d[2] * dE[2]) / nE;
  float ortho2 = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float o = d[k] - along * dE[k] / nE;
    ortho2 += o * o;
  }
  const float tol = a.gate_tol;
  const bool accepted = fabsf(nH + c_yield * alpha * trH) <= tol && fabsf(trH - trE) <= tol && along > 0.0f &&
                        sqrtf(ortho2) <= tol;
  if (a.dec_log && row < a.dec_cap) {
    float* rec = a.dec_log + (int64_t)10 * row;
    rec[0] = __int_as_float(j);
    rec[1] = eps[0];
    rec[2] = eps[1];
    rec[3] = eps[2];
    rec[4] = phi;
    rec[5] = kmu;
    rec[6] = H[0];
    rec[7] = H[1];
    rec[8] = H[2];
    rec[9] = accepted ? 1.0f : 0.0f;
  }
  float Sn[3], lnSn[3], dq;
  if (accepted) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Sn[k] = expf(H[k]);
      lnSn[k] = H[k];
    }
    dq = sqrtf((eps[0] - H[0]) * (eps[0] - H[0]) + (eps[1] - H[1]) * (eps[1] - H[1]) + (eps[2] - H[2]) * (eps[2] - H[2]));
  } else {
    const float phi0 = P.per_particle_phi ? phi_p : P.h0;
    float lnS[3];
    sand_project(P, phi0, q, S, Sn, lnS, lnSn, dq);
  }
  float Fe[9], K[6];
  sand_finish(P, U, V, Sn, lnSn, dq, q, vol, Fe, K);
#pragma unroll
  for (int k = 0; k < 9; ++k) a.st[(PF + k) * cap + j] = Fe[k];
  a.st[PPL * cap + j] = q;
  if (a.write_stress) {
#pragma unroll
    for (int e = 0; e < 6; ++e) a.st[(PK + e) * cap + j] = K[e];
  }
  return accepted;
}

__global__ __launch_bounds__(MLP_FWD_THREADS) void k_surrogate(SurrogateArgs a) {
  if (clock_skip(a.clk)) return;
  extern __shared__ __align__(16) unsigned char smem[];
  MLPShared& sh = *reinterpret_cast<MLPShared*>(smem);
  // the counter is cleared by the sort stage; a G2P stage repeated without it cannot list more than n
  const int count = min(*a.cand_count, a.n);
  if (blockIdx.x == 0 && threadIdx.x == 0 && count > 0) atomicAdd(a.candidates, (unsigned long long)count);
  const int nw = blockDim.x >> 6;
  // uniform per workgroup: nothing to do, skip the weight staging
  if ((int64_t)blockIdx.x * nw * 64 >= count) return;
  stage_weights(sh, a.wbf, a.master);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t cap = a.cap, n = a.n;
  unsigned n_acc = 0;
  for (int g = blockIdx.x * nw + w; (int64_t)g * 64 < count; g += gridDim.x * nw) {
    const int row = g * 64 + lane;
    float S[3] = {1.0f, 1.0f, 1.0f}, eps[3] = {0.0f, 0.0f, 0.0f}, phi = 0.0f, kmu = 0.0f, q = 0.0f, phi_p = 0.0f;
    int j = 0, sc = 0;
    float U[9], V[9], vol = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) U[k] = V[k] = 0.0f;
    if (row < count) {
      j = clamp_index(a.cand_slot[row], (int)n);
      const int r = clamp_index(a.cand_idx[row], (int)n);
      sc = scene_of(a.T.part_off, a.T.nscenes, j);
      const DevParams& P = a.T.params[sc];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        S[k] = a.cand[(CAND_SIGMA + k) * n + r];
        eps[k] = logf(S[k]);
      }
      q = a.st[PPL * cap + j];
      phi_p = a.st[PPHI * cap + j];
      phi = dp_phi(P, phi_p, q);
      kmu = bulk_over_shear(P);
      // what the update needs after the network: these loads complete under its arithmetic
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        U[k] = a.cand[(CAND_U + k) * n + r];
        V[k] = a.cand[(CAND_V + k) * n + r];
      }
      vol = a.st[PVOL * cap + j];
    }
    const float x[5] = {eps[0], eps[1], eps[2], phi, kmu};
    float y[3];
    wave_forward64<MLP_FWD_TILES>(sh, x, y, nullptr, 0, 0);
    // the network's registers are free from here on: keep the update's loads below this point
    __builtin_amdgcn_sched_barrier(0);
    bool accepted = false;
    if (row < count)
      accepted = surrogate_update(a, row, j, a.T.params[sc], S, eps, y, q, phi_p, phi, kmu, U, V, vol);
    n_acc += (unsigned)__popcll(__ballot(accepted));
    __builtin_amdgcn_sched_barrier(0);
  }
  // one atomic per wave
  if (lane == 0 && n_acc) atomicAdd(a.accepted, (unsigned long long)n_acc);
}

// ================================================================== cloth
// Element forces from the current directions (per element), then each vertex sums the forces of
// its incident elements in ascending element order (deterministic).
__global__ __launch_bounds__(256) void k_cloth_element_forces(ClothArgs a) {
  if (clock_skip(a.clk)) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  int fb = 0;
  if (e < a.ne) {
    const DevParams& P = a.P;
    const int pi = clamp_index(a.pos[P.mesh_e0 + e], a.n);
    float D[9], fv[9], M[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) D[i] = a.st[(PF + i) * a.cap + pi];
    cloth_element_forces(P, D, a.rest + 9 * (size_t)e, a.st[PVOL * a.cap + pi], fv, M, fb);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      a.elem_fv[9 * (size_t)e + i] = fv[i];
      a.elem_stress[9 * (size_t)e + i] = M[i];
    }
  }
  if (fb) atomicAdd(a.fallback, (unsigned long long)fb);
}

// The rest in-plane QR factor of each element, from its undeformed directions (set-up only).
__global__ __launch_bounds__(256) void k_cloth_rest(ClothArgs a, float* __restrict__ rest) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.ne) return;
  const int pi = clamp_index(a.pos[a.P.mesh_e0 + e], a.n);
  float D[9], Q[9], R[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) D[i] = a.st[(PF + i) * a.cap + pi];
  int fb = 0;
  qr3_guarded(D, Q, R, a.P.svd_ratio, fb);
  rest[9 * (size_t)e + 6] = R[0];
  rest[9 * (size_t)e + 7] = R[1];
  rest[9 * (size_t)e + 8] = R[4];
}

__global__ __launch_bounds__(256) void k_cloth_vertex_forces(ClothArgs a) {
  if (clock_skip(a.clk)) return;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.nv) return;
  float f[3] = {0.0f, 0.0f, 0.0f};
  for (int k = a.vcsr_off[v]; k < a.vcsr_off[v + 1]; ++k) {
    const int ent = a.vcsr[k];  // 3 e + corner
    const int e = ent / 3, c = ent - 3 * e;
    if (e < 0 || e >= a.ne) continue;
#pragma unroll
    for (int r = 0; r < 3; ++r) f[r] += a.elem_fv[9 * (size_t)e + 3 * c + r];
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) a.vert_force[3 * (size_t)v + r] = f[r];
}

// Fixed vertices (HybridSolver.cpp:513-550): for every in-grid node with positive weight for a fixed
// vertex, the 27 nodes (i, j, k) around it are stopped (v = v_BF = 0) through their linear index
// k r0 r1 + j r0 + i, which is checked against [0, gridNumber) only (:536-539). The union of those
// neighbourhoods is the stencil box widened by one node per axis; a neighbour one step outside the
// grid in x or y therefore stops the node its linear index names, e.g. (-1, j, k) stops
// (r0 - 1, j - 1, k), as in the reference.
template <int ORDER>
__global__ __launch_bounds__(64) void k_cloth_fixed(ClothArgs a) {
  if (clock_skip(a.clk)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nfixed) return;
  const DevParams& P = a.P;
  const int pi = clamp_index(a.pos[clamp_index(a.fixed[k], a.n)], a.n);
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float g = (a.st[(PX + d) * a.cap + pi] - P.gmin[d]) * P.inv_h[d];
    const int base = stencil_base<ORDER>(g, P.res[d]);
    lo[d] = base + ORDER + 1;
    hi[d] = base - 1;
    for (int o = 0; o <= ORDER; ++o) {
      float w, dw;
      spline<ORDER>(g - (float)(base + o), w, dw);
      if (w > 0.0f) {
        lo[d] = min(lo[d], base + o);
        hi[d] = max(hi[d], base + o);
      }
    }
    // the weight matrix holds in-grid nodes only (HybridSolver.cpp:44-47)
    lo[d] = max(lo[d], 0);
    hi[d] = min(hi[d], P.res[d] - 1);
  }
  if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return;
  const int64_t ng = (int64_t)P.res[0] * P.res[1] * P.res[2];
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int n2 = lo[2] - 1; n2 <= hi[2] + 1; ++n2)
    for (int n1 = lo[1] - 1; n1 <= hi[1] + 1; ++n1)
      for (int n0 = lo[0] - 1; n0 <= hi[0] + 1; ++n0) {
        const int64_t node = ((int64_t)n2 * P.res[1] + n1) * P.res[0] + n0;
        if (node < 0 || node >= ng) continue;
        a.gvel[2 * node] = z;
        a.gvel[2 * node + 1] = z;
      }
}

// After G2P: element position = centroid and velocity = mean of its vertices' new values
// (LagrangianMesh::updateElementPositions, HybridSolver.cpp:750-756), d1 = x2 - x1, d2 = x3 - x1
// (:584-590), d3 as updated by G2P, then the return mapping; the element's next key.
__global__ __launch_bounds__(256) void k_cloth_elements(ClothArgs a) {
  if (clock_skip(a.clk)) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  int fb = 0, oob = 0;
  if (e < a.ne) {
    const DevParams& P = a.P;
    const int id = P.mesh_e0 + e;
    // the element's and its vertices' slots, then all of their data in one round of loads
    const int pe = clamp_index(a.pos[id], a.n);
    int pv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) pv[c] = clamp_index(a.pos[clamp_index(a.faces[3 * e + c], a.n)], a.n);
    float xv[3][3], vv[3][3], d3[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      d3[d] = a.st[(PF + 3 * d + 2) * a.cap + pe];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xv[c][d] = a.st[(PX + d) * a.cap + pv[c]];
        vv[c][d] = a.st[(PV + d) * a.cap + pv[c]];
      }
    }
    float D[9], xe[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xe[d] = (xv[0][d] + xv[1][d] + xv[2][d]) / 3.0f;
      a.st[(PX + d) * a.cap + pe] = xe[d];
      a.st[(PV + d) * a.cap + pe] = (vv[0][d] + vv[1][d] + vv[2][d]) / 3.0f;
      D[3 * d] = xv[1][d] - xv[0][d];
      D[3 * d + 1] = xv[2][d] - xv[0][d];
      D[3 * d + 2] = d3[d];
    }
    cloth_return_map(P, D, fb);
#pragma unroll
    for (int i = 0; i < 9; ++i) a.st[(PF + i) * a.cap + pe] = D[i];
    a.keys[pe] = cell_key(P, a.thr, xe[0], xe[1], xe[2], oob);
  }
  if (fb) atomicAdd(a.fallback, (unsigned long long)fb);
  if (oob) atomicOr(a.status, MPM_STATUS_OUT_OF_DOMAIN);
}

// ================================================================== auxiliary kernels
// rho_p = sum_i w_ip m_i / h^3, V_p = m_p / rho_p (HybridSolver.cpp:242-249), at the current positions.
template <int ORDER>
__global__ __launch_bounds__(256) void k_volumes(SceneTable T, float* __restrict__ st, int64_t cap, int n,
                                                 const float* __restrict__ gm_all) {
  const PBlock pb = particle_block(T);
  const int i = pb.j;
  if (i >= pb.end) return;
  const DevParams& P = T.params[pb.s];
  const float* gm = gm_all + T.node_off[pb.s];
  // particles only: a cloth's vertices and elements keep their mesh volumes (HybridSolver.cpp:242-249)
  if (particle_kind(P, reinterpret_cast<const int*>(st + (size_t)PID * cap)[i]) != KIND_PARTICLE) return;
  float g[3];
  int base[3];
  float w[3][4], dw[3][4];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    g[d] = (st[(PX + d) * cap + i] - P.gmin[d]) * P.inv_h[d];
    base[d] = stencil_base<ORDER>(g[d], P.res[d]);
#pragma unroll
    for (int o = 0; o < ORDER + 1; ++o) spline<ORDER>(g[d] - (float)(base[d] + o), w[d][o], dw[d][o]);
  }
  float rho = 0.0f;
  for (int oz = 0; oz < ORDER + 1; ++oz)
    for (int oy = 0; oy < ORDER + 1; ++oy)
      for (int ox = 0; ox < ORDER + 1; ++ox) {
        int n0 = base[0] + ox, n1 = base[1] + oy, n2 = base[2] + oz;
        if (n0 < 0 || n1 < 0 || n2 < 0 || n0 >= P.res[0] || n1 >= P.res[1] || n2 >= P.res[2]) continue;
        rho += w[0][ox] * w[1][oy] * w[2][oz] * gm[((int64_t)n2 * P.res[1] + n1) * P.res[0] + n0];
      }
  rho /= P.h[0] * P.h[1] * P.h[2];
  st[PVOL * cap + i] = st[PM * cap + i] / rho;
}

// Copy planes [p0, p0+np) of the device order into caller order (out[plane][id] = st[plane][i]).
__global__ __launch_bounds__(256) void k_unpermute(const float* __restrict__ st, int64_t cap, int n, int p0, int np,
                                                   float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int id = reinterpret_cast<const int*>(st + (size_t)PID * cap)[i];
  if (id < 0 || id >= n) return;
  for (int p = 0; p < np; ++p) out[(size_t)p * n + id] = st[(p0 + p) * cap + i];
}

// Inverse of k_unpermute: planes [p0, p0+np) of the device order from caller order
// (st[plane][pos[id]] = in[plane][id]).
__global__ __launch_bounds__(256) void k_permute_in(float* __restrict__ st, int64_t cap, int n, int p0, int np,
                                                    const float* __restrict__ in, const int* __restrict__ pos) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int i = pos[id];
  if (i < 0 || i >= n) return;
  for (int p = 0; p < np; ++p) st[(p0 + p) * cap + i] = in[(size_t)p * n + id];
}

// Dense planar copy of one grid quantity (0 = m, 1 = momentum, 2 = v, 3 = v_BF), with nodes
// outside the active blocks set to zero.
__global__ __launch_bounds__(256) void k_grid_export(SceneTable T, const float* __restrict__ grid, int64_t ng, int which,
                                                     const int* __restrict__ block_slot, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ng) return;
  const int sc = scene_of(T.node_off, T.nscenes, i);
  const DevParams& P = T.params[sc];
  const int64_t li = i - T.node_off[sc];
  const int n0 = (int)(li % P.res[0]), n1 = (int)((li / P.res[0]) % P.res[1]), n2 = (int)(li / ((int64_t)P.res[0] * P.res[1]));
  const int b = T.tile_off[sc] + ((n2 >> 2) * P.tiles[1] + (n1 >> 2)) * P.tiles[0] + (n0 >> 2);
  const bool on = block_slot[b] >= 0;
  if (which == 0) {
    out[i] = on ? grid[GRID_M * ng + i] : 0.0f;
  } else if (which == 1) {
    for (int d = 0; d < 3; ++d) out[d * ng + i] = on ? grid[(GRID_MOM + d) * ng + i] : 0.0f;
  } else {
    const float* vel = grid + GRID_VEL * ng + 8 * i;
    for (int d = 0; d < 3; ++d) out[d * ng + i] = on ? vel[which == 2 ? d : 3 + d] : 0.0f;
  }
}

__global__ __launch_bounds__(256) void k_iota(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

// Number of nodes with m > 0 in the active blocks.
__global__ __launch_bounds__(64) void k_count_nodes(SceneTable T, const float* __restrict__ gm, const int* __restrict__ counters,
                                                    const int* __restrict__ active_blocks, unsigned long long* __restrict__ out) {
  const int nblk = counters[CNT_BLOCKS];
  const int lx = threadIdx.x & 3, ly = (threadIdx.x >> 2) & 3, lz = threadIdx.x >> 4;
  unsigned long long c = 0;
  for (int s = blockIdx.x; s < nblk; s += gridDim.x) {
    const int gblk = active_blocks[s];
    const int sc = scene_of(T.tile_off, T.nscenes, gblk);
    const DevParams& P = T.params[sc];
    const int blk = gblk - T.tile_off[sc];
    const int b0 = blk % P.tiles[0], b1 = (blk / P.tiles[0]) % P.tiles[1], b2 = blk / (P.tiles[0] * P.tiles[1]);
    const int n0 = 4 * b0 + lx, n1 = 4 * b1 + ly, n2 = 4 * b2 + lz;
    if (n0 >= P.res[0] || n1 >= P.res[1] || n2 >= P.res[2]) continue;
    if (gm[T.node_off[sc] + ((int64_t)n2 * P.res[1] + n1) * P.res[0] + n0] > 0.0f) ++c;
  }
  if (c) atomicAdd(out, c);
}

// Plain copy (measurement reference for the PMC byte counters).
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n4) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_fill_int(int* __restrict__ p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ================================================================== launchers
static inline int blocks_for(int64_t n, int t) { return (int)((n + t - 1) / t); }

void launch_pos_of_ids(hipStream_t s, const int* pid, int n, int* pos) {
  if (n > 0) hipLaunchKernelGGL(k_caller_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, nullptr, pid, n, nullptr, pos, nullptr);
}

void launch_keys(hipStream_t s, const SceneTable& T, int nblocks, const float* st, int64_t cap, int n,
                 const float* thr, uint32_t* keys, int* status) {
  if (n > 0) hipLaunchKernelGGL(k_keys, dim3(nblocks), dim3(256), 0, s, T, st, cap, n, thr, keys, status);
}

void launch_stress_init(hipStream_t s, const SceneTable& T, int nblocks, float* st, int64_t cap, int n,
                        unsigned long long* fallback) {
  if (n > 0) hipLaunchKernelGGL(k_stress_init, dim3(nblocks), dim3(256), 0, s, T, st, cap, n, fallback);
}

void launch_sort(hipStream_t s, const SortBufs& b, const uint32_t* keys_in, int n, int passes, int mode) {
  if (n <= 0) return;
  {
    // bins: counts (the rank is used by the counting sort) and their scan, in both modes
    uint32_t* rank = (passes & 1) ? b.k2 : b.k1;
    const int nscan = blocks_for(b.ncells, SCAN_TILE);
    hipLaunchKernelGGL(k_cell_count, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys_in, n, b.cell_count, rank,
                       b.reset, mode == SORT_COUNTING ? 1 : 0, b.occ, b.occ_clear, nscan, b.clk);
    hipLaunchKernelGGL(k_scan_sums, dim3(nscan), dim3(256), 0, s, b.cell_count, b.ncells, b.scan_agg, b.occ, b.clk);
    hipLaunchKernelGGL(k_cell_scan, dim3(nscan), dim3(256), 0, s, b.cell_count, b.ncells, b.cell_start, b.scan_agg,
                       b.occ, b.counters, b.clk);
  }
  if (mode == SORT_COUNTING) {
    uint32_t* kf = (passes & 1) ? b.k1 : b.k2;  // the result goes where the digit passes put it
    int* vf = (passes & 1) ? b.v1 : b.v2;
    const uint32_t* rank = (passes & 1) ? b.k2 : b.k1;
    int* tmp = (passes & 1) ? b.v2 : b.v1;
    hipLaunchKernelGGL(k_cell_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys_in, rank, n, b.cell_start, b.pid,
                       tmp, b.tslot, b.clk);
    hipLaunchKernelGGL(k_cell_rank, dim3(blocks_for(n, 256)), dim3(256), 0, s, b.T, keys_in, tmp, b.tslot, n,
                       b.cell_start, kf, vf, b.cell_count, b.tile_slot, b.src, b.clk);
    return;
  }
  // the digit passes are stable in their input order, which must be the caller order: the keys
  // (by physical slot) go to caller order first, and the caller id -> slot map comes with them
  uint32_t* kc = reinterpret_cast<uint32_t*>(b.tslot);
  hipLaunchKernelGGL(k_caller_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys_in, b.pid, n, kc, b.pos, b.clk);
  const int nb = blocks_for(n, SORT_TILE);
  hipLaunchKernelGGL(k_sort_hist, dim3(blocks_for(n, HIST_TILE)), dim3(256), 0, s, kc, n, passes, b.hist, b.clk);
  const uint32_t* kin = kc;
  const int* vin = nullptr;
  for (int p = 0; p < passes; ++p) {
    uint32_t* ko = (p & 1) ? b.k2 : b.k1;
    int* vo = (p & 1) ? b.v2 : b.v1;
    hipLaunchKernelGGL(k_sort_pass, dim3(nb), dim3(256), 0, s, kin, vin, ko, vo, n, p, b.hist, b.tickets,
                       b.status + (size_t)p * nb * 256, b.clk);
    kin = ko;
    vin = vo;
  }
}

void launch_fill_bins(hipStream_t s, const uint32_t* ks, int n, int64_t ncells, int* cell_start) {
  if (ncells <= 0) return;
  hipLaunchKernelGGL(k_fill_bins, dim3(blocks_for(ncells + 1, 256)), dim3(256), 0, s, ks, n, ncells, cell_start);
}

void launch_reset(hipStream_t s, const ResetArgs& r) {
  int64_t work = r.n_zero;
  if (work < r.max_tiles) work = r.max_tiles;
  int grid = blocks_for(work, 256);
  if (grid > 4096) grid = 4096;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_reset, dim3(grid), dim3(256), 0, s, r.zero_words, r.n_zero, r.counters, r.active_tiles,
                     r.tile_slot, r.active_blocks, r.block_slot, r.clk);
}

void launch_bounds(hipStream_t s, const SceneTable& T, int nblocks, const uint32_t* ks, int n, int64_t ntiles,
                   int* cell_count, int* tile_slot, int* active_tiles, int* block_slot, int* active_blocks,
                   int* counters, const int* perm, const int* pos, int* src, const Clock* clk, int mode) {
  if (n <= 0) return;
  if (mode != SORT_COUNTING)  // the counting sort's rank kernel has done this pass
    hipLaunchKernelGGL(k_bounds, dim3(nblocks), dim3(256), 0, s, T, ks, n, cell_count, tile_slot, perm, pos, src, clk);
  hipLaunchKernelGGL(k_claim, dim3(blocks_for(ntiles, 256)), dim3(256), 0, s, T, ntiles, tile_slot, active_tiles,
                     block_slot, active_blocks, counters, clk);
}

template <int RANKS>
static void launch_p2g_r(hipStream_t s, const P2GArgs& a, int order, int mode, int grid) {
  if (mode == P2G_MOM_STRESS) {
    if (order == 2) hipLaunchKernelGGL((k_p2g<2, P2G_MOM_STRESS, RANKS>), dim3(grid), dim3(64 * P2G_WAVES), 0, s, a);
    else hipLaunchKernelGGL((k_p2g<3, P2G_MOM_STRESS, RANKS>), dim3(grid), dim3(64 * P2G_WAVES), 0, s, a);
  } else if (mode == P2G_MOM) {
    if (order == 2) hipLaunchKernelGGL((k_p2g<2, P2G_MOM, RANKS>), dim3(grid), dim3(64 * P2G_WAVES), 0, s, a);
    else hipLaunchKernelGGL((k_p2g<3, P2G_MOM, RANKS>), dim3(grid), dim3(64 * P2G_WAVES), 0, s, a);
  } else {
    if (order == 2) hipLaunchKernelGGL((k_p2g<2, P2G_FORCE, RANKS>), dim3(grid), dim3(64 * P2G_WAVES), 0, s, a);
    else hipLaunchKernelGGL((k_p2g<3, P2G_FORCE, RANKS>), dim3(grid), dim3(64 * P2G_WAVES), 0, s, a);
  }
}

template <int RANKS>
static void launch_p2g_l(hipStream_t s, const P2GArgs& a, int order, int mode, int grid) {
  if (mode == P2G_MOM_STRESS) {
    if (order == 2) hipLaunchKernelGGL((k_p2g_layer<2, P2G_MOM_STRESS, RANKS>), dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_p2g_layer<3, P2G_MOM_STRESS, RANKS>), dim3(grid), dim3(256), 0, s, a);
  } else if (mode == P2G_MOM) {
    if (order == 2) hipLaunchKernelGGL((k_p2g_layer<2, P2G_MOM, RANKS>), dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_p2g_layer<3, P2G_MOM, RANKS>), dim3(grid), dim3(256), 0, s, a);
  } else {
    if (order == 2) hipLaunchKernelGGL((k_p2g_layer<2, P2G_FORCE, RANKS>), dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_p2g_layer<3, P2G_FORCE, RANKS>), dim3(grid), dim3(256), 0, s, a);
  }
}

void launch_p2g(hipStream_t s, const P2GArgs& a, int order, int mode, int grid, int ranks, int pass) {
  if (pass == P2G_LAYERS) {
    if (ranks == 2) launch_p2g_l<2>(s, a, order, mode, grid);
    else launch_p2g_l<4>(s, a, order, mode, grid);
  } else {
    if (ranks == 4) launch_p2g_r<4>(s, a, order, mode, grid);
    else launch_p2g_r<8>(s, a, order, mode, grid);
  }
}

void launch_clock_update(hipStream_t s, Clock* c, int init) {
  hipLaunchKernelGGL(k_clock_update, dim3(1), dim3(64), 0, s, c, init);
}

void launch_clock_end(hipStream_t s, Clock* c) { hipLaunchKernelGGL(k_clock_end, dim3(1), dim3(64), 0, s, c); }

void launch_grid(hipStream_t s, const GridArgs& a, int mode, int grid) {
  if (mode == GRID_FUSED) hipLaunchKernelGGL(k_grid<GRID_FUSED>, dim3(grid), dim3(64), 0, s, a);
  else if (mode == GRID_NORMALIZE) hipLaunchKernelGGL(k_grid<GRID_NORMALIZE>, dim3(grid), dim3(64), 0, s, a);
  else hipLaunchKernelGGL(k_grid<GRID_FORCE_UPDATE>, dim3(grid), dim3(64), 0, s, a);
}

void launch_g2p(hipStream_t s, const G2PArgs& a, int order, int nblocks, hipEvent_t ev_start, hipEvent_t ev_stop) {
  if (a.n <= 0) return;
  if (ev_start) {  // timed launch: the events take the dispatch's own start and end times
    if (order == 2) hipExtLaunchKernelGGL(k_g2p<2>, dim3(nblocks), dim3(256), 0, s, ev_start, ev_stop, 0, a);
    else hipExtLaunchKernelGGL(k_g2p<3>, dim3(nblocks), dim3(256), 0, s, ev_start, ev_stop, 0, a);
    return;
  }
  if (order == 2) hipLaunchKernelGGL(k_g2p<2>, dim3(nblocks), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_g2p<3>, dim3(nblocks), dim3(256), 0, s, a);
}

void launch_volumes(hipStream_t s, const SceneTable& T, int order, int nblocks, float* st, int64_t cap, int n,
                    const float* gm) {
  if (n <= 0) return;
  if (order == 2) hipLaunchKernelGGL(k_volumes<2>, dim3(nblocks), dim3(256), 0, s, T, st, cap, n, gm);
  else hipLaunchKernelGGL(k_volumes<3>, dim3(nblocks), dim3(256), 0, s, T, st, cap, n, gm);
}

void launch_unpermute(hipStream_t s, const float* st, int64_t cap, int n, int p0, int np, float* out) {
  if (n > 0) hipLaunchKernelGGL(k_unpermute, dim3(blocks_for(n, 256)), dim3(256), 0, s, st, cap, n, p0, np, out);
}

void launch_permute_in(hipStream_t s, float* st, int64_t cap, int n, int p0, int np, const float* in, const int* pos) {
  if (n > 0) hipLaunchKernelGGL(k_permute_in, dim3(blocks_for(n, 256)), dim3(256), 0, s, st, cap, n, p0, np, in, pos);
}

void launch_grid_export(hipStream_t s, const SceneTable& T, const float* grid, int64_t ng, int which, const int* block_slot,
                        float* out) {
  hipLaunchKernelGGL(k_grid_export, dim3(blocks_for(ng, 256)), dim3(256), 0, s, T, grid, ng, which, block_slot, out);
}

void launch_iota(hipStream_t s, int* p, int n) {
  if (n > 0) hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n);
}

void launch_count_nodes(hipStream_t s, const SceneTable& T, const float* gm, const int* counters, const int* active_blocks,
                        unsigned long long* out, int grid) {
  hipLaunchKernelGGL(k_count_nodes, dim3(grid), dim3(64), 0, s, T, gm, counters, active_blocks, out);
}

void launch_copy(hipStream_t s, const float4* src, float4* dst, int64_t n4) {
  int grid = blocks_for(n4, 256);
  if (grid > 8192) grid = 8192;
  if (n4 > 0) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, src, dst, n4);
}

static size_t surrogate_smem() { return sizeof(MLPShared); }

void surrogate_set_smem_limit() {
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_surrogate),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)surrogate_smem());
}

void launch_position_lists(hipStream_t s, const ListArgs& a) {
  if (a.nwaves <= 0 || a.nlists <= 0) return;
  const int ng = (a.nwaves + 63) / 64;
  hipLaunchKernelGGL(k_list_sums, dim3((ng + 3) / 4), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_list_write, dim3(ng), dim3(256), 0, s, a);
}

void launch_surrogate(hipStream_t s, const SurrogateArgs& a, int n_cu, hipEvent_t* ev) {
  if (ev && ev[0]) {  // a timed launch: the dispatch records its own start and end
    hipExtLaunchKernelGGL(k_surrogate, dim3(n_cu), dim3(MLP_FWD_THREADS), surrogate_smem(), s, ev[0], ev[1], 0, a);
    return;
  }
  hipLaunchKernelGGL(k_surrogate, dim3(n_cu), dim3(MLP_FWD_THREADS), surrogate_smem(), s, a);
}

void launch_apex_finish(hipStream_t s, const ApexArgs& a, int grid) {
  hipLaunchKernelGGL(k_apex_finish, dim3(grid), dim3(256), 0, s, a);
}

void launch_fill_int(hipStream_t s, int* p, int64_t n, int v) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
}

void launch_cloth_rest(hipStream_t s, const ClothArgs& a, float* rest) {
  if (a.ne > 0) hipLaunchKernelGGL(k_cloth_rest, dim3(blocks_for(a.ne, 256)), dim3(256), 0, s, a, rest);
}

void launch_cloth_forces(hipStream_t s, const ClothArgs& a) {
  if (a.ne <= 0) return;
  hipLaunchKernelGGL(k_cloth_element_forces, dim3(blocks_for(a.ne, 256)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_cloth_vertex_forces, dim3(blocks_for(a.nv, 256)), dim3(256), 0, s, a);
}

void launch_cloth_fixed(hipStream_t s, const ClothArgs& a) {
  if (a.nfixed <= 0) return;
  if (a.P.order == 2) hipLaunchKernelGGL(k_cloth_fixed<2>, dim3(blocks_for(a.nfixed, 64)), dim3(64), 0, s, a);
  else hipLaunchKernelGGL(k_cloth_fixed<3>, dim3(blocks_for(a.nfixed, 64)), dim3(64), 0, s, a);
}

void launch_cloth_elements(hipStream_t s, const ClothArgs& a) {
  if (a.ne > 0) hipLaunchKernelGGL(k_cloth_elements, dim3(blocks_for(a.ne, 256)), dim3(256), 0, s, a);
}

}  // namespace mpm
