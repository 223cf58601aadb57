// mpm_mlp_device.hpp -- device building blocks of the return-map surrogate MLP (BF16 MFMA, FP32
// accumulation): BF16 conversion, SiLU, LDS weight staging, the 16-row wave GEMM and forward pass.
#pragma once
#include <hip/hip_runtime.h>

#include "mpm_mlp.h"

namespace mpm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// SiLU and its derivative with the hardware exp2 and reciprocal (about 1 ulp each); the activations
// are rounded to BF16 right after, so the last FP32 bits do not matter.
__device__ __forceinline__ float sigmoid_fast(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid_fast(z); }
// log2(e) silu(z) from z' = log2(e) z (the forward pass's scaled pre-activations)
__device__ __forceinline__ float silu_log2(float zs) { return zs * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-zs)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigmoid_fast(z);
  return s * (1.0f + z * (1.0f - s));
}

// LDS row strides (in bf16 elements) padded against bank conflicts
constexpr int LDW_IN = MLP_IN_PAD + 8;   // 40
constexpr int LDW_H = MLP_HIDDEN + 8;    // 136
constexpr int LDA = MLP_HIDDEN + 8;      // activation rows of the backward pass

// Staged weights. Every row is stored in the K order of the forward MFMA operands: slot
// 32 s + 8 q + j (K step s, lane group q = lane / 16, j = 0..7) holds input feature
// 32 s + 4 q + j for j < 4 and 32 s + 16 + 4 q + (j - 4) for j >= 4. Those are the features that
// lane group q holds, four from each of the output tiles 2 s and 2 s + 1, after a 16x16 MFMA tile
// product (lane i: rows 4 (i / 16) .. +4 of column i % 16), so one layer's outputs are the next
// layer's operand without any data movement.
struct MLPShared {
  uint16_t w1[MLP_HIDDEN * LDW_IN];
  uint16_t wh[3][MLP_HIDDEN * LDW_H];
  uint16_t w5[MLP_OUT_PAD * LDW_H];
  float b[4 * MLP_HIDDEN + MLP_OUT_PAD];
};

// Slot of the first of four consecutive input features k0 .. k0 + 3 (k0 a multiple of 4).
__device__ __forceinline__ int mlp_slot4(int k0) {
  const int r = k0 & 31;
  return (k0 & ~31) + (r < 16 ? 2 * r : 2 * (r - 16) + 4);
}

static __device__ void stage_rows(uint16_t* dst, int ldw, const uint16_t* __restrict__ src, int rows, int K) {
  // 4 features (8 bytes) per copy: they stay consecutive in the staged order
  const int chunks = rows * K / 4;
  for (int i = threadIdx.x; i < chunks; i += blockDim.x) {
    const int o = i / (K / 4), k0 = 4 * (i % (K / 4));
    *reinterpret_cast<uint2*>(dst + o * ldw + mlp_slot4(k0)) = reinterpret_cast<const uint2*>(src)[i];
  }
}

static __device__ void stage_weights(MLPShared& s, const uint16_t* __restrict__ wbf, const float* __restrict__ master) {
  stage_rows(s.w1, LDW_IN, wbf + MLP_BF_W1, MLP_HIDDEN, MLP_IN_PAD);
  for (int l = 0; l < 3; ++l)
    stage_rows(s.wh[l], LDW_H, wbf + MLP_BF_WH + l * MLP_HIDDEN * MLP_HIDDEN, MLP_HIDDEN, MLP_HIDDEN);
  stage_rows(s.w5, LDW_H, wbf + MLP_BF_W5, MLP_OUT_PAD, MLP_HIDDEN);
  for (int l = 0; l < 4; ++l)
    for (int i = threadIdx.x; i < MLP_HIDDEN; i += blockDim.x)
      s.b[l * MLP_HIDDEN + i] = master[mlp_bias_offset(l) + i] * MLP_LOG2E;
  for (int i = threadIdx.x; i < MLP_OUT_PAD; i += blockDim.x) s.b[4 * MLP_HIDDEN + i] = master[mlp_bias_offset(4) + i];
}

// acc[t] (16 rows x 16 cols of output tile t) += A (16 x 32*KS, row stride lda) * W^T where W is
// [16*NT][32*KS] with row stride ldw: every operand is 8 contiguous bf16 per lane (backward pass).
template <int NT, int KS>
__device__ __forceinline__ void wave_gemm_nt(const uint16_t* A, int lda, const uint16_t* W, int ldw, f32x4* acc) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + r * lda + 32 * s + kq);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(W + (16 * t + r) * ldw + 32 * s + kq);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
  }
}

// One step of the layer pipeline over NT sample tiles (tiles T0 .. T0 + NT - 1 of the wave). Unit u
// of a layer is its output feature tiles 2u and 2u + 1 (the next layer's K step u), accumulated over
// the layer's KS K steps in ascending order into acc. The step issues those MFMAs and, between them,
// the SiLU of the previous unit's pre-activations pacc (sched_barrier keeps the order), so that the
// matrix cores and the vector ALUs of one wave are busy at the same time; the SiLU results go,
// rounded to BF16, to nb[T][pu]. Without a previous unit (has_prev false) it only issues the MFMAs.
// FEED: the previous unit is the last of the previous layer and completes this layer's operand
// (nb is b, pu = KS - 1): its SiLU runs between the MFMAs of the first KS - 1 K steps and is stored
// before the last K step reads it.
template <int KS, int NT, bool FEED>
__device__ __forceinline__ void fwd_step(const uint16_t* W, int ldw, const float* bias, int u, bf16x8 (&b)[NT][4],
                                         f32x4 (&acc)[NT][2], const f32x4 (&pacc)[NT][2], bf16x8 (&nb)[NT][4], int pu,
                                         bool has_prev, int c, int q) {
  static_assert(NT == 2 || NT == 4, "MFMAs go in pairs of sample tiles");
  static_assert(!FEED || KS > 1, "a fed step spreads the SiLU over all but its last K step");
  constexpr int NSP = FEED ? KS - 1 : KS;                 // K steps the SiLU is spread over
  constexpr int VPP = (8 * NT + NSP * NT - 1) / (NSP * NT);  // SiLU values per pair of MFMAs
  float sv[NT][2][4];
  auto pack = [&]() {
#pragma unroll
    for (int T = 0; T < NT; ++T) {
      bf16x8 o;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        o[j] = (__bf16)sv[T][0][j];
        o[4 + j] = (__bf16)sv[T][1][j];
      }
      nb[T][pu] = o;
    }
  };
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const f32x4 bb = *reinterpret_cast<const f32x4*>(&bias[16 * (2 * u + p) + 4 * q]);
#pragma unroll
    for (int T = 0; T < NT; ++T) acc[T][p] = bb;
  }
  bf16x8 a[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) a[p] = *reinterpret_cast<const bf16x8*>(W + (16 * (2 * u + p) + c) * ldw + 8 * q);
  int v0 = 0;
#pragma unroll
  for (int ss = 0; ss < KS; ++ss) {
    bf16x8 an[2];
    if (ss + 1 < KS) {
#pragma unroll
      for (int p = 0; p < 2; ++p)
        an[p] = *reinterpret_cast<const bf16x8*>(W + (16 * (2 * u + p) + c) * ldw + 32 * (ss + 1) + 8 * q);
    }
#pragma unroll
    for (int m = 0; m < 2 * NT; m += 2) {
      // MFMAs (p, T) = (m / NT, m % NT) and (m / NT, m % NT + 1)
      const int p = m / NT, T0 = m % NT;
      acc[T0][p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], b[T0][ss], acc[T0][p], 0, 0, 0);
      acc[T0 + 1][p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], b[T0 + 1][ss], acc[T0 + 1][p], 0, 0, 0);
      if (has_prev && ss < NSP) {
#pragma unroll
        for (int v = v0; v < v0 + VPP && v < 8 * NT; ++v) {
          const int T = v >> 3, pp = (v >> 2) & 1, j = v & 3;
          sv[T][pp][j] = silu_log2(pacc[T][pp][j]);
        }
        v0 += VPP;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    if (ss + 1 < KS) {
      a[0] = an[0];
      a[1] = an[1];
    }
    if (FEED && ss == KS - 2) pack();
  }
  if (!FEED && has_prev) pack();
}

// The pre-activations of unit u of layer l for tiles T0 .. T0 + NT - 1, unscaled, into zsave (training).
template <int NT>
__device__ __forceinline__ void fwd_zsave(const f32x4 (&acc)[NT][2], int T0, int l, int u, int q,
                                          float* __restrict__ zsave, int64_t row0, int64_t nrows, int c) {
#pragma unroll
  for (int T = 0; T < NT; ++T) {
    const int64_t r = row0 + 16 * (T0 + T) + c;
    if (r < nrows) {
      float* zr = zsave + (((r >> 4) * 4 + l) * 16 + (r & 15)) * MLP_HIDDEN + 4 * q;
#pragma unroll
      for (int p = 0; p < 2; ++p) *reinterpret_cast<f32x4*>(zr + 16 * (2 * u + p)) = acc[T][p] * MLP_LN2;
    }
  }
}

// SiLU and BF16 rounding of a unit on its own (the last unit of a layer before the next layer's first).
template <int NT>
__device__ __forceinline__ void fwd_silu(const f32x4 (&acc)[NT][2], bf16x8 (&nb)[NT][4], int u) {
#pragma unroll
  for (int T = 0; T < NT; ++T) {
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      o[j] = (__bf16)silu_log2(acc[T][0][j]);
      o[4 + j] = (__bf16)silu_log2(acc[T][1][j]);
    }
    nb[T][u] = o;
  }
}

// Forward pass of 64 samples held by one wave, one per lane: lane L passes the raw inputs x of
// sample L (eps(3), phi, K/mu) and receives its 3 normalised outputs in y. Each layer computes
// Z^T = W A^T on MFMA with the staged weights as the A operand (16 output features x 32 inputs)
// and the activations as the B operand (32 inputs x 16 samples), for the wave's four 16-sample
// tiles T, NT at a time (NT = 2: two passes, with half the registers of one pass over all four):
// lane i of tile T holds output features 16 t + 4 (i / 16) + 0..3 of sample 16 T + i % 16,
// which SiLU and BF16 rounding turn into the next layer's operand in place (pre-activations and
// activations scaled by log2(e), see MLP_LOG2E). When zsave is not null, the FP32 pre-activations z
// (unscaled) of sample r = row0 + L < nrows go to zsave[(r / 16) * 4 + layer][r % 16][128].
// The layers run as a pipeline of units (fwd_step): the MFMAs of one unit are issued between the
// SiLU instructions of the one before it.
template <int NT>
__device__ void wave_forward64(const MLPShared& s, const float x[5], float y[3], float* __restrict__ zsave,
                               int64_t row0, int64_t nrows) {
  const int lane = threadIdx.x & 63, c = lane & 15, q = lane >> 4;
  const float xn[5] = {x[0] * (1.0f / MLP_EPS_SCALE), x[1] * (1.0f / MLP_EPS_SCALE), x[2] * (1.0f / MLP_EPS_SCALE),
                       (x[3] - MLP_PHI_MID) * (1.0f / MLP_PHI_SPAN), (x[4] - MLP_KMU_MID) * (1.0f / MLP_KMU_SPAN)};
  float yv[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
  for (int T0 = 0; T0 < 4; T0 += NT) {
    bf16x8 b[NT][4];
#pragma unroll
    for (int T = 0; T < NT; ++T) {
      float v[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) v[k] = __shfl(xn[k], 16 * (T0 + T) + c);
      // layer-1 slots: group 0 holds inputs 0..3 (slots 0..3), group 1 input 4 (slot 0)
      bf16x8 o;
#pragma unroll
      for (int j = 0; j < 8; ++j) o[j] = (__bf16)0.0f;
      if (q == 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] = (__bf16)v[j];
      } else if (q == 1) {
        o[0] = (__bf16)v[4];
      }
      b[T][0] = o;
#pragma unroll
      for (int ss = 1; ss < 4; ++ss) b[T][ss] = o;  // not read by layer 1
    }
    // Units (l, u) in order; unit (l, u) feeds K step u of layer l + 1. Step (l, u) issues the MFMAs
    // of unit (l, u) and, between them, the SiLU of the unit before it. The last unit of a layer
    // completes the next layer's operand (its K step 3): its SiLU runs under the first three K steps
    // of the next layer's first unit (FEED), the last unit of the network's on its own.
    f32x4 acc[2][NT][2];
    bf16x8 nb[NT][4];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const uint16_t* W = l == 0 ? s.w1 : s.wh[l - 1];
      const int ldw = l == 0 ? LDW_IN : LDW_H;
      const float* bias = s.b + l * MLP_HIDDEN;
      if (l > 0) {
#pragma unroll
        for (int T = 0; T < NT; ++T)
#pragma unroll
          for (int ss = 0; ss < 3; ++ss) b[T][ss] = nb[T][ss];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (l == 0) fwd_step<1, NT, false>(W, ldw, bias, u, b, acc[u & 1], acc[(u - 1) & 1], nb, u - 1, u > 0, c, q);
        else if (u == 0) fwd_step<4, NT, true>(W, ldw, bias, 0, b, acc[0], acc[1], b, 3, true, c, q);
        else fwd_step<4, NT, false>(W, ldw, bias, u, b, acc[u & 1], acc[(u - 1) & 1], nb, u - 1, true, c, q);
        if (zsave) fwd_zsave<NT>(acc[u & 1], T0, l, u, q, zsave, row0, nrows, c);
      }
    }
    fwd_silu<NT>(acc[1], nb, 3);
#pragma unroll
    for (int T = 0; T < NT; ++T)
#pragma unroll
      for (int ss = 0; ss < 4; ++ss) b[T][ss] = nb[T][ss];
    f32x4 out[NT];
    {
      const f32x4 bb = *reinterpret_cast<const f32x4*>(&s.b[4 * MLP_HIDDEN + 4 * q]);
#pragma unroll
      for (int T = 0; T < NT; ++T) out[T] = bb;
    }
#pragma unroll
    for (int ss = 0; ss < 4; ++ss) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(s.w5 + c * LDW_H + 32 * ss + 8 * q);
#pragma unroll
      for (int T = 0; T < NT; ++T) out[T] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[T][ss], out[T], 0, 0, 0);
    }
    // outputs 0..2 of sample 16 T + c sit in lane c (group 0) of tile T
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int T = 0; T < NT; ++T) {
        const float u = __shfl(out[T][k], c);
        if (q == T0 + T) yv[k] = u;
      }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) y[k] = yv[k];
}

}  // namespace mpm
