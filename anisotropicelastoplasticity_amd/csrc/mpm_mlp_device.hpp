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

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// LDS row strides (in bf16 elements) padded against bank conflicts
constexpr int LDW_IN = MLP_IN_PAD + 8;   // 40
constexpr int LDW_H = MLP_HIDDEN + 8;    // 136
constexpr int LDA = MLP_HIDDEN + 8;      // activation rows

struct MLPShared {
  uint16_t w1[MLP_HIDDEN * LDW_IN];
  uint16_t wh[3][MLP_HIDDEN * LDW_H];
  uint16_t w5[MLP_OUT_PAD * LDW_H];
  float b[4 * MLP_HIDDEN + MLP_OUT_PAD];
};

static __device__ void stage_weights(MLPShared& s, const uint16_t* __restrict__ wbf, const float* __restrict__ master) {
  // forward BF16 copy [N][K] per layer -> padded LDS rows
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < MLP_HIDDEN * MLP_IN_PAD; i += nt) s.w1[(i / MLP_IN_PAD) * LDW_IN + i % MLP_IN_PAD] = wbf[MLP_BF_W1 + i];
  for (int l = 0; l < 3; ++l)
    for (int i = tid; i < MLP_HIDDEN * MLP_HIDDEN; i += nt)
      s.wh[l][(i / MLP_HIDDEN) * LDW_H + i % MLP_HIDDEN] = wbf[MLP_BF_WH + l * MLP_HIDDEN * MLP_HIDDEN + i];
  for (int i = tid; i < MLP_OUT_PAD * MLP_HIDDEN; i += nt) s.w5[(i / MLP_HIDDEN) * LDW_H + i % MLP_HIDDEN] = wbf[MLP_BF_W5 + i];
  for (int l = 0; l < 4; ++l)
    for (int i = tid; i < MLP_HIDDEN; i += nt) s.b[l * MLP_HIDDEN + i] = master[mlp_bias_offset(l) + i];
  for (int i = tid; i < MLP_OUT_PAD; i += nt) s.b[4 * MLP_HIDDEN + i] = master[mlp_bias_offset(4) + i];
}

// acc[t] (16 rows x 16 cols of output tile t) += A (16 x 32*KS, row stride lda) * W^T where W is
// [16*NT][32*KS] with row stride ldw: every operand is 8 contiguous bf16 per lane.
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

// Normalised input row of one sample into the wave's activation rows (row = lane & 15 is written
// by lanes 0..15; the rest of the padded row is zero).
__device__ __forceinline__ void mlp_input_row(uint16_t* Arow, const float* in5) {
  Arow[0] = f2bf(in5[0] * (1.0f / MLP_EPS_SCALE));
  Arow[1] = f2bf(in5[1] * (1.0f / MLP_EPS_SCALE));
  Arow[2] = f2bf(in5[2] * (1.0f / MLP_EPS_SCALE));
  Arow[3] = f2bf((in5[3] - MLP_PHI_MID) * (1.0f / MLP_PHI_SPAN));
  Arow[4] = f2bf((in5[4] - MLP_KMU_MID) * (1.0f / MLP_KMU_SPAN));
  for (int c = 5; c < MLP_IN_PAD; ++c) Arow[c] = 0;
}

// Forward pass of 16 rows held by one wave. `act` is the wave's [16][LDA] bf16 scratch whose first
// MLP_IN_PAD columns hold the input rows on entry. Returns the 3 outputs of row (lane & 15) in
// lanes 0..15 through out3 (others undefined). When zsave is not null, the pre-activations of the 4
// hidden layers are stored to zsave[layer][row][128] (FP32).
static __device__ void wave_forward16(const MLPShared& s, uint16_t* act, float* out3, float* zsave, int zstride_rows) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, rq = (lane >> 4) * 4;
  f32x4 acc[8];
  // layer 1
  for (int l = 0; l < 4; ++l) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const float bb = s.b[l * MLP_HIDDEN + 16 * t + col];
      acc[t] = f32x4{bb, bb, bb, bb};
    }
    if (l == 0) wave_gemm_nt<8, 1>(act, LDA, s.w1, LDW_IN, acc);
    else wave_gemm_nt<8, 4>(act, LDA, s.wh[l - 1], LDW_H, acc);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float z = acc[t][j];
        const int row = rq + j, c = 16 * t + col;
        if (zsave) zsave[(size_t)(l * zstride_rows + row) * MLP_HIDDEN + c] = z;
        act[row * LDA + c] = f2bf(silu(z));
      }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  f32x4 o[1];
  {
    const float bb = s.b[4 * MLP_HIDDEN + col];
    o[0] = f32x4{bb, bb, bb, bb};
  }
  wave_gemm_nt<1, 4>(act, LDA, s.w5, LDW_H, o);
  // redistribute: output (row, col) lives in lane col + 16*(row/4), element row%4
  __builtin_amdgcn_wave_barrier();
  float* scratch = reinterpret_cast<float*>(act);  // 16 x 16 floats fit in the activation rows
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
  for (int j = 0; j < 4; ++j) scratch[(rq + j) * 16 + col] = o[0][j];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane < 16) {
    out3[0] = scratch[lane * 16 + 0];
    out3[1] = scratch[lane * 16 + 1];
    out3[2] = scratch[lane * 16 + 2];
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}


}  // namespace mpm
