// mpm_mlp.h -- parameter layout and launchers of the return-map surrogate MLP (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

constexpr int MLP_IN = 5, MLP_IN_PAD = 32, MLP_HIDDEN = 128, MLP_OUT = 3, MLP_OUT_PAD = 16, MLP_LAYERS = 5;
// input / output normalisation (the network sees x = (eps / EPS_SCALE, (phi - PHI_MID) / PHI_SPAN,
// (K/mu - KMU_MID) / KMU_SPAN) and predicts y = (H - eps) / CORR_SCALE)
constexpr float MLP_EPS_SCALE = 0.02f;
constexpr float MLP_PHI_MID = 32.5f, MLP_PHI_SPAN = 7.5f;
constexpr float MLP_KMU_MID = 2.5f, MLP_KMU_SPAN = 1.0f;
constexpr float MLP_CORR_SCALE = 1e-3f;
// Forward kernels: 12 waves of 64 rows share one LDS copy of the weights (121 KB: one workgroup per
// CU), 3 waves per SIMD; each wave takes its 64 rows through the network MLP_FWD_TILES 16-row tiles at
// a time, which keeps it under the 168 VGPRs that 3 waves per SIMD allow (4 tiles at once need about
// 235, 2 waves per SIMD). Measured on the surrogate kernel at config 4 (3.05M candidates): 8 waves
// of 4 tiles 592 us, 12 waves of 2 tiles 584 us, 16 waves of 2 tiles (128 VGPRs, spilling) 654 us.
constexpr int MLP_FWD_THREADS = 768;
constexpr int MLP_FWD_TILES = 2;
// The forward pass works on pre-activations scaled by log2(e): the forward copy of the first layer's
// weights and the staged biases of layers 1-4 are multiplied by log2(e), so the MFMA yields
// z' = log2(e) z and SiLU is z' / (1 + 2^-z') = log2(e) silu(z) with no multiply before the exp2;
// the next layer's unscaled weights then give log2(e) times its pre-activation again, and the
// forward copy of the output layer's weights carries the ln 2 that undoes the factor. The backward
// pass works unscaled (its transposed weights are not scaled; the saved pre-activations are z).
constexpr float MLP_LOG2E = 1.44269504088896341f, MLP_LN2 = 0.693147180559945309f;

// Master (FP32) layout: for l = 0..4, W_l [N_l][K_l] then b_l [N_l];
// N = (128,128,128,128,16), K = (32,128,128,128,128).
__host__ __device__ constexpr int mlp_N(int l) { return l == 4 ? MLP_OUT_PAD : MLP_HIDDEN; }
__host__ __device__ constexpr int mlp_K(int l) { return l == 0 ? MLP_IN_PAD : MLP_HIDDEN; }
__host__ __device__ constexpr int mlp_weight_offset(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += mlp_N(i) * mlp_K(i) + mlp_N(i);
  return o;
}
__host__ __device__ constexpr int mlp_bias_offset(int l) { return mlp_weight_offset(l) + mlp_N(l) * mlp_K(l); }
constexpr int MLP_NPARAM = mlp_weight_offset(5);

// BF16 forward copy: the W blocks only, [N][K] per layer, concatenated.
constexpr int MLP_BF_W1 = 0;
constexpr int MLP_BF_WH = MLP_HIDDEN * MLP_IN_PAD;
constexpr int MLP_BF_W5 = MLP_BF_WH + 3 * MLP_HIDDEN * MLP_HIDDEN;
constexpr int MLP_BF_SIZE = MLP_BF_W5 + MLP_OUT_PAD * MLP_HIDDEN;
// BF16 transposed copy for the backward pass: W_l^T [K_l][N'_l] for l = 1..4 (N'_4 = 32, zero padded).
__host__ __device__ constexpr int mlp_t_offset(int l) { return (l - 1) * MLP_HIDDEN * MLP_HIDDEN; }
constexpr int MLP_T_SIZE = 3 * MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN * 32;

// Is master entry i a real parameter (not padding)?
__host__ __device__ inline bool mlp_param_active(int i) {
  for (int l = 0; l < 5; ++l) {
    const int w0 = mlp_weight_offset(l), b0 = mlp_bias_offset(l), e = b0 + mlp_N(l);
    if (i >= w0 && i < b0) {
      const int o = (i - w0) / mlp_K(l), k = (i - w0) % mlp_K(l);
      if (l == 0) return k < MLP_IN;
      if (l == 4) return o < MLP_OUT;
      return true;
    }
    if (i >= b0 && i < e) return l == 4 ? (i - b0) < MLP_OUT : true;
  }
  return false;
}

size_t mlp_forward_smem();
size_t mlp_backward_smem();
void mlp_set_smem_limits();
void launch_mlp_forward(hipStream_t s, const uint16_t* wbf, const float* master, const float* in, float* out, int n,
                        float* zsave, int grid);
void launch_mlp_backward(hipStream_t s, const uint16_t* wbft, const float* in, const float* y, const float* target,
                         const float* zsave, int n, float* grad_part, float* loss, int grid);
void launch_mlp_reduce(hipStream_t s, const float* part, int nparts, float* grad, const float* loss_part,
                       float* loss_acc);
void launch_mlp_adam(hipStream_t s, float* master, float* m1, float* m2, const float* grad, uint16_t* wbf,
                     uint16_t* wbft, float lr, int step);
void launch_mlp_to_bf16(hipStream_t s, const float* master, uint16_t* wbf, uint16_t* wbft);
void launch_mlp_targets(hipStream_t s, const float* in, const float* H, int n, float* tgt);
void launch_mlp_loss_mean(hipStream_t s, const float* loss_part, int nparts, int n, float* loss);

}  // namespace mpm
