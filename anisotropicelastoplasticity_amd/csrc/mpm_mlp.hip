// mpm_mlp.hip -- the learned Drucker-Prager return-map surrogate: a 5 -> 128 -> 128 -> 128 -> 128 -> 3
// MLP with SiLU, BF16 weights and activations and FP32 accumulation, evaluated and trained with
// v_mfma_f32_16x16x32_bf16 (CDNA4).
//
// Network input (one row per particle): the trial Hencky principal strains eps_1 >= eps_2 >= eps_3,
// the friction angle phi (degrees) and the ratio of bulk to shear modulus K/mu, normalised as
//   x = (eps / EPS_SCALE, (phi - PHI_MID) / PHI_SPAN, (K/mu - KMU_MID) / KMU_SPAN).
// Output: the plastic correction y = (H - eps) / CORR_SCALE of the projected strains H.
// The exact return map (HybridSolver.cpp:646-673) produces the training pairs and remains the
// fallback whenever the prediction fails the yield-residual gate.
//
// Layout. Layer l has weights W_l stored [N_l][K_l] (out x in, row-major) and bias b_l; the padded
// shapes are N = (128,128,128,128,16), K = (32,128,128,128,128): input columns 5..31 and output
// rows 3..15 are zero and stay zero (their gradients are masked). FP32 master weights live in one
// flat vector; the BF16 copies are regenerated after every optimiser step.
//
// MFMA lane layout (16x16x32 bf16): lane i holds A[i%16][8*(i/16) .. +8] and B[8*(i/16) .. +8][i%16];
// the 16x16 FP32 result gives lane i the rows 4*(i/16) .. +4 of column i%16.
#include <hip/hip_runtime.h>

#include "mpm_mlp_device.hpp"

namespace mpm {

// ------------------------------------------------------------------ batch forward (inference / training)
// in: n x 5 raw inputs (row-major), out: n x 3 normalised outputs; zsave (optional): per 16-row block b,
// 4 x 16 x 128 FP32 pre-activations at zsave + b * 4 * 16 * 128. Each wave evaluates 64 rows.
__global__ __launch_bounds__(MLP_FWD_THREADS) void k_mlp_forward(const uint16_t* __restrict__ wbf,
                                                                 const float* __restrict__ master,
                                                                 const float* __restrict__ in, float* __restrict__ out,
                                                                 int n, float* __restrict__ zsave) {
  extern __shared__ __align__(16) unsigned char smem[];
  MLPShared& s = *reinterpret_cast<MLPShared*>(smem);
  stage_weights(s, wbf, master);
  __syncthreads();
  const int nw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int g = blockIdx.x * nw + w; (int64_t)g * 64 < n; g += gridDim.x * nw) {
    const int64_t r = (int64_t)g * 64 + lane;
    float x[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (r < n)
      for (int c = 0; c < 5; ++c) x[c] = in[r * 5 + c];
    float y[3];
    wave_forward64<MLP_FWD_TILES>(s, x, y, zsave, (int64_t)g * 64, n);
    if (r < n)
      for (int c = 0; c < 3; ++c) out[r * 3 + c] = y[c];
  }
}

// ------------------------------------------------------------------ backward
// One workgroup processes 64 consecutive rows (4 waves x 16) per iteration and accumulates the
// weight gradients of its rows; the loss is L = sum over rows and outputs of (y - t)^2 / n.
// Partial gradients are written per workgroup (grad_part + blockIdx.x * MLP_NPARAM) and summed in
// workgroup order by k_mlp_reduce, so training is deterministic.
struct BackShared {
  float dz[64][MLP_HIDDEN + 4];      // dZ of the current layer, rows of the 64-row group (FP32)
  uint16_t dzt[MLP_HIDDEN][64 + 8];  // dZ^T (bf16), feature-major
  uint16_t at[MLP_HIDDEN][64 + 8];   // A_prev^T (bf16), feature-major
  uint16_t dzb[64][LDA];             // dZ rows (bf16): the A operand of dA = dZ W
};

__global__ __launch_bounds__(256) void k_mlp_backward(const uint16_t* __restrict__ wbft, const float* __restrict__ in,
                                                      const float* __restrict__ y, const float* __restrict__ target,
                                                      const float* __restrict__ zsave, int n,
                                                      float* __restrict__ grad_part, float* __restrict__ loss) {
  extern __shared__ __align__(16) unsigned char smem[];
  BackShared& s = *reinterpret_cast<BackShared*>(smem);
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int col = lane & 15, rq = (lane >> 4) * 4;
  float* g = grad_part + (size_t)blockIdx.x * MLP_NPARAM;
  for (int i = tid; i < MLP_NPARAM; i += 256) g[i] = 0.0f;
  __syncthreads();
  float lsum = 0.0f;
  const int ngrp = (n + 63) / 64;
  for (int grp = blockIdx.x; grp < ngrp; grp += gridDim.x) {
    const int row0 = grp * 64;
    const int blk0 = grp * 4;  // 16-row blocks of zsave
    // dY for layer 5: 2 (y - t) / n, in the dz array (first 16 columns, zero padded)
    for (int i = tid; i < 64 * (MLP_HIDDEN + 4); i += 256) (&s.dz[0][0])[i] = 0.0f;
    __syncthreads();
    if (tid < 64) {
      const int r = row0 + tid;
      if (r < n)
        for (int c = 0; c < 3; ++c) {
          const float d = y[(size_t)r * 3 + c] - target[(size_t)r * 3 + c];
          lsum += d * d;
          s.dz[tid][c] = 2.0f * d / (float)n;
        }
    }
    __syncthreads();
    // walk the layers from the output down: layer index L = 4 (output) .. 0 (input layer)
    for (int L = 4; L >= 0; --L) {
      const int nout = (L == 4) ? MLP_OUT_PAD : MLP_HIDDEN;
      const int nin = (L == 0) ? MLP_IN_PAD : MLP_HIDDEN;
      // A_prev^T (feature-major, bf16): SiLU(Z_{L-1}) or the normalised inputs for L = 0
      for (int i = tid; i < nin * 64; i += 256) {
        const int f = i / 64, r = i % 64;
        float a = 0.0f;
        if (row0 + r < n) {
          if (L == 0) {
            if (f < 5) {
              float x = in[(size_t)(row0 + r) * 5 + f];
              a = f < 3 ? x * (1.0f / MLP_EPS_SCALE)
                        : (f == 3 ? (x - MLP_PHI_MID) * (1.0f / MLP_PHI_SPAN) : (x - MLP_KMU_MID) * (1.0f / MLP_KMU_SPAN));
            }
          } else {
            const int b = blk0 + r / 16;
            a = silu(zsave[((size_t)b * 4 + (L - 1)) * 16 * MLP_HIDDEN + (size_t)(r % 16) * MLP_HIDDEN + f]);
          }
        }
        s.at[f][r] = f2bf(a);
      }
      // dZ^T (bf16) and dZ rows (bf16)
      for (int i = tid; i < MLP_HIDDEN * 64; i += 256) {
        const int f = i / 64, r = i % 64;
        const float d = f < nout ? s.dz[r][f] : 0.0f;
        s.dzt[f][r] = f2bf(d);
        s.dzb[r][f] = f2bf(d);
      }
      __syncthreads();
      // dW_L[o][k] += sum_r dZ[r][o] A_prev[r][k]: wave w takes output rows o in [32w, 32w+32)
      // (L = 4: only wave 0, 16 rows); K = 64 rows = 2 MFMA steps.
      {
        const int mrows = (L == 4) ? 16 : 32;
        if (16 * w < ((L == 4) ? 16 : 128) && (L != 4 || w == 0)) {
          for (int mt = 0; mt < mrows / 16; ++mt) {
            const int o0 = (L == 4 ? 0 : 32 * w) + 16 * mt;
            const int ntile = nin / 16;
            for (int t = 0; t < ntile; ++t) {
              f32x4 acc = f32x4{0, 0, 0, 0};
              const int r = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
              for (int sk = 0; sk < 2; ++sk) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&s.dzt[o0 + r][32 * sk + kq]);
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(&s.at[16 * t + r][32 * sk + kq]);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
              }
              // lane holds dW[o0 + rq + j][16 t + col]
              const int woff = mlp_weight_offset(L);
              for (int j = 0; j < 4; ++j) {
                const int o = o0 + rq + j, k = 16 * t + col;
                if (o < nout && k < nin) g[woff + o * nin + k] += acc[j];
              }
            }
          }
        }
      }
      // db_L = sum_r dZ[r][o]
      for (int o = tid; o < nout; o += 256) {
        float sb = 0.0f;
        for (int r = 0; r < 64; ++r) sb += s.dz[r][o];
        g[mlp_bias_offset(L) + o] += sb;
      }
      __syncthreads();
      if (L == 0) break;
      // dA_{L-1} = dZ_L W_L (rows x 128), then dZ_{L-1} = dA (.) SiLU'(Z_{L-1}); wave w: rows 16w..16w+15
      {
        f32x4 acc[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) acc[t] = f32x4{0, 0, 0, 0};
        // B operand [N = in feature][K = out feature] is W_L^T, read from the transposed BF16 copy
        if (L == 4) wave_gemm_nt<8, 1>(&s.dzb[16 * w][0], LDA, wbft + mlp_t_offset(4), 32, acc);
        else wave_gemm_nt<8, 4>(&s.dzb[16 * w][0], LDA, wbft + mlp_t_offset(L), MLP_HIDDEN, acc);
        __syncthreads();
        const int b = blk0 + w;
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int row = rq + j, f = 16 * t + col;
            float d = 0.0f;
            if (row0 + 16 * w + row < n)
              d = acc[t][j] * dsilu(zsave[((size_t)b * 4 + (L - 1)) * 16 * MLP_HIDDEN + (size_t)row * MLP_HIDDEN + f]);
            s.dz[16 * w + row][f] = d;
          }
        __syncthreads();
      }
    }
    __syncthreads();
  }
  // wave-reduce the loss and add it once per workgroup
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
  __shared__ float s_l[4];
  if (lane == 0) s_l[w] = lsum;
  __syncthreads();
  if (tid == 0) loss[blockIdx.x] = s_l[0] + s_l[1] + s_l[2] + s_l[3];
}

// Sum of the per-workgroup partial gradients in workgroup order (deterministic).
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, int nparts, float* __restrict__ grad,
                                                    const float* __restrict__ loss_part, float* __restrict__ loss_acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && loss_acc) {
    float l = 0.0f;
    for (int p = 0; p < nparts; ++p) l += loss_part[p];
    *loss_acc += l;
  }
  if (i >= MLP_NPARAM) return;
  float s = 0.0f;
  for (int p = 0; p < nparts; ++p) s += part[(size_t)p * MLP_NPARAM + i];
  grad[i] = s;
}

// BF16 copies of master entry i: forward [N][K] and, for layers 1..4, transposed [K][N'].
__device__ __forceinline__ void mlp_store_bf16(int i, float wv, uint16_t* wbf, uint16_t* wbft) {
  for (int l = 0; l < 5; ++l) {
    const int w0 = mlp_weight_offset(l), b0 = mlp_bias_offset(l);
    if (i >= w0 && i < b0) {
      const int K = mlp_K(l), o = (i - w0) / K, k = (i - w0) % K;
      const int fo = l == 0 ? MLP_BF_W1 : (l == 4 ? MLP_BF_W5 : MLP_BF_WH + (l - 1) * MLP_HIDDEN * MLP_HIDDEN);
      // forward copy in the log2(e)-scaled convention (MLP_LOG2E); the backward's transposed copy unscaled
      wbf[fo + o * K + k] = f2bf(l == 0 ? wv * MLP_LOG2E : (l == 4 ? wv * MLP_LN2 : wv));
      if (l >= 1) wbft[mlp_t_offset(l) + k * (l == 4 ? 32 : MLP_HIDDEN) + o] = f2bf(wv);
      return;
    }
  }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; padded entries stay zero; BF16 copies refreshed.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ master, float* __restrict__ m1, float* __restrict__ m2,
                                                  const float* __restrict__ grad, uint16_t* __restrict__ wbf,
                                                  uint16_t* __restrict__ wbft, float lr, float beta1, float beta2, float eps,
                                                  int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  const bool act = mlp_param_active(i);
  const float gi = act ? grad[i] : 0.0f;
  const float a = beta1 * m1[i] + (1.0f - beta1) * gi;
  const float b = beta2 * m2[i] + (1.0f - beta2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float ah = a / (1.0f - powf(beta1, (float)step));
  const float bh = b / (1.0f - powf(beta2, (float)step));
  float wv = act ? master[i] - lr * ah / (sqrtf(bh) + eps) : 0.0f;
  master[i] = wv;
  mlp_store_bf16(i, wv, wbf, wbft);
}

__global__ __launch_bounds__(256) void k_mlp_to_bf16(const float* __restrict__ master, uint16_t* __restrict__ wbf,
                                                     uint16_t* __restrict__ wbft) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  mlp_store_bf16(i, mlp_param_active(i) ? master[i] : 0.0f, wbf, wbft);
}

// Normalised targets t = (H - eps) / CORR_SCALE of n rows (in: n x 5, H: n x 3).
__global__ __launch_bounds__(256) void k_mlp_targets(const float* __restrict__ in, const float* __restrict__ H, int n,
                                                     float* __restrict__ tgt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * n) return;
  const int r = i / 3, c = i % 3;
  tgt[i] = (H[i] - in[(size_t)r * 5 + c]) / MLP_CORR_SCALE;
}

// Mean loss: the per-workgroup losses summed in workgroup order, over n.
__global__ void k_mlp_loss_mean(const float* __restrict__ loss_part, int nparts, int n, float* __restrict__ loss) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  float l = 0.0f;
  for (int p = 0; p < nparts; ++p) l += loss_part[p];
  *loss = l / (float)n;
}

size_t mlp_forward_smem() { return sizeof(MLPShared); }
size_t mlp_backward_smem() { return sizeof(BackShared); }

void mlp_set_smem_limits() {
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mlp_forward), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)mlp_forward_smem());
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mlp_backward), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)mlp_backward_smem());
}

void launch_mlp_forward(hipStream_t s, const uint16_t* wbf, const float* master, const float* in, float* out, int n,
                        float* zsave, int grid) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(MLP_FWD_THREADS), mlp_forward_smem(), s, wbf, master, in, out, n,
                     zsave);
}

void launch_mlp_backward(hipStream_t s, const uint16_t* wbft, const float* in, const float* y, const float* target,
                         const float* zsave, int n, float* grad_part, float* loss, int grid) {
  hipLaunchKernelGGL(k_mlp_backward, dim3(grid), dim3(256), mlp_backward_smem(), s, wbft, in, y, target, zsave, n,
                     grad_part, loss);
}

void launch_mlp_reduce(hipStream_t s, const float* part, int nparts, float* grad, const float* loss_part,
                       float* loss_acc) {
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, part, nparts, grad, loss_part,
                     loss_acc);
}

void launch_mlp_targets(hipStream_t s, const float* in, const float* H, int n, float* tgt) {
  hipLaunchKernelGGL(k_mlp_targets, dim3((3 * n + 255) / 256), dim3(256), 0, s, in, H, n, tgt);
}

void launch_mlp_loss_mean(hipStream_t s, const float* loss_part, int nparts, int n, float* loss) {
  hipLaunchKernelGGL(k_mlp_loss_mean, dim3(1), dim3(64), 0, s, loss_part, nparts, n, loss);
}

void launch_mlp_adam(hipStream_t s, float* master, float* m1, float* m2, const float* grad, uint16_t* wbf,
                     uint16_t* wbft, float lr, int step) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, master, m1, m2, grad, wbf, wbft, lr,
                     0.9f, 0.999f, 1e-8f, step);
}

void launch_mlp_to_bf16(hipStream_t s, const float* master, uint16_t* wbf, uint16_t* wbft) {
  hipLaunchKernelGGL(k_mlp_to_bf16, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, master, wbf, wbft);
}

}  // namespace mpm
