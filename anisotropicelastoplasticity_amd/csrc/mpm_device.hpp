// mpm_device.hpp -- device math for the gfx950 MPM kernels: B-spline weights, 3x3 SVD (FP32 with an
// FP64 redo for near-degenerate input), constitutive stress, return mapping and level sets.
//
// Written for CDNA4 wave64 execution: everything is branch-light scalar code that lives in VGPRs;
// no shared state. Reference behaviour is cited per function.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpm_c_abi.h"

namespace mpm {

// ------------------------------------------------------------------ parameters on the device
struct DevParams {
  int order;             // 2 or 3
  int material;          // MPM_SNOW / MPM_SAND
  int predicted;         // stress at F^ = (I + dt grad v) F (reference) or at F^n
  int friction_model;
  float gravity[3];
  float lambda0, mu0;    // Lame parameters (HybridSolver.cpp:264-265)
  float theta_c, theta_s, xi;
  float h0, h1, h2, h3;  // DP hardening, degrees
  int per_particle_phi;
  // alpha = sqrt(2/3) 2 sin(phi) / (3 - sin(phi)) when phi is the same for every particle and every q
  // (h1 = h3 = 0 and no per-particle angle), evaluated once in double on the host; -1 otherwise
  float alpha_f;
  double alpha_d;
  float friction;
  float damping;
  float svd_ratio;
  float apex_radius;
  float dinv;            // D^-1 = 3/h^2 (cubic) or 4/h^2 (quadratic), h = min spacing (HybridSolver.cpp:175-177)
  float gmin[3], h[3], inv_h[3];
  int res[3];
  int tiles[3];          // ceil(res/4)
  uint32_t div_m[2];     // multiplier and shift of an exact integer division by res[0], res[1]
  int div_s[2];          // (floor_div: cell keys to rows and planes)
  int nls;
  mpm_levelset ls[16];
  // cloth: caller ids [mesh_v0, mesh_e0) are mesh vertices, [mesh_e0, mesh_e1) mesh elements
  int mesh_v0, mesh_e0, mesh_e1;
  float c_mu, c_lambda, c_gamma, c_k, c_cf;  // in-plane Lame parameters, shear and normal stiffness, tan(friction)
};

enum ParticleKind : int { KIND_PARTICLE = 0, KIND_VERTEX = 1, KIND_ELEMENT = 2 };
__device__ __forceinline__ int particle_kind(const DevParams& P, int id) {
  return id < P.mesh_v0 ? KIND_PARTICLE : (id < P.mesh_e0 ? KIND_VERTEX : KIND_ELEMENT);
}

// ------------------------------------------------------------------ B-splines
// Cubic (interpolation.cpp:9-33) and quadratic B-spline value and derivative at r (in cells).
template <int ORDER, class T>
__device__ __forceinline__ void spline(T r, T& w, T& dw) {
  T ar = r < T(0) ? -r : r;
  T sg = r < T(0) ? T(-1) : T(1);
  if constexpr (ORDER == 3) {
    if (ar < T(1)) {
      w = T(0.5) * ar * ar * ar - ar * ar + T(2) / T(3);
      dw = sg * (T(1.5) * ar * ar - T(2) * ar);
    } else if (ar < T(2)) {
      T t = T(2) - ar;
      w = t * t * t / T(6);
      dw = -sg * T(0.5) * t * t;
    } else {
      w = T(0);
      dw = T(0);
    }
  } else {
    if (ar < T(0.5)) {
      w = T(0.75) - ar * ar;
      dw = T(-2) * r;
    } else if (ar < T(1.5)) {
      T t = T(1.5) - ar;
      w = T(0.5) * t * t;
      dw = -sg * t;
    } else {
      w = T(0);
      dw = T(0);
    }
  }
}

// The three quadratic B-spline weights and derivatives (per cell) of a particle at fx = g - base in
// [0.5, 1.5), base = floor(g - 0.5): N(fx), N(fx - 1), N(fx - 2) in closed form, without branches.
__device__ __forceinline__ void quad_weights(float fx, float* w, float* dw) {
  const float a = 1.5f - fx, b = fx - 1.0f, c = fx - 0.5f;
  w[0] = 0.5f * a * a;
  w[1] = 0.75f - b * b;
  w[2] = 0.5f * c * c;
  dw[0] = -a;
  dw[1] = -2.0f * b;
  dw[2] = c;
}

// ------------------------------------------------------------------ 3x3 helpers (row-major)
template <class T>
__device__ __forceinline__ void mat_mul(const T* A, const T* B, T* C) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) C[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
}
template <class T>
__device__ __forceinline__ T mat_det(const T* A) {
#pragma clang fp contract(off)
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// Reciprocal, square root and reciprocal square root for the Jacobi SVD: the hardware approximations
// (1 ulp) in FP32, exact operations in FP64. Any rotation angle is valid for the iteration; only
// c^2 + s^2 = 1 matters, which holds to 2 ulp with c = rsqrt(1 + t^2), s = c t. The singular values
// (column norms) and the normalised columns of U carry 1-2 ulp as well.
template <class T>
__device__ __forceinline__ T jac_rcp(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
  else return T(1) / x;
}
template <class T>
__device__ __forceinline__ T jac_sqrt(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_sqrtf(x);
  else return sqrt(x);
}
template <class T>
__device__ __forceinline__ T jac_rsqrt(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rsqf(x);
  else return T(1) / sqrt(x);
}

// One-sided (Hestenes) Jacobi SVD of a 3x3 matrix: A = U diag(S) V^T with S >= 0 in decreasing
// order (the convention of Eigen's JacobiSVD used at HybridSolver.cpp:308,620). Columns of W = A V
// are orthogonalised by plane rotations until every pair is orthogonal to working precision or
// max_sweeps is reached; S = column norms, U = normalised columns. Returns false if not converged.
template <class T>
__device__ __forceinline__ bool svd3(const T* A, T* U, T* S, T* V, int max_sweeps) {
  T w[9], v[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) { w[i] = A[i]; v[i] = T(0); }
  v[0] = v[4] = v[8] = T(1);
  const T eps = sizeof(T) == 4 ? T(1.5e-7) : T(4e-16);
  bool converged = false;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    bool rotated = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int p = k == 2 ? 1 : 0;
      const int q = k == 0 ? 1 : 2;
      T alpha = w[p] * w[p] + w[3 + p] * w[3 + p] + w[6 + p] * w[6 + p];
      T beta = w[q] * w[q] + w[3 + q] * w[3 + q] + w[6 + q] * w[6 + q];
      T gamma = w[p] * w[q] + w[3 + p] * w[3 + q] + w[6 + p] * w[6 + q];
      if (gamma * gamma > eps * eps * (alpha * beta)) {  // |gamma| > eps sqrt(alpha beta)
        rotated = true;
        T zeta = (beta - alpha) * jac_rcp(T(2) * gamma);
        T az = zeta < T(0) ? -zeta : zeta;
        T t = (zeta >= T(0) ? T(1) : T(-1)) * jac_rcp(az + jac_sqrt(T(1) + zeta * zeta));
        T c = jac_rsqrt(T(1) + t * t);
        T s = c * t;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          T wp = w[3 * r + p], wq = w[3 * r + q];
          w[3 * r + p] = c * wp - s * wq;
          w[3 * r + q] = s * wp + c * wq;
          T vp = v[3 * r + p], vq = v[3 * r + q];
          v[3 * r + p] = c * vp - s * vq;
          v[3 * r + q] = s * vp + c * vq;
        }
      }
    }
    if (!rotated) { converged = true; break; }
  }
  T s[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) s[c] = jac_sqrt(w[c] * w[c] + w[3 + c] * w[3 + c] + w[6 + c] * w[6 + c]);
  // sort columns by decreasing singular value
  int o0 = 0, o1 = 1, o2 = 2;
  if (s[o0] < s[o1]) { int t = o0; o0 = o1; o1 = t; }
  if (s[o1] < s[o2]) { int t = o1; o1 = o2; o2 = t; }
  if (s[o0] < s[o1]) { int t = o0; o0 = o1; o1 = t; }
  const int ord[3] = {o0, o1, o2};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    int oc = ord[c];
    T sc = oc == 0 ? s[0] : (oc == 1 ? s[1] : s[2]);
    S[c] = sc;
    T inv = sc > T(0) ? jac_rcp(sc) : T(0);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      T wv = oc == 0 ? w[3 * r] : (oc == 1 ? w[3 * r + 1] : w[3 * r + 2]);
      T vv = oc == 0 ? v[3 * r] : (oc == 1 ? v[3 * r + 1] : v[3 * r + 2]);
      U[3 * r + c] = wv * inv;
      V[3 * r + c] = vv;
    }
  }
  if (!(S[2] > T(0))) {  // complete U with u0 x u1
    U[2] = U[3] * U[7] - U[6] * U[4];
    U[5] = U[6] * U[1] - U[0] * U[7];
    U[8] = U[0] * U[4] - U[3] * U[1];
  }
  return converged;
}

// FP32 SVD with an FP64 redo when the input is near-degenerate (sigma_min < ratio * sigma_max),
// not converged or not finite. Counts redos in *fallback (per thread, summed by the caller).
__device__ __forceinline__ void svd3_guarded(const float* A, float* U, float* S, float* V, float ratio, int& fallback) {
  bool ok = svd3<float>(A, U, S, V, 8);
  bool finite = isfinite(S[0]) && isfinite(S[2]);
  if (!ok || !finite || S[2] < ratio * S[0]) {
    double Ad[9], Ud[9], Sd[3], Vd[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Ad[i] = A[i];
    svd3<double>(Ad, Ud, Sd, Vd, 30);
#pragma unroll
    for (int i = 0; i < 9; ++i) { U[i] = float(Ud[i]); V[i] = float(Vd[i]); }
    S[0] = float(Sd[0]); S[1] = float(Sd[1]); S[2] = float(Sd[2]);
    fallback += 1;
  }
}

// Stress matrix V_p * P(Fh) * G^T (HybridSolver.cpp:279-339) with P the first Piola-Kirchhoff
// stress of fixed-corotated snow with hardening (:281-287, 314-325) or of Hencky sand (:326-339).
__device__ __forceinline__ void stress_matrix(const DevParams& P, const float* Fh, const float* G, float Jp, float vol,
                                              float* out, int& fallback) {
  float U[9], S[3], V[9];
  svd3_guarded(Fh, U, S, V, P.svd_ratio, fallback);
  float Pk[9];
  if (P.material == MPM_SNOW) {
    float sc = expf(P.xi * (1.0f - Jp));
    float lambda = P.lambda0 * sc, mu = P.mu0 * sc;
    float J = mat_det(Fh);
    // J F^-T = cofactor matrix
    float C[9];
    C[0] = Fh[4] * Fh[8] - Fh[5] * Fh[7];
    C[1] = Fh[5] * Fh[6] - Fh[3] * Fh[8];
    C[2] = Fh[3] * Fh[7] - Fh[4] * Fh[6];
    C[3] = Fh[2] * Fh[7] - Fh[1] * Fh[8];
    C[4] = Fh[0] * Fh[8] - Fh[2] * Fh[6];
    C[5] = Fh[1] * Fh[6] - Fh[0] * Fh[7];
    C[6] = Fh[1] * Fh[5] - Fh[2] * Fh[4];
    C[7] = Fh[2] * Fh[3] - Fh[0] * Fh[5];
    C[8] = Fh[0] * Fh[4] - Fh[1] * Fh[3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float R = U[3 * r] * V[3 * c] + U[3 * r + 1] * V[3 * c + 1] + U[3 * r + 2] * V[3 * c + 2];
        Pk[3 * r + c] = 2.0f * mu * (Fh[3 * r + c] - R) + lambda * (J - 1.0f) * C[3 * r + c];
      }
  } else {
    float l0 = logf(S[0]), l1 = logf(S[1]), l2 = logf(S[2]);
    float tr = l0 + l1 + l2;
    float d0 = (2.0f * P.mu0 * l0 + P.lambda0 * tr) / S[0];
    float d1 = (2.0f * P.mu0 * l1 + P.lambda0 * tr) / S[1];
    float d2 = (2.0f * P.mu0 * l2 + P.lambda0 * tr) / S[2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) Pk[3 * r + c] = U[3 * r] * d0 * V[3 * c] + U[3 * r + 1] * d1 * V[3 * c + 1] + U[3 * r + 2] * d2 * V[3 * c + 2];
  }
  // out = vol * Pk * G^T
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      out[3 * r + c] = vol * (Pk[3 * r] * G[3 * c] + Pk[3 * r + 1] * G[3 * c + 1] + Pk[3 * r + 2] * G[3 * c + 2]);
}

// Drucker-Prager projection of principal stretches (HybridSolver.cpp:634-673). alpha_c >= 0 is the
// scene's constant alpha (phi independent of q and of the particle); otherwise alpha follows from
// phi(q) = phi0 + (h1 q - h3) exp(-h2 q). lnS receives ln Sin (the trial Hencky strains) and lnSout
// ln Sout: the trial strains (elastic), 0 (tip) or H = eps - dgamma dev(eps)/|dev(eps)| (cone).
template <class T>
__device__ __forceinline__ void dp_project(T lambda, T mu, T phi0, T h1, T h2, T h3, T q, T alpha_c, const T* Sin,
                                           T* Sout, T* lnS, T* lnSout, T& dq) {
#pragma clang fp contract(off)
  T alpha;
  if (alpha_c >= T(0)) {
    alpha = alpha_c;
  } else {
    const T pi = T(3.14159265358979323846);
    T fa = (phi0 + (h1 * q - h3) * exp(-h2 * q)) * pi / T(180);
    T sf = sin(fa);
    alpha = sqrt(T(2) / T(3)) * T(2) * sf / (T(3) - sf);
  }
  T l0 = log(Sin[0]), l1 = log(Sin[1]), l2 = log(Sin[2]);
  lnS[0] = l0; lnS[1] = l1; lnS[2] = l2;
  T tr = l0 + l1 + l2;
  T e0 = l0 - tr / T(3), e1 = l1 - tr / T(3), e2 = l2 - tr / T(3);
  T en = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
  T dg = en + (T(3) * lambda + T(2) * mu) / T(2) / mu * tr * alpha;
  dq = T(0);
  Sout[0] = Sin[0]; Sout[1] = Sin[1]; Sout[2] = Sin[2];
  lnSout[0] = l0; lnSout[1] = l1; lnSout[2] = l2;
  if (dg <= T(0)) {
  } else if (en == T(0) || tr > T(0)) {
    Sout[0] = Sout[1] = Sout[2] = T(1);
    lnSout[0] = lnSout[1] = lnSout[2] = T(0);
    dq = sqrt(l0 * l0 + l1 * l1 + l2 * l2);
  } else {
    T k = dg / en;
    lnSout[0] = l0 - k * e0;
    lnSout[1] = l1 - k * e1;
    lnSout[2] = l2 - k * e2;
    Sout[0] = exp(lnSout[0]);
    Sout[1] = exp(lnSout[1]);
    Sout[2] = exp(lnSout[2]);
    dq = dg;
  }
}

// K = U diag(d) U^T as its six entries (xx, yy, zz, xy, xz, yz).
__device__ __forceinline__ void kirchhoff_assemble(const float* U, const float* d, float* K) {
#pragma clang fp contract(off)
  const int R[6] = {0, 1, 2, 0, 0, 1}, C[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int e = 0; e < 6; ++e)
    K[e] = U[3 * R[e]] * d[0] * U[3 * C[e]] + U[3 * R[e] + 1] * d[1] * U[3 * C[e] + 1] + U[3 * R[e] + 2] * d[2] * U[3 * C[e] + 2];
}

// Sand: V tau = V U diag(2 mu ln s_i + lambda sum ln s) U^T from the Hencky strains ln s_i.
__device__ __forceinline__ void kirchhoff_sand_ln(const DevParams& P, const float* U, const float* l, float vol, float* K) {
#pragma clang fp contract(off)
  const float tr = P.lambda0 * (l[0] + l[1] + l[2]);
  const float d[3] = {vol * (2.0f * P.mu0 * l[0] + tr), vol * (2.0f * P.mu0 * l[1] + tr), vol * (2.0f * P.mu0 * l[2] + tr)};
  kirchhoff_assemble(U, d, K);
}

// Volume-weighted Kirchhoff stress K = V_p P(F) F^T of F = U diag(S) V^T, stored as the six entries
// (xx, yy, zz, xy, xz, yz) of a symmetric matrix. With P of HybridSolver.cpp:314-339:
//   sand: P F^T = U diag(2 mu ln s_i + lambda sum ln s) U^T
//   snow: P F^T = 2 mu (F F^T - R F^T) + lambda (J-1) J I = U diag(2 mu (s_i^2 - s_i) + lambda (J-1) J) U^T
// (R = U V^T, J = det F), with lambda, mu scaled by exp(xi (1 - J_P)) for snow (:281-287).
__device__ __forceinline__ void kirchhoff(const DevParams& P, const float* U, const float* S, float J, float Jp,
                                          float vol, float* K) {
#pragma clang fp contract(off)
  float d[3];
  if (P.material == MPM_SNOW) {
    float sc = expf(P.xi * (1.0f - Jp));
    float lambda = P.lambda0 * sc, mu = P.mu0 * sc;
    float iso = lambda * (J - 1.0f) * J;
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = vol * (2.0f * mu * (S[k] * S[k] - S[k]) + iso);
  } else {
    const float l[3] = {logf(S[0]), logf(S[1]), logf(S[2])};
    kirchhoff_sand_ln(P, U, l, vol, K);
    return;
  }
  kirchhoff_assemble(U, d, K);
}

// K of a given F (used once after the caller sets the state; afterwards G2P produces K).
__device__ __forceinline__ void kirchhoff_of_F(const DevParams& P, const float* F, float Jp, float vol, float* K, int& fb) {
  float U[9], S[3], V[9];
  svd3_guarded(F, U, S, V, P.svd_ratio, fb);
  kirchhoff(P, U, S, mat_det(F), Jp, vol, K);
}

// The sand return map after the SVD and the projection, as in return_map (used by the surrogate
// kernel's fallback so that a rejected prediction gives bitwise the exact result): F_E = U Sn V^T and
// V tau from the projected Hencky strains ln Sn.
__device__ __forceinline__ void sand_finish(const DevParams& P, const float* U, const float* V, const float* Sn,
                                            const float* lnSn, float dq, float& plastic, float vol, float* Fe, float* K) {
#pragma clang fp contract(off)
  plastic += dq;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Fe[3 * r + c] = U[3 * r] * Sn[0] * V[3 * c] + U[3 * r + 1] * Sn[1] * V[3 * c + 1] + U[3 * r + 2] * Sn[2] * V[3 * c + 2];
  kirchhoff_sand_ln(P, U, lnSn, vol, K);
}

// Surrogate hooks of the return map: pair logging (exact path) or deferral of cone-branch
// particles to the surrogate kernel.
struct SurrogateIO {
  int mode;          // 0 = exact, 1 = exact + log cone pairs, 2 = defer cone-branch particles
  float* pair_log;   // [cap][8]: eps(3), phi, K/mu, H(3)
  int64_t pair_cap;
  unsigned long long* pair_count;
  // filled when mode == 2 and the particle is deferred
  bool deferred;
  float sigma[3];
  float U[9];
  float V[9];
  // FP64 apex projections: deferred to a kernel of their own (defer_apex), which then calls
  // return_map again on the unchanged trial state with defer_apex off and gets the same bits
  bool defer_apex;
  bool apex_deferred;
};

// Friction angle (degrees) and K/mu of one particle: phi(q) = h0 + (h1 q - h3) exp(-h2 q)
// (HybridSolver.cpp:646-647), h0 replaced by the per-particle angle when configured.
__device__ __forceinline__ float dp_phi(const DevParams& P, float phi_p, float q) {
  return (P.per_particle_phi ? phi_p : P.h0) + (P.h1 * q - P.h3) * expf(-P.h2 * q);
}
__device__ __forceinline__ float bulk_over_shear(const DevParams& P) { return P.lambda0 / P.mu0 + 2.0f / 3.0f; }

// FP32 Drucker-Prager projection as called by the return map; the surrogate's exact fallback calls
// the same function on the same operands (no reassociation without fast-math), giving the same bits.
__device__ __forceinline__ void sand_project(const DevParams& P, float phi0, float q, const float* S, float* Sn,
                                             float* lnS, float* lnSn, float& dq) {
  dp_project<float>(P.lambda0, P.mu0, phi0, P.h1, P.h2, P.h3, q, P.alpha_f, S, Sn, lnS, lnSn, dq);
}

// Training pairs of the surrogate (mode 1): the trial Hencky strains l of a cone projection and its
// result ln Sn (the tip branch, Sigma = 1, and the tensile side are not logged).
__device__ __forceinline__ void sand_log_pair(const DevParams& P, SurrogateIO& io, const float* l, const float* Sn,
                                              const float* lnSn, float dq, float phi_p, float q) {
  if (io.mode != 1 || !(dq > 0.0f)) return;
  const float tr = l[0] + l[1] + l[2];
  if (tr <= 0.0f && Sn[0] != 1.0f) {
    const unsigned long long idx = atomicAdd(io.pair_count, 1ull);
    if ((int64_t)idx < io.pair_cap) {
      float* rec = io.pair_log + 8 * idx;
      rec[0] = l[0]; rec[1] = l[1]; rec[2] = l[2];
      rec[3] = dp_phi(P, phi_p, q);
      rec[4] = bulk_over_shear(P);
      rec[5] = lnSn[0]; rec[6] = lnSn[1]; rec[7] = lnSn[2];
    }
  }
}

// A sand particle whose projection lies within the apex radius: the FP64 projection from its trial
// SVD, then F_E and V tau (the apex kernel calls this with the SVD that k_g2p stored, return_map
// with its own, so both give the same bits). l = ln S in FP32 (for the pair log).
__device__ __forceinline__ void sand_apex_finish(const DevParams& P, const float* U, const float* S, const float* V,
                                                 const float* l, float phi0, float phi_p, float& plastic, float vol,
                                                 float* Fe, float* K, int& fb_apex, SurrogateIO& io) {
  double Sd[3] = {S[0], S[1], S[2]}, So[3], lSd[3], lSo[3], dqd;
  dp_project<double>(P.lambda0, P.mu0, phi0, P.h1, P.h2, P.h3, plastic, P.alpha_d, Sd, So, lSd, lSo, dqd);
  float Sn[3], lnSn[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Sn[k] = float(So[k]);
    lnSn[k] = float(lSo[k]);
  }
  const float dq = float(dqd);
  fb_apex += 1;
  sand_log_pair(P, io, l, Sn, lnSn, dq, phi_p, plastic);
  sand_finish(P, U, V, Sn, lnSn, dq, plastic, vol, Fe, K);
}

// Return mapping of one particle (HybridSolver.cpp:616-681). Fh is replaced by the elastic part
// U Sigma V^T; plastic holds J_P (snow, whose update det(F_P') = J_P * prod(Sigma)/prod(Sigma') is
// the determinant of F_P' = V Sigma'^-1 U^T F^ F_P, :676-677) or plasticAmount q (sand). K receives
// the Kirchhoff stress V tau of the new elastic state (same SVD).
__device__ __forceinline__ void return_map(const DevParams& P, float* Fh, float& plastic, float phi_p, float vol,
                                           float* K, int& fb_svd, int& fb_apex, SurrogateIO& io) {
  // No FMA contraction in the return map: the surrogate kernel's exact fallback evaluates the same
  // functions in another kernel and must reproduce these results bit for bit.
#pragma clang fp contract(off)
  float U[9], S[3], V[9];
  svd3_guarded(Fh, U, S, V, P.svd_ratio, fb_svd);
  if (io.mode == 2 && P.material == MPM_SAND) {
    // cone branch (trial plastic, compressive): leave F^ as is and hand the particle to the surrogate
    float l0 = logf(S[0]), l1 = logf(S[1]), l2 = logf(S[2]);
    float tr = l0 + l1 + l2;
    float e0 = l0 - tr / 3.0f, e1 = l1 - tr / 3.0f, e2 = l2 - tr / 3.0f;
    float en = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
    const float fa = dp_phi(P, phi_p, plastic) * 3.14159265358979323846f / 180.0f;
    const float sf = sinf(fa);
    const float alpha = sqrtf(2.0f / 3.0f) * 2.0f * sf / (3.0f - sf);
    const float dg = en + (3.0f * P.lambda0 + 2.0f * P.mu0) / (2.0f * P.mu0) * tr * alpha;
    if (dg > 0.0f && en > 0.0f && !(tr > 0.0f) && l0 * l0 + l1 * l1 + l2 * l2 >= P.apex_radius * P.apex_radius) {
      io.deferred = true;
      io.sigma[0] = S[0]; io.sigma[1] = S[1]; io.sigma[2] = S[2];
#pragma unroll
      for (int i = 0; i < 9; ++i) { io.U[i] = U[i]; io.V[i] = V[i]; }
      return;
    }
  }
  const float sgn = mat_det(U) * mat_det(V) < 0.0f ? -1.0f : 1.0f;
  float Sn[3];
  if (P.material == MPM_SNOW) {
    float lo = 1.0f - P.theta_c, hi = 1.0f + P.theta_s;
#pragma unroll
    for (int k = 0; k < 3; ++k) Sn[k] = S[k] > hi ? hi : (S[k] < lo ? lo : S[k]);
    plastic = plastic * ((S[0] / Sn[0]) * (S[1] / Sn[1]) * (S[2] / Sn[2]));
  } else {
    float phi0 = P.per_particle_phi ? phi_p : P.h0;
    float dq, l[3], lnSn[3];
    sand_project(P, phi0, plastic, S, Sn, l, lnSn, dq);
    // a projection onto the cone happened (the tip branch, tr > 0, gives Sigma = 1 exactly): close
    // to the cone apex it is redone in FP64
    if (dq > 0.0f && !(l[0] + l[1] + l[2] > 0.0f) &&
        l[0] * l[0] + l[1] * l[1] + l[2] * l[2] < P.apex_radius * P.apex_radius) {
      if (io.defer_apex) {  // F^, q and the stress are left as they are; the SVD goes to the apex kernel
        io.apex_deferred = true;
        io.sigma[0] = S[0]; io.sigma[1] = S[1]; io.sigma[2] = S[2];
#pragma unroll
        for (int i = 0; i < 9; ++i) { io.U[i] = U[i]; io.V[i] = V[i]; }
        return;
      }
      sand_apex_finish(P, U, S, V, l, phi0, phi_p, plastic, vol, Fh, K, fb_apex, io);
      return;
    }
    sand_log_pair(P, io, l, Sn, lnSn, dq, phi_p, plastic);
    sand_finish(P, U, V, Sn, lnSn, dq, plastic, vol, Fh, K);
    return;
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Fh[3 * r + c] = U[3 * r] * Sn[0] * V[3 * c] + U[3 * r + 1] * Sn[1] * V[3 * c + 1] + U[3 * r + 2] * Sn[2] * V[3 * c + 2];
  kirchhoff(P, U, Sn, sgn * Sn[0] * Sn[1] * Sn[2], plastic, vol, K);
}

// ------------------------------------------------------------------ cloth (LagrangianMesh)
// Gram-Schmidt QR of D = [d1 d2 d3] (columns; D row-major), GramSchmidtOrthonomalization
// (geometry.cpp:31-62): Q orthonormal columns, R upper triangular with R(i,i) = column norms.
// No FMA contraction: the rest factors and the per-substep factors are computed in different kernels
// and must agree bit for bit on an undeformed element.
template <class T>
__device__ __forceinline__ void qr3(const T* D, T* Q, T* R) {
#pragma clang fp contract(off)
  T q1[3], q2[3], q3[3];
  const T r11 = sqrt(D[0] * D[0] + D[3] * D[3] + D[6] * D[6]);
#pragma unroll
  for (int a = 0; a < 3; ++a) q1[a] = D[3 * a] / r11;
  const T r12 = D[1] * q1[0] + D[4] * q1[1] + D[7] * q1[2];
#pragma unroll
  for (int a = 0; a < 3; ++a) q2[a] = D[3 * a + 1] - r12 * q1[a];
  const T r22 = sqrt(q2[0] * q2[0] + q2[1] * q2[1] + q2[2] * q2[2]);
#pragma unroll
  for (int a = 0; a < 3; ++a) q2[a] /= r22;
  const T r13 = D[2] * q1[0] + D[5] * q1[1] + D[8] * q1[2];
  const T r23 = D[2] * q2[0] + D[5] * q2[1] + D[8] * q2[2];
#pragma unroll
  for (int a = 0; a < 3; ++a) q3[a] = D[3 * a + 2] - r13 * q1[a] - r23 * q2[a];
  const T r33 = sqrt(q3[0] * q3[0] + q3[1] * q3[1] + q3[2] * q3[2]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    q3[a] /= r33;
    Q[3 * a] = q1[a];
    Q[3 * a + 1] = q2[a];
    Q[3 * a + 2] = q3[a];
  }
  R[0] = r11; R[1] = r12; R[2] = r13;
  R[3] = T(0); R[4] = r22; R[5] = r23;
  R[6] = T(0); R[7] = T(0); R[8] = r33;
}

// FP32 QR, redone in FP64 when D is near-degenerate (a column nearly in the span of the previous
// ones: R(1,1) < ratio |d2| or R(2,2) < ratio |d3|) or the result is not finite.
__device__ __forceinline__ void qr3_guarded(const float* D, float* Q, float* R, float ratio, int& fallback) {
#pragma clang fp contract(off)
  qr3<float>(D, Q, R);
  const float n2 = sqrtf(D[1] * D[1] + D[4] * D[4] + D[7] * D[7]);
  const float n3 = sqrtf(D[2] * D[2] + D[5] * D[5] + D[8] * D[8]);
  const bool finite = isfinite(R[0]) && isfinite(R[4]) && isfinite(R[8]) && isfinite(R[5]);
  if (!finite || R[4] < ratio * n2 || R[8] < ratio * n3) {
    double Dd[9], Qd[9], Rd[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Dd[i] = D[i];
    qr3<double>(Dd, Qd, Rd);
#pragma unroll
    for (int i = 0; i < 9; ++i) { Q[i] = float(Qd[i]); R[i] = float(Rd[i]); }
    fallback += 1;
  }
}

// Per-element cloth forces from the current directions D (HybridSolver.cpp:370-455 and
// LagrangianMesh::computeVertexInPlaneForces, LagrangianMesh.cpp:382-460):
//  - in-plane: A = R_rest^-1 R (upper-left 2x2 blocks of the QR factors), P = 2 mu (A - rot(A)) +
//    lambda (J - 1) cof(A) with rot the orthogonal polar factor and J the product of A's diagonal;
//    vertex forces f2 = -(P11 i11 + P12 i12) q1, f3 = -P12 i22 q1 - P22 i22 q2, f1 = -(f2 + f3);
//  - normal/shear: dR = [[S11 S12 gamma r13], [0 S22 gamma r23], [0 0 -k (1 - r33)^2 or 0 if r33 > 1]]
//    with S = R_rest^-1 P, K = dR R^T, dF3 = Q sym_upper(K) R^-T cn, element stress V dF3 d3^T.
// rest = (i11, i12, i22, cn0, cn1, cn2, R11, R12, R22): the inverse rest in-plane R, the normal-force
// vector, and the rest in-plane R as this file's qr3 computes it from the element's rest directions.
// A is formed as I + E with E = R_rest^-1 (R - R_rest), and A - rot(A), J - 1 from E without
// cancellation. This equals the reference's A in exact arithmetic; in FP32 it makes an undeformed
// element exactly force-free (A = I bit for bit), where R_rest^-1 R carries a rounding strain of
// about 1e-7 that, for fine meshes, the unscaled in-plane force turns into accelerations of m/s^2.
__device__ __forceinline__ void cloth_element_forces(const DevParams& P, const float* D, const float* rest, float vol,
                                                     float* fv /* f1 f2 f3 */, float* M /* 3x3 */, int& fb) {
  float Q[9], R[9];
  qr3_guarded(D, Q, R, P.svd_ratio, fb);
  const float i11 = rest[0], i12 = rest[1], i22 = rest[2];
  const float d11 = R[0] - rest[6], d12 = R[1] - rest[7], d22 = R[4] - rest[8];
  const float e00 = i11 * d11, e01 = i11 * d12 + i12 * d22, e11 = i22 * d22;
  // A = [[1 + e00, e01], [0, 1 + e11]] has a non-negative diagonal (R and R_rest^-1 do), so det A >= 0
  // and rot(A) = [[c, -s], [s, c]] with (c, s) = (2 + e00 + e11, -e01) / norm.
  const float pp = 2.0f + e00 + e11, qq = -e01;
  const float nn = sqrtf(pp * pp + qq * qq);
  float s = 0.0f, omc = 0.0f;  // sin and 1 - cos of the polar rotation
  if (nn > 0.0f) {
    s = qq / nn;
    omc = qq * qq / (nn * (nn + pp));
  }
  const float m00 = e00 + omc, m01 = e01 + s, m10 = -s, m11 = e11 + omc;  // A - rot(A)
  const float Jm1 = e00 + e11 + e00 * e11;
  const float p00 = 2.0f * P.c_mu * m00 + P.c_lambda * Jm1 * (1.0f + e11);
  const float p01 = 2.0f * P.c_mu * m01;
  const float p10 = 2.0f * P.c_mu * m10 + P.c_lambda * Jm1 * (-e01);
  const float p11 = 2.0f * P.c_mu * m11 + P.c_lambda * Jm1 * (1.0f + e00);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float q1 = Q[3 * a], q2 = Q[3 * a + 1];
    const float f2 = -(p00 * i11 + p01 * i12) * q1;
    const float f3 = -p01 * i22 * q1 - p11 * i22 * q2;
    fv[3 + a] = f2;
    fv[6 + a] = f3;
    fv[a] = -(f2 + f3);
  }
  const float s11 = i11 * p00 + i12 * p10, s12 = i11 * p01 + i12 * p11, s22 = i22 * p11;
  const float dR[9] = {s11, s12, P.c_gamma * R[2], 0.0f, s22, P.c_gamma * R[5], 0.0f, 0.0f,
                       R[8] > 1.0f ? 0.0f : -P.c_k * (1.0f - R[8]) * (1.0f - R[8])};
  float K[9];  // dR R^T
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) K[3 * r + c] = dR[3 * r] * R[3 * c] + dR[3 * r + 1] * R[3 * c + 1] + dR[3 * r + 2] * R[3 * c + 2];
  float Ks[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Ks[3 * r + c] = r <= c ? K[3 * r + c] : K[3 * c + r];
  // R^-T cn: solve R^T y = cn (R^T lower triangular)
  float y[3];
  y[0] = rest[3] / R[0];
  y[1] = (rest[4] - R[1] * y[0]) / R[4];
  y[2] = (rest[5] - R[2] * y[0] - R[5] * y[1]) / R[8];
  float t[3], dF3[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) t[r] = Ks[3 * r] * y[0] + Ks[3 * r + 1] * y[1] + Ks[3 * r + 2] * y[2];
#pragma unroll
  for (int r = 0; r < 3; ++r) dF3[r] = Q[3 * r] * t[0] + Q[3 * r + 1] * t[1] + Q[3 * r + 2] * t[2];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) M[3 * r + c] = vol * dF3[r] * D[3 * c + 2];
}

// Cloth return mapping on the QR factor of D (HybridSolver.cpp:684-722): r33 > 1 (separation)
// clears the normal/shear part; otherwise the shear |(r13, r23)| is limited by the friction cone
// c_F k (r33 - 1)^2 / gamma. d3 <- Q R(:,2).
__device__ __forceinline__ void cloth_return_map(const DevParams& P, float* D, int& fb) {
  float Q[9], R[9];
  qr3_guarded(D, Q, R, P.svd_ratio, fb);
  if (R[8] > 1.0f) {
    R[8] = 1.0f;
    R[2] = R[5] = 0.0f;
  } else {
    const float fn = P.c_k * (R[8] - 1.0f) * (R[8] - 1.0f);
    const float ft = P.c_gamma * sqrtf(R[2] * R[2] + R[5] * R[5]);
    if (ft > P.c_cf * fn) {
      R[2] *= P.c_cf * fn / ft;
      R[5] *= P.c_cf * fn / ft;
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) D[3 * a + 2] = Q[3 * a] * R[2] + Q[3 * a + 1] * R[5] + Q[3 * a + 2] * R[8];
}

// ------------------------------------------------------------------ level sets (LevelSet.cpp:8-42)
__device__ __forceinline__ void levelset(const DevParams& P, const float* x, float& phi, float* n) {
  phi = 3.0e38f;
  n[0] = n[1] = n[2] = 0.0f;
  for (int k = 0; k < P.nls; ++k) {
    const mpm_levelset& pr = P.ls[k];
    float p, nn[3] = {0.0f, 0.0f, 0.0f};
    if (pr.type == MPM_LS_PLANE) {
      float len = sqrtf(pr.a[0] * pr.a[0] + pr.a[1] * pr.a[1] + pr.a[2] * pr.a[2]);
      nn[0] = pr.a[0] / len; nn[1] = pr.a[1] / len; nn[2] = pr.a[2] / len;
      p = nn[0] * x[0] + nn[1] * x[1] + nn[2] * x[2] - pr.a[3];
    } else if (pr.type == MPM_LS_WALL2GROUND) {
      float wx = pr.a[0], wy = pr.a[1], gz = pr.a[2];
      p = fminf(fminf(x[2] - gz, wx - x[0]), wy - x[1]);
      float dz = fabsf(x[2] - gz), dx = fabsf(wx - x[0]), dy = fabsf(wy - x[1]);
      if (dz <= dx && dz <= dy) nn[2] = 1.0f;
      else if (dy <= dx) nn[1] = -1.0f;
      else nn[0] = -1.0f;
    } else if (pr.type == MPM_LS_BOX) {
      p = 3.0e38f;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        float dlo = x[ax] - pr.a[ax], dhi = pr.b[ax] - x[ax];
        if (dlo < p) { p = dlo; nn[0] = nn[1] = nn[2] = 0.0f; nn[ax] = 1.0f; }
        if (dhi < p) { p = dhi; nn[0] = nn[1] = nn[2] = 0.0f; nn[ax] = -1.0f; }
      }
    } else {
      float d0 = x[0] - pr.a[0], d1 = x[1] - pr.a[1], d2 = x[2] - pr.a[2];
      float len = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
      p = len - pr.a[3];
      if (len > 0.0f) { nn[0] = d0 / len; nn[1] = d1 / len; nn[2] = d2 / len; }
      else nn[1] = 1.0f;
    }
    if (p < phi) { phi = p; n[0] = nn[0]; n[1] = nn[1]; n[2] = nn[2]; }
  }
}

// Grid collision of one node (HybridSolver.cpp:478-507); v in/out, vbf = velocity before friction.
__device__ __forceinline__ void collide(const DevParams& P, const float* xnode, float* v, float* vbf) {
  vbf[0] = v[0]; vbf[1] = v[1]; vbf[2] = v[2];
  if (P.nls == 0) return;
  float phi, n[3];
  levelset(P, xnode, phi, n);
  if (!(phi <= 0.0f)) return;
  if (P.friction_model == MPM_FRICTION_STICKY) {
    v[0] = v[1] = v[2] = 0.0f;
    vbf[0] = vbf[1] = vbf[2] = 0.0f;
    return;
  }
  float vn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
  if (!(vn < 0.0f)) return;
  float vt[3] = {v[0] - vn * n[0], v[1] - vn * n[1], v[2] - vn * n[2]};
  float vtn = sqrtf(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
  vbf[0] = vt[0]; vbf[1] = vt[1]; vbf[2] = vt[2];
  float mu = P.friction;
  if (P.friction_model == MPM_FRICTION_REFERENCE) {
    if (vtn < -mu * vn) { v[0] = v[1] = v[2] = 0.0f; }
    else { v[0] = vt[0]; v[1] = vt[1]; v[2] = vt[2]; }
  } else {
    if (vtn <= -mu * vn) { v[0] = v[1] = v[2] = 0.0f; }
    else {
      float k = mu * vn / vtn;
      v[0] = vt[0] + k * vt[0]; v[1] = vt[1] + k * vt[1]; v[2] = vt[2] + k * vt[2];
    }
  }
}

}  // namespace mpm
