// mpm_oracle.cpp -- CPU restatement of the reference explicit MPM substep.
//
// TEST INFRASTRUCTURE ONLY. This file is the correctness oracle for the HIP library: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as the
// checker or as the timed CPU baseline. The product path never calls it.
//
// It restates, without Eigen, the algorithm of
//   AnisotropicElastoplasticity/HybridSolver.cpp  18-97 (weights), 113-250 (P2G), 252-368 (forces),
//                                                 457 (gravity), 460-511 (collision), 553-577 (F),
//                                                 612-681 (return map), 725-737 (grid velocity),
//                                                 739-825 (G2P, APIC), 942-950 (advection)
//   AnisotropicElastoplasticity/interpolation.cpp 9-50 (cubic B-spline, derivative, clamp)
//   AnisotropicElastoplasticity/LevelSet.cpp       8-42 (ground plane, walls + ground)
//   AnisotropicElastoplasticity/RegularGrid.cpp    117-177 (node layout, toIndex)
// following the reference literally where that is cheap: per-particle candidate window
// i_fl-2..i_fl+2 with the positivity filter, node sums in ascending particle index, stress at the
// predicted F^ times F^n^T, the full plastic deformation gradient F_P for snow (HybridSolver.cpp:
// 676-677), the reference forms of advection (sum_i w (x_i + dt v_i)) and of B
// (sum_i w v_i x_i^T - v_p x_p^T) when `reference_forms` is set. Every deviation the
// configurations need is a parameter switch (mpm_params), not a second implementation.
//
// Templated on the scalar type: double is "the FP64 run of the reference", float is a diagnostic.
// The 3x3 SVD is a one-sided (Hestenes) Jacobi iterated to convergence with singular values
// non-negative and sorted in decreasing order (Eigen JacobiSVD's convention; every use in the
// reference is an isotropic function of Sigma recombined with the same U and V, so the sign
// and ordering conventions of U, V are not observable).
#include "mpm_c_abi.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <type_traits>
#include <vector>

// The all-cores CPU line builds this file a second time with OpenMP: the particle and node loops
// run in parallel and grid scatters become atomic adds (the sums then run in thread order, not in
// ascending particle order). Without OpenMP the macros are empty and the code is the serial oracle.
#ifdef _OPENMP
#include <omp.h>
#define ORC_PARALLEL _Pragma("omp parallel")
#define ORC_FOR _Pragma("omp for schedule(static)")
#define ORC_PAR_FOR _Pragma("omp parallel for schedule(static)")
#define ORC_ATOMIC _Pragma("omp atomic")
#else
#define ORC_PARALLEL
#define ORC_FOR
#define ORC_PAR_FOR
#define ORC_ATOMIC
#endif

namespace orc {

// ---------------------------------------------------------------- B-splines
// Cubic B-spline and derivative exactly as interpolation.cpp:9-33.
template <class T> T cubic_N(T x) {
  T ax = std::fabs(x);
  return (ax >= T(2)) ? T(0)
                      : (ax >= T(1) ? T(-1) / T(6) * ax * ax * ax + ax * ax - T(2) * ax + T(4) / T(3)
                                    : T(0.5) * ax * ax * ax - ax * ax + T(2) / T(3));
}
template <class T> T cubic_dN(T x) {
  return (x >= T(2))    ? T(0)
         : (x >= T(1))  ? T(-0.5) * x * x + T(2) * x - T(2)
         : (x >= T(0))  ? T(1.5) * x * x - T(2) * x
         : (x >= T(-1)) ? T(-1.5) * x * x - T(2) * x
         : (x >= T(-2)) ? T(0.5) * x * x + T(2) * x + T(2)
                        : T(0);
}
// Quadratic B-spline: N(r) = 3/4 - r^2 (|r| < 1/2), (3/2 - |r|)^2 / 2 (|r| < 3/2), 0 otherwise.
template <class T> T quad_N(T x) {
  T ax = std::fabs(x);
  if (ax < T(0.5)) return T(0.75) - ax * ax;
  if (ax < T(1.5)) return T(0.5) * (T(1.5) - ax) * (T(1.5) - ax);
  return T(0);
}
template <class T> T quad_dN(T x) {
  T ax = std::fabs(x);
  if (ax < T(0.5)) return T(-2) * x;
  if (ax < T(1.5)) return (x > 0 ? T(-1) : T(1)) * (T(1.5) - ax);
  return T(0);
}
template <class T> T spline_N(int order, T x) { return order == 3 ? cubic_N(x) : quad_N(x); }
template <class T> T spline_dN(int order, T x) { return order == 3 ? cubic_dN(x) : quad_dN(x); }

// clamp as interpolation.cpp:35-50
template <class T> T ref_clamp(T x, T lo, T hi) { return x > hi ? hi : (x < lo ? lo : x); }

// ---------------------------------------------------------------- 3x3 linear algebra
template <class T> struct M3 {
  T a[9];  // row-major
  T& operator()(int r, int c) { return a[3 * r + c]; }
  T operator()(int r, int c) const { return a[3 * r + c]; }
  static M3 zero() { M3 m; std::fill(m.a, m.a + 9, T(0)); return m; }
  static M3 eye() { M3 m = zero(); m.a[0] = m.a[4] = m.a[8] = T(1); return m; }
};
template <class T> M3<T> mul(const M3<T>& A, const M3<T>& B) {
  M3<T> C = M3<T>::zero();
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) C(r, c) += A(r, k) * B(k, c);
  return C;
}
template <class T> M3<T> transpose(const M3<T>& A) {
  M3<T> B;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) B(r, c) = A(c, r);
  return B;
}
template <class T> T det(const M3<T>& A) {
  return A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) - A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
         A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
}
// inverse-transpose (cofactor / det)
template <class T> M3<T> inv_transpose(const M3<T>& A) {
  M3<T> C;
  C(0, 0) = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
  C(0, 1) = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2);
  C(0, 2) = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
  C(1, 0) = A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2);
  C(1, 1) = A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0);
  C(1, 2) = A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1);
  C(2, 0) = A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1);
  C(2, 1) = A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2);
  C(2, 2) = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
  T d = det(A);
  for (int i = 0; i < 9; ++i) C.a[i] /= d;
  return C;
}

// One-sided Jacobi SVD: A = U diag(S) V^T, S >= 0 sorted decreasing.
template <class T> void svd3(const M3<T>& A, M3<T>& U, T S[3], M3<T>& V) {
  M3<T> W = A;
  V = M3<T>::eye();
  const T eps = std::numeric_limits<T>::epsilon();
  for (int sweep = 0; sweep < 100; ++sweep) {
    bool rotated = false;
    const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int k = 0; k < 3; ++k) {
      int p = P[k], q = Q[k];
      T alpha = 0, beta = 0, gamma = 0;
      for (int r = 0; r < 3; ++r) {
        alpha += W(r, p) * W(r, p);
        beta += W(r, q) * W(r, q);
        gamma += W(r, p) * W(r, q);
      }
      if (gamma == T(0) || std::fabs(gamma) <= eps * std::sqrt(alpha * beta)) continue;
      rotated = true;
      T zeta = (beta - alpha) / (T(2) * gamma);
      T t = (zeta >= 0 ? T(1) : T(-1)) / (std::fabs(zeta) + std::sqrt(T(1) + zeta * zeta));
      T c = T(1) / std::sqrt(T(1) + t * t), s = c * t;
      for (int r = 0; r < 3; ++r) {
        T wp = W(r, p), wq = W(r, q);
        W(r, p) = c * wp - s * wq;
        W(r, q) = s * wp + c * wq;
        T vp = V(r, p), vq = V(r, q);
        V(r, p) = c * vp - s * vq;
        V(r, q) = s * vp + c * vq;
      }
    }
    if (!rotated) break;
  }
  T s[3];
  for (int c = 0; c < 3; ++c) s[c] = std::sqrt(W(0, c) * W(0, c) + W(1, c) * W(1, c) + W(2, c) * W(2, c));
  int ord[3] = {0, 1, 2};
  std::sort(ord, ord + 3, [&](int a, int b) { return s[a] > s[b]; });
  M3<T> V2 = V;
  for (int c = 0; c < 3; ++c) {
    S[c] = s[ord[c]];
    for (int r = 0; r < 3; ++r) {
      V(r, c) = V2(r, ord[c]);
      U(r, c) = S[c] > T(0) ? W(r, ord[c]) / S[c] : T(0);
    }
  }
  // complete U where singular values vanish
  auto cross = [&](int a, int b, int out) {
    U(0, out) = U(1, a) * U(2, b) - U(2, a) * U(1, b);
    U(1, out) = U(2, a) * U(0, b) - U(0, a) * U(2, b);
    U(2, out) = U(0, a) * U(1, b) - U(1, a) * U(0, b);
  };
  if (S[1] == T(0)) {
    // pick any unit vector orthogonal to u0
    T ax = std::fabs(U(0, 0)), ay = std::fabs(U(1, 0)), az = std::fabs(U(2, 0));
    T e[3] = {0, 0, 0};
    e[(ax <= ay && ax <= az) ? 0 : (ay <= az ? 1 : 2)] = 1;
    T d = e[0] * U(0, 0) + e[1] * U(1, 0) + e[2] * U(2, 0);
    T w[3] = {e[0] - d * U(0, 0), e[1] - d * U(1, 0), e[2] - d * U(2, 0)};
    T nw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int r = 0; r < 3; ++r) U(r, 1) = w[r] / nw;
  }
  if (S[2] == T(0)) cross(0, 1, 2);
}

// Gram-Schmidt QR exactly as geometry.cpp:31-62 (columns d1, d2, d3).
template <class T> void gram_schmidt_qr(const M3<T>& D, M3<T>& Q, M3<T>& R) {
  T d1[3] = {D(0, 0), D(1, 0), D(2, 0)}, d2[3] = {D(0, 1), D(1, 1), D(2, 1)}, d3[3] = {D(0, 2), D(1, 2), D(2, 2)};
  auto dot = [](const T* a, const T* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  T r11 = std::sqrt(dot(d1, d1));
  T q1[3] = {d1[0] / r11, d1[1] / r11, d1[2] / r11};
  T r12 = dot(d2, q1);
  T q2[3] = {d2[0] - r12 * q1[0], d2[1] - r12 * q1[1], d2[2] - r12 * q1[2]};
  T r22 = std::sqrt(dot(q2, q2));
  for (int i = 0; i < 3; ++i) q2[i] /= r22;
  T r13 = dot(d3, q1), r23 = dot(d3, q2);
  T q3[3];
  for (int i = 0; i < 3; ++i) q3[i] = d3[i] - r13 * q1[i] - r23 * q2[i];
  T r33 = std::sqrt(dot(q3, q3));
  for (int i = 0; i < 3; ++i) q3[i] /= r33;
  for (int i = 0; i < 3; ++i) {
    Q(i, 0) = q1[i];
    Q(i, 1) = q2[i];
    Q(i, 2) = q3[i];
  }
  R = M3<T>::zero();
  R(0, 0) = r11; R(0, 1) = r12; R(0, 2) = r13;
  R(1, 1) = r22; R(1, 2) = r23;
  R(2, 2) = r33;
}

// ---------------------------------------------------------------- constitutive pieces
// Lame parameters (HybridSolver.cpp:264-265, 637-638).
template <class T> void lame(const mpm_params& P, T& lambda, T& mu) {
  T E = P.youngs_modulus, nu = P.poisson_ratio;
  lambda = E * nu / (T(1) + nu) / (T(1) - T(2) * nu);
  mu = E / T(2) / (T(1) + nu);
}

// First Piola-Kirchhoff stress P(Fh) (HybridSolver.cpp:279-339).
template <class T> M3<T> piola(const mpm_params& P, const M3<T>& Fh, T Jp) {
  T lambda0, mu0;
  lame<T>(P, lambda0, mu0);
  M3<T> U, V;
  T S[3];
  svd3(Fh, U, S, V);
  M3<T> out = M3<T>::zero();
  if (P.material == MPM_SNOW) {
    T sc = std::exp(T(P.hardening) * (T(1) - Jp));
    T lambda = lambda0 * sc, mu = mu0 * sc;
    M3<T> R = mul(U, transpose(V));
    T J = det(Fh);
    M3<T> Fit = inv_transpose(Fh);
    for (int i = 0; i < 9; ++i) out.a[i] = T(2) * mu * (Fh.a[i] - R.a[i]) + lambda * (J - T(1)) * J * Fit.a[i];
  } else {
    T ln[3] = {std::log(S[0]), std::log(S[1]), std::log(S[2])};
    T tr = ln[0] + ln[1] + ln[2];
    T d[3];
    for (int k = 0; k < 3; ++k) d[k] = T(2) * mu0 * (ln[k] / S[k]) + lambda0 * tr / S[k];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) out(r, c) = U(r, 0) * d[0] * V(c, 0) + U(r, 1) * d[1] * V(c, 1) + U(r, 2) * d[2] * V(c, 2);
  }
  return out;
}

// Drucker-Prager projection on principal stretches (HybridSolver.cpp:634-673), returns the
// new Sigma and the increment of plasticAmount. phi0 replaces h0 (degrees).
template <class T> void dp_project(const mpm_params& P, T phi0, T q, const T Sin[3], T Sout[3], T& dq) {
  T lambda, mu;
  lame<T>(P, lambda, mu);
  const T pi = T(3.14159265358979323846);
  T h1 = P.dp_h[1], h2 = P.dp_h[2], h3 = P.dp_h[3];
  T frictionAngle = (phi0 + (h1 * q - h3) * std::exp(-h2 * q)) * pi / T(180);
  T alpha = std::sqrt(T(2) / T(3)) * T(2) * std::sin(frictionAngle) / (T(3) - std::sin(frictionAngle));
  T ln[3] = {std::log(Sin[0]), std::log(Sin[1]), std::log(Sin[2])};
  T tr = ln[0] + ln[1] + ln[2];
  T dev[3] = {ln[0] - tr / T(3), ln[1] - tr / T(3), ln[2] - tr / T(3)};
  T devn = std::sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  T dgamma = devn + (T(3) * lambda + T(2) * mu) / T(2) / mu * tr * alpha;
  dq = 0;
  for (int k = 0; k < 3; ++k) Sout[k] = Sin[k];
  if (dgamma <= T(0)) {
    // elastic
  } else if (devn == T(0) || tr > T(0)) {
    for (int k = 0; k < 3; ++k) Sout[k] = T(1);
    dq = std::sqrt(ln[0] * ln[0] + ln[1] * ln[1] + ln[2] * ln[2]);
  } else {
    for (int k = 0; k < 3; ++k) Sout[k] = std::exp(ln[k] - dgamma * dev[k] / devn);
    dq = dgamma;
  }
}

// ---------------------------------------------------------------- level sets
template <class T> void levelset(const std::vector<mpm_levelset>& prims, const T x[3], T& phi, T n[3]) {
  phi = std::numeric_limits<T>::infinity();
  n[0] = n[1] = n[2] = 0;
  for (const auto& pr : prims) {
    T p, nn[3] = {0, 0, 0};
    if (pr.type == MPM_LS_PLANE) {
      T a0 = pr.a[0], a1 = pr.a[1], a2 = pr.a[2];
      T len = std::sqrt(a0 * a0 + a1 * a1 + a2 * a2);
      nn[0] = a0 / len; nn[1] = a1 / len; nn[2] = a2 / len;
      p = nn[0] * x[0] + nn[1] * x[1] + nn[2] * x[2] - T(pr.a[3]);
    } else if (pr.type == MPM_LS_WALL2GROUND) {
      // LevelSet.cpp:29-42 (wallX, wallY, groundZ)
      T wx = pr.a[0], wy = pr.a[1], gz = pr.a[2];
      p = std::min(std::min(x[2] - gz, wx - x[0]), wy - x[1]);
      T dz = std::fabs(x[2] - gz), dx = std::fabs(wx - x[0]), dy = std::fabs(wy - x[1]);
      if (dz <= dx && dz <= dy) nn[2] = 1;
      else if (dy <= dx) nn[1] = -1;
      else nn[0] = -1;
    } else if (pr.type == MPM_LS_BOX) {
      p = std::numeric_limits<T>::infinity();
      for (int ax = 0; ax < 3; ++ax) {
        T dlo = x[ax] - T(pr.a[ax]), dhi = T(pr.b[ax]) - x[ax];
        if (dlo < p) { p = dlo; nn[0] = nn[1] = nn[2] = 0; nn[ax] = 1; }
        if (dhi < p) { p = dhi; nn[0] = nn[1] = nn[2] = 0; nn[ax] = -1; }
      }
    } else {  // SPHERE
      T d[3] = {x[0] - T(pr.a[0]), x[1] - T(pr.a[1]), x[2] - T(pr.a[2])};
      T len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      p = len - T(pr.a[3]);
      if (len > T(0)) { nn[0] = d[0] / len; nn[1] = d[1] / len; nn[2] = d[2] / len; }
      else nn[1] = 1;
    }
    if (p < phi) { phi = p; n[0] = nn[0]; n[1] = nn[1]; n[2] = nn[2]; }
  }
}

// Forces of one cloth element from its current and rest directions (HybridSolver.cpp:370-455,
// LagrangianMesh::computeVertexInPlaneForces LagrangianMesh.cpp:382-460): the in-plane vertex forces
// f (f1, f2, f3 at f[0..2], f[3..5], f[6..8]) and the element stress Ms = V dF3 d3^T.
template <class T>
void cloth_element(const M3<T>& D, const M3<T>& Dr, T vol, T c_mu, T c_lambda, T c_gamma, T c_k, bool ref_normal,
                   T* fout, M3<T>& Ms_) {
  M3<T> Q, R, rQ, rR;
  gram_schmidt_qr(D, Q, R);
  gram_schmidt_qr(Dr, rQ, rR);
  // inverseR of the rest in-plane R (geometry.cpp:67-73)
  const T i11 = T(1) / rR(0, 0), i12 = -rR(0, 1) / rR(0, 0) / rR(1, 1), i22 = T(1) / rR(1, 1);
  // A = invRest * inPlaneR (both upper triangular)
  const T a00 = i11 * R(0, 0), a01 = i11 * R(0, 1) + i12 * R(1, 1), a10 = T(0), a11 = i22 * R(1, 1);
  // polar rotation U V^T of A (orthogonal factor of its SVD)
  T r00, r01, r10, r11;
  const T dA = a00 * a11 - a01 * a10;
  if (dA >= T(0)) {
    T pp = a00 + a11, qq = a10 - a01, nn = std::sqrt(pp * pp + qq * qq);
    if (nn > T(0)) { pp /= nn; qq /= nn; } else { pp = T(1); qq = T(0); }
    r00 = pp; r01 = -qq; r10 = qq; r11 = pp;
  } else {
    T pp = a00 - a11, qq = a01 + a10, nn = std::sqrt(pp * pp + qq * qq);
    pp /= nn; qq /= nn;
    r00 = pp; r01 = qq; r10 = qq; r11 = -pp;
  }
  const T J = a00 * a11;  // product of the diagonal (:443)
  // P = 2 mu (A - R) + lambda (J - 1) [a11 -a01; 0 a00]^T (:445-446)
  const T p00 = T(2) * c_mu * (a00 - r00) + c_lambda * (J - T(1)) * a11;
  const T p01 = T(2) * c_mu * (a01 - r01);
  const T p10 = T(2) * c_mu * (a10 - r10) + c_lambda * (J - T(1)) * (-a01);
  const T p11 = T(2) * c_mu * (a11 - r11) + c_lambda * (J - T(1)) * a00;
  // vertex forces (:450-458)
  for (int a = 0; a < 3; ++a) {
    const T q1 = Q(a, 0), q2 = Q(a, 1);
    const T f2 = -(p00 * i11 + p01 * i12) * q1;
    const T f3 = -p01 * i22 * q1 - p11 * i22 * q2;
    fout[3 + a] = f2;
    fout[6 + a] = f3;
    fout[a] = -(f2 + f3);
  }
  // in-plane PK of the element: invRest * P (:448)
  const T d11 = i11 * p00 + i12 * p10, d12 = i11 * p01 + i12 * p11, d22 = i22 * p11;
  // normal and shear response (HybridSolver.cpp:406-430)
  M3<T> dR = M3<T>::zero();
  dR(0, 0) = d11; dR(0, 1) = d12; dR(0, 2) = c_gamma * R(0, 2);
  dR(1, 1) = d22; dR(1, 2) = c_gamma * R(1, 2);
  dR(2, 2) = R(2, 2) > T(1) ? T(0) : -c_k * (T(1) - R(2, 2)) * (T(1) - R(2, 2));
  M3<T> K = mul(dR, transpose(R));
  M3<T> Ks;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Ks(i, j) = K(std::min(i, j), std::max(i, j));
  M3<T> Rit = inv_transpose(R);
  T cn[3];
  if (ref_normal) {
    for (int i = 0; i < 3; ++i) cn[i] = Dr(2, i);  // (D_rest^T).col(2) (:432)
  } else {
    M3<T> Drit = inv_transpose(Dr);  // D_rest^-T e3
    for (int i = 0; i < 3; ++i) cn[i] = Drit(i, 2);
  }
  M3<T> QK = mul(mul(Q, Ks), Rit);
  T dF3[3];
  for (int i = 0; i < 3; ++i) dF3[i] = QK(i, 0) * cn[0] + QK(i, 1) * cn[1] + QK(i, 2) * cn[2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Ms_(i, j) = vol * dF3[i] * D(j, 2);  // V_e dF3 d3^T (:434)
}

// Cloth return mapping on the QR factor of D (HybridSolver.cpp:684-722); returns the new D.
template <class T> M3<T> cloth_return_map(M3<T> D, T c_gamma, T c_k, T c_cf) {
  M3<T> Q, R;
  gram_schmidt_qr(D, Q, R);
  if (R(2, 2) > T(1)) {
    R(2, 2) = T(1);
    R(0, 2) = R(1, 2) = T(0);
  } else {
    const T fn = c_k * (R(2, 2) - T(1)) * (R(2, 2) - T(1));
    const T ft = c_gamma * std::sqrt(R(0, 2) * R(0, 2) + R(1, 2) * R(1, 2));
    if (ft > c_cf * fn) {
      R(0, 2) *= c_cf * fn / ft;
      R(1, 2) *= c_cf * fn / ft;
    }
  }
  for (int a = 0; a < 3; ++a) D(a, 2) = Q(a, 0) * R(0, 2) + Q(a, 1) * R(1, 2) + Q(a, 2) * R(2, 2);
  return D;
}

// ---------------------------------------------------------------- the simulation
struct NodeW {
  int idx;
  double w, dw[3];
};

template <class T> struct Sim {
  mpm_params P;
  bool reference_forms = true;
  T gmin[3], gmax[3], h[3];
  int res[3];
  int ng = 0;
  std::vector<mpm_levelset> ls;
  int n = 0;
  std::vector<T> x, v, B, F, FP, plastic, m, V, phi;
  std::vector<T> gm, gp, gf, gv, gvbf;
  std::vector<T> gnode;  // node positions (3 planes)
  uint32_t status = 0;

  // cloth: LagrangianMesh (LagrangianMesh.h:13-117). Vertices and elements (triangles) are MPM
  // particles; element e carries the directions D = [d1 d2 d3] (row-major M3 per element) and its
  // rest directions; planes are [component][index].
  int nv = 0, ne = 0;
  std::vector<T> vx, vv, vB, vm;      // vertices: x, v, B, m
  std::vector<T> ex, ev, eB, em, eV;  // elements: x (centroid), v, B, m, volume
  std::vector<T> eD, eDr;             // directions, rest directions (9 planes each)
  std::vector<int> faces;             // [e][3] vertex indices
  std::vector<uint8_t> vfixed;
  T c_mu = 0, c_lambda = 0, c_gamma = 0, c_k = 0, c_cf = 0;
  bool c_ref_normal = false;  // true: the reference's (D_rest^T).col(2); false: D_rest^-T e3

  void set_grid(const double mn[3], const double mx[3], const int r[3]) {
    for (int a = 0; a < 3; ++a) {
      gmin[a] = T(mn[a]);
      gmax[a] = T(mx[a]);
      res[a] = r[a];
      h[a] = (gmax[a] - gmin[a]) / T(res[a]);  // RegularGrid.cpp:144-146
    }
    ng = res[0] * res[1] * res[2];
    gm.assign(ng, 0); gp.assign(3 * (size_t)ng, 0); gf.assign(3 * (size_t)ng, 0);
    gv.assign(3 * (size_t)ng, 0); gvbf.assign(3 * (size_t)ng, 0); gnode.assign(3 * (size_t)ng, 0);
    for (int k = 0; k < res[2]; ++k)
      for (int j = 0; j < res[1]; ++j)
        for (int i = 0; i < res[0]; ++i) {
          int id = to_index(i, j, k);
          gnode[id] = gmin[0] + T(i) * h[0];  // RegularGrid.cpp:154-159
          gnode[ng + id] = gmin[1] + T(j) * h[1];
          gnode[2 * (size_t)ng + id] = gmin[2] + T(k) * h[2];
        }
  }
  int to_index(int i, int j, int k) const { return k * res[0] * res[1] + j * res[0] + i; }  // :164-168
  T hmin() const { return std::min(h[0], std::min(h[1], h[2])); }
  T dinv() const { T hm = hmin(); return (P.kernel_order == 3 ? T(3) : T(4)) / hm / hm; }

  // evaluateInterpolationWeights_ for one particle (HybridSolver.cpp:30-79)
  void weights(int p, std::vector<NodeW>& out) const {
    const T xp[3] = {x[p], x[n + p], x[2 * (size_t)n + p]};
    weights_at(xp, out);
  }
  void weights_at(const T xp[3], std::vector<NodeW>& out) const {
    out.clear();
    T pr[3] = {xp[0] - gmin[0], xp[1] - gmin[1], xp[2] - gmin[2]};
    int fl[3];
    for (int a = 0; a < 3; ++a) fl[a] = static_cast<int>(pr[a] / h[a]);
    for (int i = fl[0] - 2; i <= fl[0] + 2; ++i)
      for (int j = fl[1] - 2; j <= fl[1] + 2; ++j)
        for (int k = fl[2] - 2; k <= fl[2] + 2; ++k) {
          if (!(0 <= i && i < res[0] && 0 <= j && j < res[1] && 0 <= k && k < res[2])) continue;
          T ri = pr[0] / h[0] - T(i), rj = pr[1] / h[1] - T(j), rk = pr[2] / h[2] - T(k);
          T wi = spline_N(P.kernel_order, ri), wj = spline_N(P.kernel_order, rj), wk = spline_N(P.kernel_order, rk);
          if (!(wi > 0 && wj > 0 && wk > 0)) continue;
          T di = spline_dN(P.kernel_order, ri) / h[0], dj = spline_dN(P.kernel_order, rj) / h[1],
            dk = spline_dN(P.kernel_order, rk) / h[2];
          NodeW nw;
          nw.idx = to_index(i, j, k);
          nw.w = double(wi * wj * wk);
          nw.dw[0] = double(di * wj * wk);
          nw.dw[1] = double(wi * dj * wk);
          nw.dw[2] = double(wi * wj * dk);
          out.push_back(nw);
        }
  }

  M3<T> getM(const std::vector<T>& arr, int p) const {
    M3<T> M;
    for (int c = 0; c < 9; ++c) M.a[c] = arr[(size_t)c * n + p];
    return M;
  }
  void setM(std::vector<T>& arr, int p, const M3<T>& M) {
    for (int c = 0; c < 9; ++c) arr[(size_t)c * n + p] = M.a[c];
  }
  T X(int a, int p) const { return x[(size_t)a * n + p]; }

  // particleToGrid_ (HybridSolver.cpp:113-240): masses, momenta, velocities.
  void p2g() {
    std::fill(gm.begin(), gm.end(), T(0));
    std::fill(gp.begin(), gp.end(), T(0));
    T Di = dinv();
    ORC_PARALLEL
    {
      std::vector<NodeW> nw;
      ORC_FOR
      for (int p = 0; p < n; ++p) {
        weights(p, nw);
        M3<T> Bp = getM(B, p);
        for (const auto& q : nw) {
          T w = T(q.w);
          ORC_ATOMIC
          gm[q.idx] += w * m[p];
          T dx[3];
          for (int a = 0; a < 3; ++a) dx[a] = gnode[(size_t)a * ng + q.idx] - X(a, p);
          for (int a = 0; a < 3; ++a) {
            T aff = Bp(a, 0) * dx[0] + Bp(a, 1) * dx[1] + Bp(a, 2) * dx[2];
            ORC_ATOMIC
            gp[(size_t)a * ng + q.idx] += w * m[p] * (v[(size_t)a * n + p] + Di * aff);
          }
        }
      }
    }
    // cloth vertices and elements (:121-125, 137-141, 216-230)
    std::vector<NodeW> nw;
    for (int k = 0; k < nv; ++k) p2g_one(vx, vv, vB, vm, nv, k, Di, nw);
    for (int k = 0; k < ne; ++k) p2g_one(ex, ev, eB, em, ne, k, Di, nw);
    ORC_PAR_FOR
    for (int c = 0; c < ng; ++c)
      for (int a = 0; a < 3; ++a)
        gv[(size_t)a * ng + c] = gm[c] > T(0) ? gp[(size_t)a * ng + c] / gm[c] : T(0);  // :233-240
  }
  // mass and APIC momentum of one particle of a planar set (cloth vertices or elements)
  void p2g_one(const std::vector<T>& X_, const std::vector<T>& V_, const std::vector<T>& B_, const std::vector<T>& M_,
               int cnt, int k, T Di, std::vector<NodeW>& nw) {
    const T xp[3] = {X_[k], X_[cnt + k], X_[2 * (size_t)cnt + k]};
    weights_at(xp, nw);
    for (const auto& q : nw) {
      T w = T(q.w);
      gm[q.idx] += w * M_[k];
      T dx[3];
      for (int a = 0; a < 3; ++a) dx[a] = gnode[(size_t)a * ng + q.idx] - xp[a];
      for (int a = 0; a < 3; ++a) {
        T aff = B_[(size_t)(3 * a) * cnt + k] * dx[0] + B_[(size_t)(3 * a + 1) * cnt + k] * dx[1] +
                B_[(size_t)(3 * a + 2) * cnt + k] * dx[2];
        gp[(size_t)a * ng + q.idx] += w * M_[k] * (V_[(size_t)a * cnt + k] + Di * aff);
      }
    }
  }

  // Particle part of computeGridForces_ (HybridSolver.cpp:252-368); gravity is added by grid_update.
  void forces(T dt) {
    std::fill(gf.begin(), gf.end(), T(0));
    ORC_PARALLEL
    {
    std::vector<NodeW> nw;
    ORC_FOR
    for (int p = 0; p < n; ++p) {
      weights(p, nw);
      M3<T> Fn = getM(F, p);
      M3<T> Fh = Fn;
      if (P.stress_at_predicted_F) {
        M3<T> G = M3<T>::zero();  // dt * grad v (:269-271, 297-301)
        for (const auto& q : nw)
          for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) G(a, k) += dt * gv[(size_t)a * ng + q.idx] * T(q.dw[k]);
        M3<T> GF = mul(G, Fn);
        for (int i = 0; i < 9; ++i) Fh.a[i] = Fn.a[i] + GF.a[i];  // :306
      }
      T Jp = (P.material == MPM_SNOW) ? jp_of(p) : T(1);
      M3<T> Pk = piola<T>(P, Fh, Jp);
      M3<T> st = mul(Pk, transpose(Fn));  // volume * P * EDG^T (:321-338)
      for (int i = 0; i < 9; ++i) st.a[i] *= V[p];
      for (const auto& q : nw)
        for (int a = 0; a < 3; ++a) {
          ORC_ATOMIC
          gf[(size_t)a * ng + q.idx] -= st(a, 0) * T(q.dw[0]) + st(a, 1) * T(q.dw[1]) + st(a, 2) * T(q.dw[2]);
        }
    }
    }
    cloth_forces();
  }
  T jp_of(int p) const { return det(getM(FP, p)); }  // :283

  M3<T> elemM(const std::vector<T>& arr, int e) const {
    M3<T> M;
    for (int c = 0; c < 9; ++c) M.a[c] = arr[(size_t)c * ne + e];
    return M;
  }
  // Cloth forces (HybridSolver.cpp:370-455): in-plane fixed-corotated forces on the vertices
  // (LagrangianMesh::computeVertexInPlaneForces, LagrangianMesh.cpp:382-460) spread by the vertex
  // weights, and the element stress V_e dF3 d3^T of the normal/shear response spread by the element
  // weight gradients.
  void cloth_forces() {
    if (ne == 0) return;
    std::vector<T> fv(3 * (size_t)nv, T(0));
    std::vector<M3<T>> Ms(ne);
    for (int e = 0; e < ne; ++e) {
      T f[9];
      cloth_element<T>(elemM(eD, e), elemM(eDr, e), eV[e], c_mu, c_lambda, c_gamma, c_k, c_ref_normal, f, Ms[e]);
      for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) fv[(size_t)a * nv + faces[3 * e + c]] += f[3 * c + a];
    }
    std::vector<NodeW> nw;
    for (int k = 0; k < nv; ++k) {
      const T xp[3] = {vx[k], vx[nv + k], vx[2 * (size_t)nv + k]};
      weights_at(xp, nw);
      for (const auto& q : nw)
        for (int a = 0; a < 3; ++a) gf[(size_t)a * ng + q.idx] += T(q.w) * fv[(size_t)a * nv + k];
    }
    for (int e = 0; e < ne; ++e) {
      const T xp[3] = {ex[e], ex[ne + e], ex[2 * (size_t)ne + e]};
      weights_at(xp, nw);
      for (const auto& q : nw)
        for (int a = 0; a < 3; ++a)
          gf[(size_t)a * ng + q.idx] -= Ms[e](a, 0) * T(q.dw[0]) + Ms[e](a, 1) * T(q.dw[1]) + Ms[e](a, 2) * T(q.dw[2]);
    }
  }

  // gravity (:457), updateGridVelocities_ (:725-737), gridCollisionHandling_ (:460-511)
  void grid_update(T dt) {
    velocity_update(dt);
    collide();
  }
  void velocity_update(T dt) {
    ORC_PAR_FOR
    for (int c = 0; c < ng; ++c) {
      if (!(gm[c] > T(0))) continue;
      for (int a = 0; a < 3; ++a) {
        T f = gf[(size_t)a * ng + c] + gm[c] * T(P.gravity[a]);
        gv[(size_t)a * ng + c] += dt * f / gm[c];
      }
    }
  }
  // RegularGrid::max_velocity (RegularGrid.cpp:188-200): largest node speed (zero where m = 0)
  double max_speed() const {
    double vm = 0.0;
    for (int c = 0; c < ng; ++c) {
      double s2 = 0.0;
      for (int a = 0; a < 3; ++a) s2 += double(gv[(size_t)a * ng + c]) * double(gv[(size_t)a * ng + c]);
      vm = std::max(vm, std::sqrt(s2));
    }
    return vm;
  }
  void collide() {
    gvbf = gv;
    const T mu = T(P.friction);
    ORC_PAR_FOR
    for (int c = 0; c < ng; ++c) {
      T xi[3] = {gnode[c], gnode[ng + c], gnode[2 * (size_t)ng + c]};
      T phi_, nrm[3];
      if (ls.empty()) continue;
      levelset<T>(ls, xi, phi_, nrm);
      if (!(phi_ <= T(0))) continue;
      T vel[3] = {gv[c], gv[ng + c], gv[2 * (size_t)ng + c]};
      if (P.friction_model == MPM_FRICTION_STICKY) {
        for (int a = 0; a < 3; ++a) gv[(size_t)a * ng + c] = gvbf[(size_t)a * ng + c] = T(0);
        continue;
      }
      T vn = vel[0] * nrm[0] + vel[1] * nrm[1] + vel[2] * nrm[2];
      if (!(vn < T(0))) continue;
      T vt[3] = {vel[0] - vn * nrm[0], vel[1] - vn * nrm[1], vel[2] - vn * nrm[2]};
      T vtn = std::sqrt(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
      T out[3];
      if (P.friction_model == MPM_FRICTION_REFERENCE) {
        if (vtn < -mu * vn) out[0] = out[1] = out[2] = 0;
        else for (int a = 0; a < 3; ++a) out[a] = vt[a];
      } else {
        if (vtn <= -mu * vn) out[0] = out[1] = out[2] = 0;
        else for (int a = 0; a < 3; ++a) out[a] = vt[a] + mu * vn * vt[a] / vtn;
      }
      for (int a = 0; a < 3; ++a) {
        gvbf[(size_t)a * ng + c] = vt[a];
        gv[(size_t)a * ng + c] = out[a];
      }
    }
    fix_vertices();
  }
  // Fixed cloth vertices (HybridSolver.cpp:513-550): for every node (ri, rj, rk) with positive weight
  // for a fixed vertex (in-grid nodes only, as the weight matrix holds them), each neighbour
  // (i, j, k) in ri-1..ri+1 x rj-1..rj+1 x rk-1..rk+1 (the test `<= 12` at :532-535 holds for all 27)
  // is stopped (v = v_BF = 0) through its linear index toIndex(i, j, k) = k r0 r1 + j r0 + i, which
  // is only checked against [0, gridNumber) (:536-539). A neighbour one step outside the grid in x
  // or y therefore stops the node its linear index names: (-1, j, k) stops (r0 - 1, j - 1, k).
  void fix_vertices() {
    std::vector<NodeW> nw;
    for (int k = 0; k < nv; ++k) {
      if (!vfixed[k]) continue;
      const T xp[3] = {vx[k], vx[nv + k], vx[2 * (size_t)nv + k]};
      weights_at(xp, nw);
      for (const auto& q : nw) {
        const int ri = q.idx % res[0], rj = (q.idx / res[0]) % res[1], rk = q.idx / (res[0] * res[1]);
        for (int i = ri - 1; i <= ri + 1; ++i)
          for (int j = rj - 1; j <= rj + 1; ++j)
            for (int kk = rk - 1; kk <= rk + 1; ++kk) {
              const long long id = (long long)kk * res[0] * res[1] + (long long)j * res[0] + i;
              if (id < 0 || id >= (long long)ng) continue;
              for (int a = 0; a < 3; ++a) gv[(size_t)a * ng + id] = gvbf[(size_t)a * ng + id] = T(0);
            }
      }
    }
  }

  // G2P (:739-825), advection (:942-950), F update (:553-577), plasticity (:612-681).
  void g2p(T dt) {
    ORC_PARALLEL
    {
    std::vector<NodeW> nw;
    ORC_FOR
    for (int p = 0; p < n; ++p) {
      weights(p, nw);
      T vp[3] = {0, 0, 0}, xs[3] = {0, 0, 0}, vbfs[3] = {0, 0, 0};
      M3<T> Bn = M3<T>::zero(), Gv = M3<T>::zero(), VX = M3<T>::zero();
      T wsum = 0;
      for (const auto& q : nw) {
        T w = T(q.w);
        wsum += w;
        T vi[3], vb[3], xi[3];
        for (int a = 0; a < 3; ++a) {
          vi[a] = gv[(size_t)a * ng + q.idx];
          vb[a] = gvbf[(size_t)a * ng + q.idx];
          xi[a] = gnode[(size_t)a * ng + q.idx];
        }
        for (int a = 0; a < 3; ++a) {
          vp[a] += w * vi[a];
          vbfs[a] += w * vb[a];
          xs[a] += w * (xi[a] + dt * vb[a]);
          for (int b = 0; b < 3; ++b) {
            VX(a, b) += w * vi[a] * xi[b];
            Bn(a, b) += w * vi[a] * (xi[b] - X(b, p));
            Gv(a, b) += vb[a] * T(q.dw[b]);
          }
        }
      }
      (void)wsum;
      if (reference_forms) {
        // B = omegas * diag(v_a) * x_grid - diag(v_p,a) * x_p (:797-806)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) Bn(a, b) = VX(a, b) - vp[a] * X(b, p);
      }
      // damping C <- skew + (1-d) sym (:808-824)
      T d = T(P.apic_damping);
      if (d != T(0)) {
        M3<T> C = Bn;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            T sym = T(0.5) * (C(a, b) + C(b, a));
            Bn(a, b) = (C(a, b) - sym) + (T(1) - d) * sym;
          }
      }
      setM(B, p, Bn);
      for (int a = 0; a < 3; ++a) v[(size_t)a * n + p] = vp[a];
      T xnew[3];
      for (int a = 0; a < 3; ++a) xnew[a] = reference_forms ? xs[a] : X(a, p) + dt * vbfs[a];
      // F^ = F + dt (sum vbf grad w^T) F, weights at x^n (:560-577)
      M3<T> Fn = getM(F, p);
      M3<T> GF = mul(Gv, Fn);
      M3<T> Fh;
      for (int i = 0; i < 9; ++i) Fh.a[i] = Fn.a[i] + dt * GF.a[i];
      for (int a = 0; a < 3; ++a) x[(size_t)a * n + p] = xnew[a];
      return_map(p, Fh);
    }
    }
    g2p_mesh(dt);
  }

  // Velocity, affine momentum (damping 1: the symmetric part is dropped, HybridSolver.cpp:808-824)
  // and the weight-gradient product sum_i v_BF,i grad w^T of one cloth particle at position xp.
  void g2p_one(const T xp[3], T dt, T vp[3], M3<T>& Bn, M3<T>& Gv, T xnew[3], std::vector<NodeW>& nw) {
    weights_at(xp, nw);
    T xs[3] = {0, 0, 0}, vbfs[3] = {0, 0, 0};
    M3<T> VX = M3<T>::zero();
    Bn = M3<T>::zero();
    Gv = M3<T>::zero();
    for (int a = 0; a < 3; ++a) vp[a] = T(0);
    for (const auto& q : nw) {
      T w = T(q.w);
      for (int a = 0; a < 3; ++a) {
        const T vi = gv[(size_t)a * ng + q.idx], vb = gvbf[(size_t)a * ng + q.idx];
        const T xi = gnode[(size_t)a * ng + q.idx];
        vp[a] += w * vi;
        vbfs[a] += w * vb;
        xs[a] += w * (xi + dt * vb);
        for (int b = 0; b < 3; ++b) {
          const T xb = gnode[(size_t)b * ng + q.idx];
          VX(a, b) += w * vi * xb;
          Bn(a, b) += w * vi * (xb - xp[b]);
          Gv(a, b) += vb * T(q.dw[b]);
        }
      }
    }
    if (reference_forms)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Bn(a, b) = VX(a, b) - vp[a] * xp[b];
    M3<T> C = Bn;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) Bn(a, b) = C(a, b) - T(0.5) * (C(a, b) + C(b, a));
    for (int a = 0; a < 3; ++a) xnew[a] = reference_forms ? xs[a] : xp[a] + dt * vbfs[a];
  }

  // G2P of the cloth (HybridSolver.cpp:739-758, 760-825, 942-950, 578-609) and the element return
  // mapping (:684-722).
  void g2p_mesh(T dt) {
    if (nv == 0 && ne == 0) return;
    std::vector<NodeW> nw;
    T vp[3], xn[3];
    M3<T> Bn, Gv;
    for (int k = 0; k < nv; ++k) {
      const T xp[3] = {vx[k], vx[nv + k], vx[2 * (size_t)nv + k]};
      g2p_one(xp, dt, vp, Bn, Gv, xn, nw);
      for (int c = 0; c < 9; ++c) vB[(size_t)c * nv + k] = Bn.a[c];
      for (int a = 0; a < 3; ++a) {
        vv[(size_t)a * nv + k] = vp[a];
        vx[(size_t)a * nv + k] = xn[a];
      }
    }
    for (int e = 0; e < ne; ++e) {
      const T xp[3] = {ex[e], ex[ne + e], ex[2 * (size_t)ne + e]};
      g2p_one(xp, dt, vp, Bn, Gv, xn, nw);
      for (int c = 0; c < 9; ++c) eB[(size_t)c * ne + e] = Bn.a[c];
      const int v0 = faces[3 * e], v1 = faces[3 * e + 1], v2 = faces[3 * e + 2];
      M3<T> D = elemM(eD, e);
      T d3[3];
      for (int a = 0; a < 3; ++a) d3[a] = D(a, 2) + dt * (Gv(a, 0) * D(0, 2) + Gv(a, 1) * D(1, 2) + Gv(a, 2) * D(2, 2));
      for (int a = 0; a < 3; ++a) {
        const T p0 = vx[(size_t)a * nv + v0], p1 = vx[(size_t)a * nv + v1], p2 = vx[(size_t)a * nv + v2];
        ex[(size_t)a * ne + e] = (p0 + p1 + p2) / T(3);
        ev[(size_t)a * ne + e] = (vv[(size_t)a * nv + v0] + vv[(size_t)a * nv + v1] + vv[(size_t)a * nv + v2]) / T(3);
        D(a, 0) = p1 - p0;
        D(a, 1) = p2 - p0;
        D(a, 2) = d3[a];
      }
      // return mapping on the QR factor of D (:684-722)
      D = cloth_return_map<T>(D, c_gamma, c_k, c_cf);
      for (int c = 0; c < 9; ++c) eD[(size_t)c * ne + e] = D.a[c];
    }
  }

  // LagrangianMesh::ObjMesh (LagrangianMesh.cpp:300-352) from given vertices and faces: element
  // volume 0.25 area thickness, added to each of its vertices; masses density x volume; rest and
  // current directions v2 - v1, v3 - v1 and the unit normal; element positions at the centroids.
  void set_mesh(int nv_, const double* x_, const double* v_, int ne_, const int* f_, double density,
                double thickness, double E, double nu, double gamma, double k, double angle_deg,
                const uint8_t* fixed, int ref_normal) {
    nv = nv_;
    ne = ne_;
    faces.assign(f_, f_ + 3 * (size_t)ne);
    vx.assign(x_, x_ + 3 * (size_t)nv);
    vv.assign(3 * (size_t)nv, T(0));
    if (v_) for (size_t i = 0; i < 3 * (size_t)nv; ++i) vv[i] = T(v_[i]);
    vB.assign(9 * (size_t)nv, T(0));
    std::vector<double> vvol(nv, 0.0);
    ex.assign(3 * (size_t)ne, T(0)); ev.assign(3 * (size_t)ne, T(0)); eB.assign(9 * (size_t)ne, T(0));
    em.assign(ne, T(0)); eV.assign(ne, T(0)); eD.assign(9 * (size_t)ne, T(0)); eDr.assign(9 * (size_t)ne, T(0));
    for (int e = 0; e < ne; ++e) {
      double p[3][3];
      for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) p[c][a] = x_[(size_t)a * nv + faces[3 * e + c]];
      auto len = [](const double* a, const double* b) {
        return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
      };
      const double la = len(p[0], p[1]), lb = len(p[1], p[2]), lc = len(p[2], p[0]), hp = 0.5 * (la + lb + lc);
      const double area = std::sqrt(hp * (hp - la) * (hp - lb) * (hp - lc));  // geometry.cpp:12-23
      const double vol = 0.25 * area * thickness;
      eV[e] = T(vol);
      em[e] = T(density * vol);
      for (int c = 0; c < 3; ++c) vvol[faces[3 * e + c]] += vol;
      double d1[3], d2[3], nrm[3];
      for (int a = 0; a < 3; ++a) {
        d1[a] = p[1][a] - p[0][a];
        d2[a] = p[2][a] - p[0][a];
      }
      nrm[0] = d1[1] * d2[2] - d1[2] * d2[1];
      nrm[1] = d1[2] * d2[0] - d1[0] * d2[2];
      nrm[2] = d1[0] * d2[1] - d1[1] * d2[0];
      const double nl = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
      for (int a = 0; a < 3; ++a) {
        const double col[3] = {d1[a], d2[a], nrm[a] / nl};
        for (int c = 0; c < 3; ++c) eDr[(size_t)(3 * a + c) * ne + e] = eD[(size_t)(3 * a + c) * ne + e] = T(col[c]);
        ex[(size_t)a * ne + e] = T((p[0][a] + p[1][a] + p[2][a]) / 3.0);
      }
    }
    vm.assign(nv, T(0));
    for (int k2 = 0; k2 < nv; ++k2) vm[k2] = T(density * vvol[k2]);
    for (int e = 0; e < ne; ++e)
      for (int a = 0; a < 3; ++a)
        ev[(size_t)a * ne + e] = (vv[(size_t)a * nv + faces[3 * e]] + vv[(size_t)a * nv + faces[3 * e + 1]] +
                                  vv[(size_t)a * nv + faces[3 * e + 2]]) / T(3);
    c_lambda = T(E * nu / (1.0 + nu) / (1.0 - 2.0 * nu));
    c_mu = T(E / 2.0 / (1.0 + nu));
    c_gamma = T(gamma);
    c_k = T(k);
    c_cf = T(std::tan(angle_deg * 3.14159265358979323846 / 180.0));
    c_ref_normal = ref_normal != 0;
    vfixed.assign(nv, 0);
    if (fixed) for (int i = 0; i < nv; ++i) vfixed[i] = fixed[i];
  }

  // updatePlasticity_ particle branch (HybridSolver.cpp:616-681)
  void return_map(int p, const M3<T>& Fh) {
    M3<T> total = mul(Fh, getM(FP, p));  // :618-619
    M3<T> U, Vm;
    T S[3];
    svd3(Fh, U, S, Vm);
    T Sn[3] = {S[0], S[1], S[2]};
    if (P.material == MPM_SNOW) {
      for (int k = 0; k < 3; ++k)
        Sn[k] = ref_clamp<T>(S[k], T(1) - T(P.critical_compression), T(1) + T(P.critical_stretch));
    } else {
      T phi0 = P.per_particle_phi ? phi[p] : T(P.dp_h[0]);
      T dq;
      dp_project<T>(P, phi0, plastic[p], S, Sn, dq);
      plastic[p] += dq;
    }
    M3<T> FE, FPn;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) FE(r, c) = U(r, 0) * Sn[0] * Vm(c, 0) + U(r, 1) * Sn[1] * Vm(c, 1) + U(r, 2) * Sn[2] * Vm(c, 2);
    // F_P = V Sigma^-1 U^T (F^ F_P) (:676-677)
    M3<T> VSU;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        VSU(r, c) = Vm(r, 0) / Sn[0] * U(c, 0) + Vm(r, 1) / Sn[1] * U(c, 1) + Vm(r, 2) / Sn[2] * U(c, 2);
    FPn = mul(VSU, total);
    setM(F, p, FE);
    setM(FP, p, FPn);
    if (P.material == MPM_SNOW) plastic[p] = det(FPn);
  }

  void estimate_volumes() {
    p2g();
    std::vector<NodeW> nw;
    T gvol = h[0] * h[1] * h[2];
    for (int p = 0; p < n; ++p) {
      weights(p, nw);
      T rho = 0;
      for (const auto& q : nw) rho += T(q.w) * gm[q.idx];
      rho /= gvol;  // :246
      V[p] = m[p] / rho;  // :247
    }
  }
};

}  // namespace orc

// ---------------------------------------------------------------- C interface for the tests
namespace {
struct Handle {
  int prec;
  orc::Sim<double> d;
  orc::Sim<float> f;
};
template <class T> void load(std::vector<T>& dst, const double* src, size_t count, T fill) {
  dst.assign(count, fill);
  if (src)
    for (size_t i = 0; i < count; ++i) dst[i] = T(src[i]);
}
template <class T> void store(const std::vector<T>& src, double* dst) {
  if (dst)
    for (size_t i = 0; i < src.size(); ++i) dst[i] = double(src[i]);
}
template <class T> void set_particles(orc::Sim<T>& s, int n, const double* x, const double* v, const double* B,
                                      const double* F, const double* plastic, const double* m, const double* V,
                                      const double* phi) {
  s.n = n;
  load(s.x, x, 3 * (size_t)n, T(0));
  load(s.v, v, 3 * (size_t)n, T(0));
  load(s.B, B, 9 * (size_t)n, T(0));
  load(s.F, F, 9 * (size_t)n, T(0));
  load(s.plastic, plastic, n, T(s.P.material == MPM_SNOW ? 1 : 0));
  load(s.m, m, n, T(0));
  load(s.V, V, n, T(0));
  load(s.phi, phi, n, T(0));
  if (!F)
    for (int p = 0; p < n; ++p)
      for (int a = 0; a < 3; ++a) s.F[(size_t)(4 * a) * n + p] = T(1);
  // F_P with det F_P = J_P (isotropic); only its determinant is ever observed (:283)
  s.FP.assign(9 * (size_t)n, T(0));
  for (int p = 0; p < n; ++p) {
    T jp = s.P.material == MPM_SNOW ? s.plastic[p] : T(1);
    T c = std::cbrt(jp);
    for (int a = 0; a < 3; ++a) s.FP[(size_t)(4 * a) * n + p] = c;
  }
}
// HybridSolver::solve (HybridSolver.cpp:854-990) with the adaptive Dt: first P2G, Dt = CFL /
// max(300, vmax/h_min) (:860); per iteration forces and velocity update with the current Dt, new Dt
// from the updated velocities (:878), frame clamp to 1/60 s (:880-892), collision, G2P with the new
// Dt, then the next P2G. dts[i] is the Dt of substep i, frame_end[i] is 1 where substep i completed
// a frame. Returns the number of substeps.
template <class T>
int solve(orc::Sim<T>& s, double cfl, double maxt, int max_sub, double* dts, int* frame_end, double* t_out) {
  const double hmin = std::min(s.h[0], std::min(s.h[1], s.h[2]));
  s.p2g();
  double Dt = cfl / std::max(3e2, s.max_speed() / hmin);
  double t = 0.0, inner = 0.0;
  int k = 0;
  while (t <= maxt && k < max_sub) {
    s.forces(T(Dt));
    s.velocity_update(T(Dt));
    Dt = cfl / std::max(3e2, s.max_speed() / hmin);
    int fe = 0;
    if (inner + Dt >= 1.0 / 60.0) {
      Dt = 1.0 / 60.0 - inner;
      t += 1.0 / 60.0;
      inner = 0.0;
      fe = 1;
    } else {
      inner += Dt;
    }
    s.collide();
    s.g2p(T(Dt));
    if (dts) dts[k] = Dt;
    if (frame_end) frame_end[k] = fe;
    ++k;
    s.p2g();
  }
  if (t_out) *t_out = t;
  return k;
}

template <class T> int stage(orc::Sim<T>& s, int st, double dt) {
  switch (st) {
    case MPM_STAGE_P2G: s.p2g(); s.forces(T(dt)); return 0;
    case MPM_STAGE_GRID: s.grid_update(T(dt)); return 0;
    case MPM_STAGE_G2P: s.g2p(T(dt)); return 0;
    default: return 0;
  }
}
}  // namespace

namespace {
template <class T> static void get_mesh(const orc::Sim<T>& s, double* vx, double* vv, double* D) {
  for (size_t i = 0; i < 3 * (size_t)s.nv; ++i) {
    if (vx) vx[i] = double(s.vx[i]);
    if (vv) vv[i] = double(s.vv[i]);
  }
  if (D)
    for (size_t i = 0; i < 9 * (size_t)s.ne; ++i) D[i] = double(s.eD[i]);
}
}  // namespace

extern "C" {

void* orc_create(int precision, const mpm_params* P, const double mn[3], const double mx[3], const int32_t res[3],
                 int reference_forms) {
  Handle* h = new Handle();
  h->prec = precision;
  h->d.P = *P;
  h->f.P = *P;
  h->d.reference_forms = h->f.reference_forms = reference_forms != 0;
  int r[3] = {res[0], res[1], res[2]};
  h->d.set_grid(mn, mx, r);
  h->f.set_grid(mn, mx, r);
  return h;
}
void orc_destroy(void* hp) { delete static_cast<Handle*>(hp); }

int orc_set_levelset(void* hp, const mpm_levelset* prims, int n) {
  Handle* h = static_cast<Handle*>(hp);
  h->d.ls.assign(prims, prims + n);
  h->f.ls.assign(prims, prims + n);
  return 0;
}

int orc_set_particles(void* hp, int n, const double* x, const double* v, const double* B, const double* F,
                      const double* plastic, const double* m, const double* V, const double* phi) {
  Handle* h = static_cast<Handle*>(hp);
  if (h->prec == 64) set_particles(h->d, n, x, v, B, F, plastic, m, V, phi);
  else set_particles(h->f, n, x, v, B, F, plastic, m, V, phi);
  return 0;
}

int orc_get_particles(void* hp, double* x, double* v, double* B, double* F, double* plastic, double* V) {
  Handle* h = static_cast<Handle*>(hp);
  auto go = [&](auto& s) {
    store(s.x, x); store(s.v, v); store(s.B, B); store(s.F, F); store(s.plastic, plastic); store(s.V, V);
  };
  if (h->prec == 64) go(h->d);
  else go(h->f);
  return 0;
}

int orc_stage(void* hp, int st, double dt) {
  Handle* h = static_cast<Handle*>(hp);
  return h->prec == 64 ? stage(h->d, st, dt) : stage(h->f, st, dt);
}

int orc_substep(void* hp, double dt) {
  for (int st = MPM_STAGE_P2G; st <= MPM_STAGE_G2P; ++st) orc_stage(hp, st, dt);
  return 0;
}

// Cloth of the oracle (LagrangianMesh::ObjMesh from given vertices and faces).
int orc_set_mesh(void* hp, int nv, const double* x, const double* v, int nf, const int* faces, double density,
                 double thickness, double E, double nu, double gamma, double k, double angle_deg, const uint8_t* fixed,
                 int ref_normal) {
  Handle* h = static_cast<Handle*>(hp);
  h->d.set_mesh(nv, x, v, nf, faces, density, thickness, E, nu, gamma, k, angle_deg, fixed, ref_normal);
  h->f.set_mesh(nv, x, v, nf, faces, density, thickness, E, nu, gamma, k, angle_deg, fixed, ref_normal);
  return 0;
}
// One cloth element in FP64: vertex forces f[9] and element stress M[9] (row-major) from D and the
// rest directions Dr (both row-major, columns d1 d2 d3).
void orc_cloth_element(const double* D, const double* Dr, double vol, double E, double nu, double gamma, double k,
                       int ref_normal, double* f, double* M) {
  orc::M3<double> A, B, Ms;
  for (int i = 0; i < 9; ++i) { A.a[i] = D[i]; B.a[i] = Dr[i]; }
  const double lambda = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu), mu = E / 2.0 / (1.0 + nu);
  orc::cloth_element<double>(A, B, vol, mu, lambda, gamma, k, ref_normal != 0, f, Ms);
  for (int i = 0; i < 9; ++i) M[i] = Ms.a[i];
}
void orc_cloth_return_map(const double* D, double gamma, double k, double angle_deg, double* out) {
  orc::M3<double> A;
  for (int i = 0; i < 9; ++i) A.a[i] = D[i];
  orc::M3<double> B = orc::cloth_return_map<double>(A, gamma, k, std::tan(angle_deg * 3.14159265358979323846 / 180.0));
  for (int i = 0; i < 9; ++i) out[i] = B.a[i];
}
int orc_get_mesh(void* hp, double* vx, double* vv, double* D) {
  Handle* h = static_cast<Handle*>(hp);
  if (h->prec == 64) get_mesh(h->d, vx, vv, D);
  else get_mesh(h->f, vx, vv, D);
  return 0;
}
int orc_solve(void* hp, double cfl, double maxt, int max_substeps, double* dts, int* frame_end, double* t_out) {
  Handle* h = static_cast<Handle*>(hp);
  return h->prec == 64 ? solve(h->d, cfl, maxt, max_substeps, dts, frame_end, t_out)
                       : solve(h->f, cfl, maxt, max_substeps, dts, frame_end, t_out);
}
int orc_run(void* hp, double dt, int nsteps) {
  for (int i = 0; i < nsteps; ++i) orc_substep(hp, dt);
  return 0;
}

// The particle force pass alone, on grid velocities chosen by the caller: gv (3 x ng planar) replaces
// the grid velocity that the P2G would leave (the predicted F^ is built from it), then forces(dt)
// runs and f (3 x ng planar) receives the grid forces without gravity.
int orc_forces_at(void* hp, const double* gv, double dt, double* f) {
  Handle* h = static_cast<Handle*>(hp);
  auto go = [&](auto& s) {
    using T = typename std::decay_t<decltype(s.gv)>::value_type;
    for (size_t i = 0; i < s.gv.size(); ++i) s.gv[i] = T(gv[i]);
    s.forces(T(dt));
    store(s.gf, f);
  };
  if (h->prec == 64) go(h->d);
  else go(h->f);
  return 0;
}

// The nodes stopped by the fixed cloth vertices: every node's v and v_BF are set to 1, the
// fixed-vertex pass runs, and mask[node] = 1 where it zeroed them.
int orc_fixed_nodes(void* hp, uint8_t* mask) {
  Handle* h = static_cast<Handle*>(hp);
  auto go = [&](auto& s) {
    using T = typename std::decay_t<decltype(s.gv)>::value_type;
    std::fill(s.gv.begin(), s.gv.end(), T(1));
    std::fill(s.gvbf.begin(), s.gvbf.end(), T(1));
    s.fix_vertices();
    for (int c = 0; c < s.ng; ++c) mask[c] = (s.gv[c] == T(0) && s.gvbf[c] == T(0)) ? 1 : 0;
  };
  if (h->prec == 64) go(h->d);
  else go(h->f);
  return 0;
}

int orc_estimate_volumes(void* hp) {
  Handle* h = static_cast<Handle*>(hp);
  if (h->prec == 64) h->d.estimate_volumes();
  else h->f.estimate_volumes();
  return 0;
}

// m[ng], p[3ng], f[3ng] (stress part), v[3ng], vbf[3ng]
int orc_get_grid(void* hp, double* m, double* p, double* f, double* v, double* vbf) {
  Handle* h = static_cast<Handle*>(hp);
  auto go = [&](auto& s) { store(s.gm, m); store(s.gp, p); store(s.gf, f); store(s.gv, v); store(s.gvbf, vbf); };
  if (h->prec == 64) go(h->d);
  else go(h->f);
  return 0;
}

// Cell keys as HybridSolver.cpp:33-36 (truncating cast) and RegularGrid::toIndex, clamped to the grid.
void orc_cell_keys(int n, const double* x, const double mn[3], const double h[3], const int32_t res[3], uint32_t* keys) {
  for (int p = 0; p < n; ++p) {
    int c[3];
    for (int a = 0; a < 3; ++a) {
      c[a] = static_cast<int>((x[(size_t)a * n + p] - mn[a]) / h[a]);
      c[a] = std::max(0, std::min(res[a] - 1, c[a]));
    }
    keys[p] = uint32_t(c[2] * res[0] * res[1] + c[1] * res[0] + c[0]);
  }
}

// Stable sort by key (the ordering implied by setFromTriplets: ascending particle index within a cell).
void orc_stable_sort(int n, const uint32_t* keys, int32_t* perm, int32_t ncells, int32_t* bins) {
  std::vector<int32_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return keys[a] < keys[b]; });
  if (perm) std::copy(idx.begin(), idx.end(), perm);
  if (bins) {
    std::vector<int32_t> cnt(ncells + 1, 0);
    for (int p = 0; p < n; ++p) cnt[keys[p] + 1]++;
    for (int c = 0; c < ncells; ++c) cnt[c + 1] += cnt[c];
    std::copy(cnt.begin(), cnt.end(), bins);
  }
}

void orc_svd3(const double* A, double* U, double* S, double* V) {
  orc::M3<double> a, u, vv;
  std::copy(A, A + 9, a.a);
  orc::svd3(a, u, S, vv);
  std::copy(u.a, u.a + 9, U);
  std::copy(vv.a, vv.a + 9, V);
}
void orc_qr3(const double* D, double* Q, double* R) {
  orc::M3<double> d, q, r;
  std::copy(D, D + 9, d.a);
  orc::gram_schmidt_qr(d, q, r);
  std::copy(q.a, q.a + 9, Q);
  std::copy(r.a, r.a + 9, R);
}
void orc_spline(int order, int n, const double* xs, double* w, double* dw) {
  for (int i = 0; i < n; ++i) {
    w[i] = orc::spline_N(order, xs[i]);
    dw[i] = orc::spline_dN(order, xs[i]);
  }
}
// Piola stress of one F (row-major) for material params P (snow uses Jp).
void orc_piola(const mpm_params* P, const double* Fh, double Jp, double* out) {
  orc::M3<double> f;
  std::copy(Fh, Fh + 9, f.a);
  orc::M3<double> r = orc::piola<double>(*P, f, Jp);
  std::copy(r.a, r.a + 9, out);
}
// Drucker-Prager projection of principal stretches.
#ifdef _OPENMP
// threads of the OpenMP build (the all-cores CPU line)
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
#endif
void orc_dp_project(const mpm_params* P, double phi0, double q, const double* Sin, double* Sout, double* dq) {
  orc::dp_project<double>(*P, phi0, q, Sin, Sout, *dq);
}
void orc_levelset(const mpm_levelset* prims, int nprims, const double* x, double* phi, double* n) {
  std::vector<mpm_levelset> v(prims, prims + nprims);
  orc::levelset<double>(v, x, *phi, n);
}

}  // extern "C"
