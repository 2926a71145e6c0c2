// mpm_oracle.hpp -- CPU restatement of the reference MLS-MPM substep.
//
// TEST INFRASTRUCTURE ONLY. This is the checker the parity tests, smoke() and
// the bench's CPU-baseline leg compare the HIP engine against; the product
// (libmpm_hip.so, mpm_amd/) never includes, links or calls it.
//
// Templated on the scalar type S:
//   * S = float  restates the reference's GPU arithmetic (kekeblom/mpm,
//     src/mpm.cu:14-178) operation by operation: FP32 everywhere the reference
//     computes in float, FP64 where it computes in double (the exp and the
//     volumetric scalar of MMSnow, MaterialModel.cuh:88,92), no FMA
//     contraction (built with -ffp-contract=off), and a correctly rounded
//     reciprocal square root standing in for the one svd3_cuda.h:119 calls.
//   * S = double is the FP64 oracle of the same step: every operation in
//     double, with a converged Jacobi SVD instead of the 4-sweep
//     approximate one.
//
// Particle scatter order: serial in particle order unless built with OpenMP
// and called with threads > 1, in which case P2G uses `omp atomic`, the
// reference's own OpenMP scatter idiom (include/mesh_builder.h:57-83).
//
// Parity status of each function (what pins it, see tests/):
//   rsqrt_rn<float>     exhaustive-style checks against exact arithmetic.
//   svd3 / polar        numpy.linalg.svd, reconstruction, SO(3) checks and the
//                       reference KATs (tests/test_linalg.cu:25-91).
//   weights             partition of unity and moment identities.
//   p2g                 mass / momentum conservation, brute-force FP64 sums.
//   grid_update         closed-form free fall and wall rules.
//   g2p                 affine-field reproduction (APIC), rigid motion.
//   material stress     zero stress on rotations, closed form on diag(F).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/mpm_hip.h"

namespace oracle {

// ---------------------------------------------------------------------------
// Correctly rounded reciprocal square root (FP32) -- the reference calls the
// CUDA intrinsic __frsqrt_rn, whose result is the float nearest to x^-1/2.
// Computed here from a double estimate and fixed up with exact float128
// comparisons of (midpoint^2 * x) against 1, so the result is exactly
// round-to-nearest-even of the real value.
// ---------------------------------------------------------------------------
inline float rsqrt_rn(float x) {
  if (!(x > 0.0f)) {
    if (x == 0.0f) return std::copysign(INFINITY, x);
    return NAN;  // negative or NaN
  }
  if (std::isinf(x)) return 0.0f;
  float f = (float)(1.0 / std::sqrt((double)x));
  if (!std::isfinite(f) || f == 0.0f) return f;
  // up to two corrections are ever needed; loop defensively
  for (int it = 0; it < 4; ++it) {
    float up = std::nextafter(f, INFINITY);
    float dn = std::nextafter(f, 0.0f);
    __float128 X = (__float128)x;
    __float128 mhi = ((__float128)f + (__float128)up) / 2;
    __float128 mlo = ((__float128)f + (__float128)dn) / 2;
    __float128 thi = mhi * mhi * X;  // < 1  <=> x^-1/2 > mhi
    __float128 tlo = mlo * mlo * X;  // > 1  <=> x^-1/2 < mlo
    if (thi < 1) { f = up; continue; }
    if (tlo > 1) { f = dn; continue; }
    if (thi == 1) {  // exact tie with the upper midpoint: round to even
      uint32_t b; std::memcpy(&b, &f, 4);
      if (b & 1u) f = up;
    } else if (tlo == 1) {
      uint32_t b; std::memcpy(&b, &f, 4);
      if (b & 1u) f = dn;
    }
    break;
  }
  return f;
}
inline double rsqrt_rn(double x) { return 1.0 / std::sqrt(x); }

template <class S> inline S rsqrt(S x) { return rsqrt_rn(x); }

// ---------------------------------------------------------------------------
// 3x3 SVD, the McAdams-Selle-Tamstorf-Teran-Sifakis method (UW-Madison TR1690,
// 2011), which the reference vendors as include/svd3_cuda.h:35-1043:
//   1. S = A^T A (svd3_cuda.h:63-97)
//   2. Jacobi eigen-sweeps on S with quaternion-accumulated rotations; each
//      sweep visits the pairs (1,2), (2,3), (3,1) (:104-411). FP32: 4 sweeps
//      of approximate Givens rotations whose half-angle is taken from
//      (s_pp - s_qq, s_qp / 2) and replaced by pi/8 when that is too steep.
//   3. quaternion normalisation (rsqrt + one Newton step) and V (:417-468)
//   4. B = A V, columns sorted by decreasing norm with sign fixes that keep V
//      a rotation (:474-707)
//   5. Givens QR of B -> U, sigma (:713-1028).
// Output: A = U diag(sigma) V^T, U and V proper rotations, |sigma0| >= |sigma1|
// >= |sigma2|, sigma0, sigma1 >= 0, sigma2 carrying the sign of det A.
// Matrices here are row-major m[r][c].
// ---------------------------------------------------------------------------
namespace svdc {
// Constants of the published method. sin(pi/8) is the correctly rounded float;
// the cos(pi/8) entry is one ulp above the correctly rounded float and is kept
// at that value (0x3F6C835F) because the algorithm's table uses it.
inline float sin_pi8_f() { uint32_t b = 1053028117u; float f; std::memcpy(&f, &b, 4); return f; }
inline float cos_pi8_f() { uint32_t b = 1064076127u; float f; std::memcpy(&f, &b, 4); return f; }
constexpr float four_gamma_sq_f = 5.8284273147583007813f;  // float(3 + 2 sqrt 2)
constexpr float tiny_f = 1.e-20f;
constexpr float small_f = 1.e-12f;
}  // namespace svdc

// The quaternion q = (qs; qv[0..2]) accumulates V.
// One Jacobi conjugation on the (p, q) plane of the symmetric matrix; r is the
// third index. s_pp, s_qq, s_rr are diagonal entries, s_qp, s_rp, s_rq the
// off-diagonal ones. qp, qq, qr are the quaternion's vector components along
// axes p, q, r.
template <class S>
inline void jacobi_conjugate_approx(S& s_pp, S& s_qp, S& s_qq, S& s_rp, S& s_rq, S& s_rr,
                                    S& qs, S& qp, S& qq, S& qr) {
  const S half = s_qp * S(0.5);
  const S diff = s_pp - s_qq;
  const S half2 = half * half;
  const bool exact = half2 >= S(svdc::tiny_f);
  S sh = exact ? half : S(0);
  S ch = exact ? diff : S(1);
  S sh2 = sh * sh;
  S ch2 = ch * ch;
  const S w = rsqrt<S>(sh2 + ch2);
  sh = w * sh;
  ch = w * ch;
  if (ch2 <= S(svdc::four_gamma_sq_f) * sh2) {  // rotation too steep: use pi/8
    sh = S(svdc::sin_pi8_f());
    ch = S(svdc::cos_pi8_f());
  }
  sh2 = sh * sh;
  ch2 = ch * ch;
  const S c = ch2 - sh2;
  S s = ch * sh;
  s = s + s;
  // conjugation S' = Q^T S Q with the unnormalised quaternion scale folded in
  const S scale = sh2 + ch2;
  s_rr = (s_rr * scale) * scale;
  s_rp = s_rp * scale;
  s_rq = s_rq * scale;
  const S t_rp = s * s_rp;
  const S t_rq = s * s_rq;
  s_rp = c * s_rp + t_rq;
  s_rq = c * s_rq - t_rp;
  const S s2 = s * s;
  const S c2 = c * c;
  const S a = s_qq * s2;
  const S b = s_pp * s2;
  s_pp = s_pp * c2 + a;
  s_qq = s_qq * c2 + b;
  const S c2ms2 = c2 - s2;
  S twice = s_qp + s_qp;
  s_qp = s_qp * c2ms2;
  const S cs = c * s;
  twice = twice * cs;
  const S dcs = diff * cs;
  s_pp = s_pp + twice;
  s_qp = s_qp - dcs;
  s_qq = s_qq - twice;
  // accumulate the rotation about axis r into the quaternion
  const S t_p = sh * qp, t_q = sh * qq, t_r = sh * qr, t_s = sh * qs;
  qs = ch * qs; qp = ch * qp; qq = ch * qq; qr = ch * qr;
  qr = qr + t_s;
  qs = qs - t_r;
  qp = qp + t_q;
  qq = qq - t_p;
}

// FP64 variant: the exact Jacobi rotation that zeroes s_qp, with the smaller
// of the two admissible angles (|theta| <= pi/4, the classical choice that
// makes the sweeps converge quadratically): cot 2theta = (s_pp - s_qq) / (2 s_qp),
// t = tan theta = sgn(tau) / (|tau| + sqrt(1 + tau^2)). Same conjugation and
// quaternion bookkeeping as the FP32 variant.
inline void jacobi_conjugate_exact(double& s_pp, double& s_qp, double& s_qq, double& s_rp,
                                   double& s_rq, double& s_rr, double& qs, double& qp,
                                   double& qq, double& qr) {
  (void)s_rr;
  if (s_qp == 0.0) return;
  const double diff = s_pp - s_qq;
  const double tau = diff / (2.0 * s_qp);
  const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
  const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
  const double ch = std::sqrt(0.5 * (1.0 + c));  // cos(theta / 2)
  const double sh = s / (2.0 * ch);               // sin(theta / 2)
  const double t_rp = s * s_rp, t_rq = s * s_rq;
  s_rp = c * s_rp + t_rq;
  s_rq = c * s_rq - t_rp;
  const double s2 = s * s, c2 = c * c, cs = c * s;
  const double npp = s_pp * c2 + s_qq * s2 + 2.0 * s_qp * cs;
  const double nqq = s_qq * c2 + s_pp * s2 - 2.0 * s_qp * cs;
  const double nqp = s_qp * (c2 - s2) - diff * cs;
  s_pp = npp; s_qq = nqq; s_qp = nqp;
  const double t_p = sh * qp, t_q = sh * qq, t_r = sh * qr, t_s = sh * qs;
  qs = ch * qs; qp = ch * qp; qq = ch * qq; qr = ch * qr;
  qr = qr + t_s;
  qs = qs - t_r;
  qp = qp + t_q;
  qq = qq - t_p;
}

// rsqrt followed by one Newton step, r' = (r + r/2) - n2 r^3 / 2
template <class S> inline S rsqrt_newton(S n2) {
  S r = rsqrt<S>(n2);
  const S h = r * S(0.5);
  S t = r * h;
  t = r * t;
  t = n2 * t;
  r = r + h;
  r = r - t;
  return r;
}

// Givens rotation for the QR step that zeroes b_low against b_diag.
// Thresholds of the QR step: the method's FP32 constant, and a double one far
// below any representable product of FP32 data for the FP64 oracle.
template <class S> inline S qr_small() { return S(svdc::small_f); }
template <> inline double qr_small<double>() { return 1e-90; }  // rsqrt(small^2)^3 stays finite

template <class S> inline void qr_givens(S b_diag, S b_low, S& c, S& s) {
  S sh = (b_low * b_low >= qr_small<S>()) ? b_low : S(0);
  S ch = std::max(std::max(S(0) - b_diag, b_diag), qr_small<S>());
  const bool pos = b_diag >= S(0);
  S n2 = ch * ch + sh * sh;
  S r = rsqrt_newton<S>(n2);
  r = r * n2;  // ~ sqrt(n2)
  ch = ch + r;
  if (!pos) std::swap(ch, sh);
  n2 = ch * ch + sh * sh;
  r = rsqrt_newton<S>(n2);
  ch = ch * r;
  sh = sh * r;
  c = ch * ch - sh * sh;
  s = sh * ch;
  s = s + s;
}

template <class S>
inline void svd3(const S A[3][3], S U[3][3], S sigma[3], S V[3][3]) {
  // 1. normal equations S = A^T A, summed over rows 0,1,2 (svd3_cuda.h:63-97)
  auto dot_cols = [&](int a, int b) {
    S acc = A[0][a] * A[0][b];
    acc = A[1][a] * A[1][b] + acc;
    acc = A[2][a] * A[2][b] + acc;
    return acc;
  };
  S s11 = dot_cols(0, 0), s21 = dot_cols(1, 0), s31 = dot_cols(2, 0);
  S s22 = dot_cols(1, 1), s32 = dot_cols(2, 1), s33 = dot_cols(2, 2);
  S qs = 1, qx = 0, qy = 0, qz = 0;
  // 2. Jacobi sweeps: (p,q,r) = (1,2,3), (2,3,1), (3,1,2)
  if constexpr (std::is_same<S, float>::value) {
    for (int sweep = 0; sweep < 4; ++sweep) {
      jacobi_conjugate_approx<S>(s11, s21, s22, s31, s32, s33, qs, qx, qy, qz);
      jacobi_conjugate_approx<S>(s22, s32, s33, s21, s31, s11, qs, qy, qz, qx);
      jacobi_conjugate_approx<S>(s33, s31, s11, s32, s21, s22, qs, qz, qx, qy);
    }
  } else {
    for (int sweep = 0; sweep < 40; ++sweep) {
      const double off = std::fabs(s21) + std::fabs(s32) + std::fabs(s31);
      const double dia = std::fabs(s11) + std::fabs(s22) + std::fabs(s33);
      if (off <= 1e-17 * dia || off == 0.0) break;
      jacobi_conjugate_exact(s11, s21, s22, s31, s32, s33, qs, qx, qy, qz);
      jacobi_conjugate_exact(s22, s32, s33, s21, s31, s11, qs, qy, qz, qx);
      jacobi_conjugate_exact(s33, s31, s11, s32, s21, s22, qs, qz, qx, qy);
    }
  }
  // 3. normalise the quaternion, then V = R(q) (svd3_cuda.h:417-468)
  S n2 = qs * qs;
  n2 = qx * qx + n2;
  n2 = qy * qy + n2;
  n2 = qz * qz + n2;
  const S rn = rsqrt_newton<S>(n2);
  qs = qs * rn; qx = qx * rn; qy = qy * rn; qz = qz * rn;
  {
    const S xx = qx * qx, yy = qy * qy, zz = qz * qz, ss = qs * qs;
    V[0][0] = ((ss + xx) - yy) - zz;
    V[1][1] = ((ss - xx) + yy) - zz;
    V[2][2] = ((ss - xx) - yy) + zz;
    const S x2 = qx + qx, y2 = qy + qy, z2 = qz + qz;
    const S sx = qs * x2, sy = qs * y2, sz = qs * z2;
    const S xy = qy * x2, yz = qz * y2, zx = qx * z2;
    V[0][1] = xy - sz;
    V[1][0] = xy + sz;
    V[1][2] = yz - sx;
    V[2][1] = yz + sx;
    V[2][0] = zx - sy;
    V[0][2] = zx + sy;
  }
  // 4. B = A V (each entry summed over k = 0,1,2 in that order)
  S B[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      S acc = V[0][c] * A[r][0];
      acc = acc + V[1][c] * A[r][1];
      acc = acc + V[2][c] * A[r][2];
      B[r][c] = acc;
    }
  S rho[3];
  for (int c = 0; c < 3; ++c) {
    S acc = B[0][c] * B[0][c];
    acc = acc + B[1][c] * B[1][c];
    acc = acc + B[2][c] * B[2][c];
    rho[c] = acc;
  }
  // sort columns by decreasing norm; each swap negates one of the two columns
  // so that V stays a proper rotation (svd3_cuda.h:532-707)
  auto swap_cols = [&](int a, int b, int negate) {
    if (rho[a] < rho[b]) {
      for (int r = 0; r < 3; ++r) {
        std::swap(B[r][a], B[r][b]);
        std::swap(V[r][a], V[r][b]);
      }
      std::swap(rho[a], rho[b]);
      for (int r = 0; r < 3; ++r) {
        B[r][negate] = -B[r][negate];
        V[r][negate] = -V[r][negate];
      }
    }
  };
  swap_cols(0, 1, 1);
  swap_cols(0, 2, 0);
  swap_cols(1, 2, 2);
  // 5. QR of B by three Givens rotations: (row0,row1), (row0,row2), (row1,row2)
  S Um[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  auto rotate = [&](int a, int b) {
    S c, s;
    qr_givens<S>(B[a][a], B[b][a], c, s);
    for (int col = 0; col < 3; ++col) {
      const S ta = s * B[a][col], tb = s * B[b][col];
      B[a][col] = c * B[a][col] + tb;
      B[b][col] = c * B[b][col] - ta;
    }
    for (int row = 0; row < 3; ++row) {
      const S ta = s * Um[row][a], tb = s * Um[row][b];
      Um[row][a] = c * Um[row][a] + tb;
      Um[row][b] = c * Um[row][b] - ta;
    }
  };
  rotate(0, 1);
  rotate(0, 2);
  rotate(1, 2);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) U[r][c] = Um[r][c];
  sigma[0] = B[0][0];
  sigma[1] = B[1][1];
  sigma[2] = B[2][2];
}

// ---------------------------------------------------------------------------
// Small 3x3 helpers, in the reference's operation order.
// ---------------------------------------------------------------------------
// Three-term inner product in the order Eigen 3.3's unrolled reduction uses
// for the reference's fixed-size products (redux splits 3 terms as 1 + 2):
// p0 + (p1 + p2).
template <class S> inline S dot3(S a0, S b0, S a1, S b1, S a2, S b2) {
  const S p0 = a0 * b0, p1 = a1 * b1, p2 = a2 * b2;
  return p0 + (p1 + p2);
}
template <class S> inline void matmul(const S a[3][3], const S b[3][3], S out[3][3]) {
  S t[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      t[r][c] = dot3<S>(a[r][0], b[0][c], a[r][1], b[1][c], a[r][2], b[2][c]);
  std::memcpy(out, t, sizeof(t));
}
// a * b^T
template <class S> inline void matmul_bt(const S a[3][3], const S b[3][3], S out[3][3]) {
  S t[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      t[r][c] = dot3<S>(a[r][0], b[c][0], a[r][1], b[c][1], a[r][2], b[c][2]);
  std::memcpy(out, t, sizeof(t));
}

// linalg::determinant (src/linalg.cu:48-53): expansion along row 0.
template <class S> inline S det3(const S m[3][3]) {
  const S sub1 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  const S sub2 = m[1][0] * m[2][2] - m[1][2] * m[2][0];
  const S sub3 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
  return m[0][0] * sub3 - m[0][1] * sub2 + m[0][2] * sub1;
}
// Eigen's 3x3 determinant (cofactors of column 0), used by MMJelly
// (MaterialModel.cuh:145,147 call F.determinant()).
template <class S> inline S det3_col0(const S m[3][3]) {
  return m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
         m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
         m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
}

// polar decomposition through the SVD (linalg.cu:18-33): R = U V^T
template <class S> inline void polar_R(const S F[3][3], S R[3][3]) {
  S U[3][3], sig[3], V[3][3];
  svd3<S>(F, U, sig, V);
  matmul_bt<S>(U, V, R);
}

template <class S> inline S clampv(S v, S lo, S hi) { return std::max(std::min(v, hi), lo); }

// ---------------------------------------------------------------------------
// Particle and step.
// ---------------------------------------------------------------------------
template <class S> struct Particle {
  S x[3], v[3];
  S C[3][3], F[3][3];
  S Jp;
  uint8_t mat;
};

// Interpolation kernels (the InterpolationKernel plugin, InterpolationKernel.cuh):
// the reference's quadratic B-spline with its constant D^-1, and a cubic
// B-spline whose D^-1 goes through the generic per-particle inverse.
enum Interp : int { INTERP_QUADRATIC = 0, INTERP_CUBIC = 1 };

template <class S> struct StepConsts {
  S dt, dx, dx_inv, dinv, gravity, bfrac;
  uint32_t N;
  int boundary_mode;
  int compat;
  int bcells;
  int interp;  // Interp
};

template <class S> inline StepConsts<S> make_consts(const mpm_params& p, int interp = INTERP_QUADRATIC) {
  StepConsts<S> k;
  k.interp = interp;
  k.N = p.N;
  k.dt = S(p.dt);
  // TransferScheme.h:26-27: dx = 1.0/N (double) stored as real; dx_inv = 1.0/dx
  k.dx = S(1.0 / double(p.N));
  k.dx_inv = S(1.0 / double(k.dx));
  // QuadraticInterpolationKernel::D_inv_const (InterpolationKernel.cuh:71-73):
  // Identity * 4 * dx_inv * dx_inv
  k.dinv = (S(4) * k.dx_inv) * k.dx_inv;
  k.gravity = S(p.gravity);
  k.boundary_mode = p.boundary_mode;
  k.bfrac = S(p.boundary);
  k.bcells = (int)p.boundary;
  k.compat = p.compat;
  return k;
}

// QuadraticInterpolationKernel::weights_per_direction (InterpolationKernel.cuh:57-69).
// base = int(x/dx - 0.5) truncates toward zero.
template <class S> inline void quad_weights(const S x[3], S dx_inv, int base[3], S w[3][3]) {
  for (int a = 0; a < 3; ++a) {
    const S xg = x[a] * dx_inv;
    base[a] = (int)(xg - S(0.5));
    const S fx = xg - S(base[a]);
    const S t0 = S(1.5) - fx;
    const S t1 = fx - S(1.0);
    const S t2 = fx - S(0.5);
    w[a][0] = S(0.5) * (t0 * t0);
    w[a][1] = S(0.75) - t1 * t1;
    w[a][2] = S(0.5) * (t2 * t2);
  }
}

// Cubic B-spline N(t), t = |node - particle| / dx: (1/2) t^3 - t^2 + 2/3 on
// [0, 1), (2 - t)^3 / 6 on [1, 2), 0 beyond.
template <class S> inline S cubic_bspline(S t) {
  if (t < S(1)) {
    const S t2 = t * t;
    return ((S(0.5) * t2) * t - t2) + S(2.0 / 3.0);
  }
  if (t < S(2)) {
    const S u = S(2) - t;
    return ((u * u) * u) * S(1.0 / 6.0);
  }
  return S(0);
}

// Cubic kernel weights over 4 nodes per direction: base = int(x/dx - 1)
// (truncated toward zero, as the quadratic kernel's .cast<int>()), node o at
// distance |x/dx - base - o|.
template <class S> inline void cubic_weights(const S x[3], S dx_inv, int base[3], S w[3][4]) {
  for (int a = 0; a < 3; ++a) {
    const S xg = x[a] * dx_inv;
    base[a] = (int)(xg - S(1.0));
    const S fx = xg - S(base[a]);
    for (int o = 0; o < 4; ++o) w[a][o] = cubic_bspline<S>(std::fabs(fx - S(o)));
  }
}

inline int kernel_size(int interp) { return interp == INTERP_CUBIC ? 4 : 3; }

// InterpolationKernel::weights_per_direction of the selected kernel; w[a][o]
// for o < kernel_size.
template <class S>
inline void kernel_weights(const StepConsts<S>& k, const S x[3], int base[3], S w[3][4]) {
  if (k.interp == INTERP_CUBIC) {
    cubic_weights<S>(x, k.dx_inv, base, w);
    return;
  }
  S q[3][3];
  quad_weights<S>(x, k.dx_inv, base, q);
  for (int a = 0; a < 3; ++a) {
    for (int o = 0; o < 3; ++o) w[a][o] = q[a][o];
    w[a][3] = S(0);
  }
}

// Inverse of a 3x3 matrix by its adjugate: cofactors of column 0 give the
// determinant, every entry is cofactor / det.
template <class S> inline void inverse3(const S m[3][3], S out[3][3]) {
  S cof[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      cof[r][c] = m[r1][c1] * m[r2][c2] - m[r1][c2] * m[r2][c1];
    }
  const S det = dot3<S>(cof[0][0], m[0][0], cof[1][0], m[1][0], cof[2][0], m[2][0]);
  const S inv = S(1) / det;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) out[r][c] = cof[c][r] * inv;
}

// InterpolationKernelBase::D_inv (InterpolationKernel.cuh:21-50): D = sum over
// the stencil of w_i w_j w_k (x_node - x_p)(x_node - x_p)^T, inverted. The
// reference evaluates it when the kernel's d_is_const() is false.
template <class S>
inline void generic_dinv(const S x[3], const int base[3], const S w[3][4], int ns, S dx, S Dinv[3][3]) {
  S D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  S d[3];
  for (int i = 0; i < ns; ++i) {
    d[0] = S(base[0] + i) * dx - x[0];
    for (int j = 0; j < ns; ++j) {
      d[1] = S(base[1] + j) * dx - x[1];
      for (int kk = 0; kk < ns; ++kk) {
        d[2] = S(base[2] + kk) * dx - x[2];
        const S wt = (w[0][i] * w[1][j]) * w[2][kk];
        for (int r = 0; r < 3; ++r) {
          const S wd = wt * d[r];
          for (int c = 0; c < 3; ++c) D[r][c] = D[r][c] + wd * d[c];
        }
      }
    }
  }
  inverse3<S>(D, Dinv);
}

// The Kirchhoff stress the MaterialModel plugin returns (computePF).
// MMSnow: MaterialModel.cuh:85-93; MMFixedCorotated: :56-61; MMJelly: :131-139.
// The reference evaluates exp(hardening (1 - Jp)) and lambda ((J - 1) J) in
// double and rounds to float; that is kept for S = float.
// `lame` (mu, lambda), when given, replaces the material's mu0, lambda0 for
// this particle (the differentiable step's per-particle parameters).
template <class S>
inline void material_stress(const Particle<S>& p, const mpm_material& m, int compat,
                            const S R[3][3], S PF[3][3], const S* lame = nullptr) {
  if (m.kind == MPM_MATERIAL_WC_LIQUID) {
    const S J = p.Jp;
    const S pr = S(double(S(m.bulk_modulus)) * ((double(J) - 1.0) * double(J)));
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) PF[r][c] = (r == c) ? pr : S(0);
    return;
  }
  const S mu0 = lame ? lame[0] : S(m.mu0), lambda0 = lame ? lame[1] : S(m.lambda0);
  S mu = mu0, lambda = lambda0;
  if (m.kind != MPM_MATERIAL_FIXED_COROTATED) {
    const S e = S(std::exp(double(S(m.hardening)) * (1.0 - double(p.Jp))));
    mu = mu0 * e;
    lambda = lambda0 * e;
  }
  // volumetric J: the reference uses the plastic Jp; compat FIXED uses det F
  const S J = (compat == MPM_COMPAT_FIXED) ? det3<S>(p.F) : p.Jp;
  S FmR[3][3];
  const S two_mu = S(2.0 * double(mu));
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) FmR[r][c] = two_mu * (p.F[r][c] - R[r][c]);
  S dev[3][3];
  matmul_bt<S>(FmR, p.F, dev);
  const S vol = S(double(lambda) * ((double(J) - 1.0) * double(J)));
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) PF[r][c] = dev[r][c] + ((r == c) ? vol : S(0));
}

// Per-particle P2G preparation (MLS_APIC_Scheme::p2g_prepare_particle,
// TransferScheme.h:66-86): affine = -Dinv dt V PF + m C.
// Dinv: the particle's D^-1 when the kernel's is not constant (nullptr: the
// quadratic kernel's constant k.dinv I).
template <class S>
inline void p2g_affine(const Particle<S>& p, const mpm_material& m, const StepConsts<S>& k,
                       S affine[3][3], const S* lame = nullptr, const S (*Dinv)[3] = nullptr) {
  S R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  if (m.kind != MPM_MATERIAL_WC_LIQUID) polar_R<S>(p.F, R);
  S PF[3][3];
  material_stress<S>(p, m, k.compat, R, PF, lame);
  if (Dinv) {
    // stress = ((-Dinv) dt V) PF, a full matrix product (TransferScheme.h:83)
    S sD[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) sD[r][c] = ((-Dinv[r][c]) * k.dt) * S(m.particle_volume);
    S st[3][3];
    matmul<S>(sD, PF, st);
    const S mass = S(m.particle_mass);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) affine[r][c] = st[r][c] + mass * p.C[r][c];
    return;
  }
  // stress = ((-Dinv) * dt) * V * PF with diagonal Dinv
  const S d = ((-k.dinv) * k.dt) * S(m.particle_volume);
  const S mass = S(m.particle_mass);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) affine[r][c] = d * PF[r][c] + mass * p.C[r][c];
}

// Grid stored as N^3 x 4 (momentum xyz, mass), index N*N*i + N*j + k.
template <class S, class Add>
inline void p2g_particle(const Particle<S>& p, const mpm_material& m, const StepConsts<S>& k,
                         Add&& add, const S* lame = nullptr) {
  const uint32_t N = k.N;
  const int ns = kernel_size(k.interp);
  int base[3];
  S w[3][4];
  kernel_weights<S>(k, p.x, base, w);
  // out-of-domain particles are skipped (mpm.cu:31-35)
  for (int a = 0; a < 3; ++a)
    if (base[a] + ns < 0 || base[a] >= (int)N) return;
  S affine[3][3];
  if (k.interp == INTERP_QUADRATIC) {
    p2g_affine<S>(p, m, k, affine, lame);
  } else {
    S Dinv[3][3];
    generic_dinv<S>(p.x, base, w, ns, k.dx, Dinv);
    p2g_affine<S>(p, m, k, affine, lame, Dinv);
  }
  const S mass = S(m.particle_mass);
  int beg[3], end[3];
  for (int a = 0; a < 3; ++a) {
    beg[a] = std::max(0, -base[a]);
    // mpm.cu:42-44 bounds the upper end by N + base (compat REFERENCE); the
    // correct bound, used by G2P (mpm.cu:140-142) and compat FIXED, is N - base
    const int lim = (k.compat == MPM_COMPAT_FIXED) ? (int)N - base[a] : (int)N + base[a];
    end[a] = std::min(ns, lim);
  }
  for (int i = beg[0]; i < end[0]; ++i) {
    const uint32_t ig = (uint32_t)(base[0] + i);
    const S d0 = S(ig) * k.dx - p.x[0];
    for (int j = beg[1]; j < end[1]; ++j) {
      const uint32_t jg = (uint32_t)(base[1] + j);
      const S d1 = S(jg) * k.dx - p.x[1];
      for (int kk = beg[2]; kk < end[2]; ++kk) {
        const uint32_t kg = (uint32_t)(base[2] + kk);
        const S d2 = S(kg) * k.dx - p.x[2];
        S ad[3];  // affine * d (TransferScheme.h:93)
        for (int r = 0; r < 3; ++r) ad[r] = dot3<S>(affine[r][0], d0, affine[r][1], d1, affine[r][2], d2);
        const S wt = (w[0][i] * w[1][j]) * w[2][kk];
        S c4[4];
        for (int r = 0; r < 3; ++r) c4[r] = wt * (p.v[r] * mass + ad[r]);
        c4[3] = wt * mass;
        // linear index as the reference forms it in u32; j or k = N, N+1 alias
        // into the next row (compat REFERENCE); beyond N^3 the reference writes
        // past its allocation, which is dropped here
        const uint64_t idx = (uint64_t)N * N * ig + (uint64_t)N * jg + kg;
        if (idx >= (uint64_t)N * N * N) continue;
        add(idx, c4);
      }
    }
  }
}

// Wall classification of node (xi, yi, zi) (mpm.cu:92-105). The decision is
// discrete, so it is always made in float as the reference makes it
// (real(index) / N against 0.05f and 1 - 0.05f), also in the FP64 oracle.
inline void node_walls(uint32_t xi, uint32_t yi, uint32_t zi, uint32_t N, int mode, float bfrac,
                       int bcells, bool& sticky, bool& floor_) {
  if (mode == MPM_BOUNDARY_CELLS) {
    const int b = bcells, n = (int)N;
    sticky = (int)xi < b || (int)xi >= n - b || (int)yi >= n - b || (int)zi < b || (int)zi >= n - b;
    floor_ = (int)yi < b;
  } else {
    const float b = bfrac;
    const float one_mb = 1.0f - b;
    const float n = (float)(int)N;
    const float x = (float)xi / n, y = (float)yi / n, z = (float)zi / n;
    sticky = x < b || x > one_mb || y > one_mb || z < b || z > one_mb;
    floor_ = y < b;
  }
}

// gridOpKernel (mpm.cu:76-107) for one node.
template <class S>
inline void grid_update_node(S cell[4], uint32_t xi, uint32_t yi, uint32_t zi,
                             const StepConsts<S>& k) {
  if (!(cell[3] > S(0))) return;
  for (int w = 0; w < 4; ++w) cell[w] = cell[w] / cell[3];  // cell[3] -> 1 at w = 3
  cell[1] = cell[1] + k.dt * k.gravity;
  bool sticky, floor_;
  node_walls(xi, yi, zi, k.N, k.boundary_mode, float(k.bfrac), k.bcells, sticky, floor_);
  if (sticky) { cell[0] = 0; cell[1] = 0; cell[2] = 0; }
  if (floor_) cell[1] = std::max(S(0), cell[1]);
}

// End-of-step plasticity (MaterialModel plugin's endOfStepMutation).
template <class S>
inline void plasticity(Particle<S>& p, const mpm_material& m) {
  if (m.kind == MPM_MATERIAL_FIXED_COROTATED) return;  // MaterialModel.cuh:63
  if (m.kind == MPM_MATERIAL_JELLY) {                  // MaterialModel.cuh:141-149
    const S oldJ = det3_col0<S>(p.F);
    p.Jp = clampv<S>(p.Jp * oldJ / det3_col0<S>(p.F), S(0.6f), S(20.0f));
    return;
  }
  if (m.kind == MPM_MATERIAL_WC_LIQUID) {
    // J tracked in Jp; F reset to J^(1/3) I
    const S J = det3<S>(p.F);
    const S cj = std::cbrt(J);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) p.F[r][c] = (r == c) ? cj : S(0);
    p.Jp = J;
    return;
  }
  // MMSnow::endOfStepMutation (MaterialModel.cuh:95-114)
  S U[3][3], sig[3], V[3][3];
  svd3<S>(p.F, U, sig, V);
  for (int a = 0; a < 3; ++a) sig[a] = clampv<S>(sig[a], S(m.clamp_lo), S(m.clamp_hi));
  const S oldJ = det3<S>(p.F);
  S US[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) US[r][c] = U[r][c] * sig[c];
  matmul_bt<S>(US, V, p.F);
  const S Fdet = det3<S>(p.F);
  p.Jp = clampv<S>(p.Jp * oldJ / Fdet, S(0.6f), S(20.0f));
}

// gridToParticle body (mpm.cu:109-178) for one particle; grid holds velocities.
template <class S>
inline void g2p_particle(Particle<S>& p, const mpm_material& m, const StepConsts<S>& k,
                         const S* grid) {
  const uint32_t N = k.N;
  const int ns = kernel_size(k.interp);
  int base[3];
  S w[3][4];
  kernel_weights<S>(k, p.x, base, w);
  for (int a = 0; a < 3; ++a)
    if (base[a] + ns < 0 || base[a] >= (int)N) return;  // mpm.cu:129-133: left untouched
  S Dinv[3][3];
  const bool generic = k.interp != INTERP_QUADRATIC;
  if (generic) generic_dinv<S>(p.x, base, w, ns, k.dx, Dinv);
  S v[3] = {0, 0, 0};
  S C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  int beg[3], end[3];
  for (int a = 0; a < 3; ++a) {
    beg[a] = std::max(0, -base[a]);
    end[a] = std::min(ns, (int)N - base[a]);
  }
  for (int i = beg[0]; i < end[0]; ++i) {
    const uint32_t ig = (uint32_t)(base[0] + i);
    const S d0 = S(ig) * k.dx - p.x[0];
    for (int j = beg[1]; j < end[1]; ++j) {
      const uint32_t jg = (uint32_t)(base[1] + j);
      const S d1 = S(jg) * k.dx - p.x[1];
      for (int kk = beg[2]; kk < end[2]; ++kk) {
        const uint32_t kg = (uint32_t)(base[2] + kk);
        const S d2 = S(kg) * k.dx - p.x[2];
        const S* cell = grid + 4 * ((size_t)N * N * ig + (size_t)N * jg + kg);
        const S wt = (w[0][i] * w[1][j]) * w[2][kk];
        // TransferScheme.h:128,131: v += w vg; C += (w vg)(d^T Dinv)
        S dd[3] = {d0 * k.dinv, d1 * k.dinv, d2 * k.dinv};
        if (generic)
          for (int c = 0; c < 3; ++c) dd[c] = dot3<S>(d0, Dinv[0][c], d1, Dinv[1][c], d2, Dinv[2][c]);
        for (int r = 0; r < 3; ++r) {
          const S wv = wt * cell[r];
          v[r] = v[r] + wv;
          for (int c = 0; c < 3; ++c) C[r][c] = C[r][c] + wv * dd[c];
        }
      }
    }
  }
  for (int r = 0; r < 3; ++r) p.v[r] = v[r];
  std::memcpy(p.C, C, sizeof(C));
  // g2p_finish_particle (TransferScheme.h:134-142): F = (I + dt C) F
  S A[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[r][c] = ((r == c) ? S(1) : S(0)) + k.dt * C[r][c];
  matmul<S>(A, p.F, p.F);
  plasticity<S>(p, m);
  // advection (mpm.cu:175)
  for (int a = 0; a < 3; ++a) p.x[a] = p.x[a] + k.dt * p.v[a];
}

// One substep (Simulation::advance, mpm.cu:323-329). `grid` is scratch of
// 4 N^3 scalars; after the call it holds the updated grid velocities.
// `lame` (2 per particle: mu, lambda), when given, replaces the materials'
// mu0, lambda0 per particle.
template <class S>
inline void step(std::vector<Particle<S>>& parts, const mpm_material* mats, int nmat,
                 const StepConsts<S>& k, std::vector<S>& grid, int threads,
                 const S* lame = nullptr) {
  const uint32_t N = k.N;
  const size_t nodes = (size_t)N * N * N;
  grid.assign(4 * nodes, S(0));
  S* g = grid.data();
  const int64_t P = (int64_t)parts.size();
  (void)nmat;
#ifdef _OPENMP
  if (threads > 1) {
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t pi = 0; pi < P; ++pi) {
      p2g_particle<S>(
          parts[pi], mats[parts[pi].mat], k,
          [&](uint64_t idx, const S c4[4]) {
            for (int w = 0; w < 4; ++w) {
#pragma omp atomic
              g[4 * idx + w] += c4[w];
            }
          },
          lame ? lame + 2 * pi : nullptr);
    }
  } else
#endif
  {
    for (int64_t pi = 0; pi < P; ++pi)
      p2g_particle<S>(
          parts[pi], mats[parts[pi].mat], k,
          [&](uint64_t idx, const S c4[4]) {
            for (int w = 0; w < 4; ++w) g[4 * idx + w] += c4[w];
          },
          lame ? lame + 2 * pi : nullptr);
  }
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads > 1 ? threads : 1) schedule(static)
#endif
  for (int64_t n = 0; n < (int64_t)nodes; ++n) {
    const uint32_t zi = (uint32_t)(n % N), yi = (uint32_t)((n / N) % N), xi = (uint32_t)(n / ((size_t)N * N));
    grid_update_node<S>(g + 4 * n, xi, yi, zi, k);
  }
  const int64_t guard = (int64_t)N * N * N;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads > 1 ? threads : 1) schedule(static)
#endif
  for (int64_t pi = 0; pi < P; ++pi) {
    // mpm.cu:114 skips particles whose index exceeds N^3 (compat REFERENCE)
    if (k.compat == MPM_COMPAT_REFERENCE && pi > guard) continue;
    g2p_particle<S>(parts[pi], mats[parts[pi].mat], k, g);
  }
}

}  // namespace oracle
