// device_math.hpp -- per-particle arithmetic of the MLS-MPM substep, in
// registers, for gfx950. Everything here is __device__ inline; the kernels in
// substep.hip and adjoint.hip compose it.
//
// Reference correspondence (kekeblom/mpm):
//   quad_weights        QuadraticInterpolationKernel::weights_per_direction,
//                       include/InterpolationKernel.cuh:57-69
//   svd3                McAdams et al. TR1690 3x3 SVD, include/svd3_cuda.h:35-1043
//   det3                linalg::determinant, src/linalg.cu:48-53
//   kirchhoff_stress    MaterialModel::computePF, include/MaterialModel.cuh:56-61,85-93,131-139
//   plasticity          MaterialModel::endOfStepMutation, MaterialModel.cuh:63,95-114,141-149
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine_types.hpp"

namespace mpmdev {

// ---------------------------------------------------------------------------
// Quadratic B-spline weights. base = int(x/dx - 0.5) truncates toward zero
// (the reference's .cast<int>()), so particles within half a cell below the
// origin keep base 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void quad_weights(const float x[3], float dx_inv, int base[3],
                                             float w[3][3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float xg = x[a] * dx_inv;
    base[a] = (int)(xg - 0.5f);
    const float fx = xg - (float)base[a];
    const float t0 = 1.5f - fx, t1 = fx - 1.0f, t2 = fx - 0.5f;
    w[a][0] = 0.5f * (t0 * t0);
    w[a][1] = 0.75f - t1 * t1;
    w[a][2] = 0.5f * (t2 * t2);
  }
}

__device__ __forceinline__ bool outside_domain(const int base[3], int N) {
  return base[0] + 3 < 0 || base[0] >= N || base[1] + 3 < 0 || base[1] >= N ||
         base[2] + 3 < 0 || base[2] >= N;
}

// ---------------------------------------------------------------------------
// Morton order of grid blocks: 3-way bit interleave of 10-bit coordinates.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ uint32_t compact3(uint32_t v) {
  v &= 0x09249249u;
  v = (v | (v >> 2)) & 0x030c30c3u;
  v = (v | (v >> 4)) & 0x0300f00fu;
  v = (v | (v >> 8)) & 0x030000ffu;
  v = (v | (v >> 16)) & 0x000003ffu;
  return v;
}
__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
}

// ---------------------------------------------------------------------------
// gridOpKernel for one node (reference src/mpm.cu:76-107): momentum / mass
// (three sequential divides; mass becomes 1), gravity on y, then the walls.
// Nodes without mass are left as they are.
// ---------------------------------------------------------------------------
// The walls of node (xi, yi, zi): sticky (all velocity removed) on five
// sides, a separating floor (no downward velocity) at the bottom.
__device__ __forceinline__ void wall_flags(int xi, int yi, int zi, const StepConsts& k,
                                           bool& sticky, bool& floor_) {
  const int N = k.N;
  if (k.boundary_mode == MPM_BOUNDARY_CELLS) {
    const int b = k.bcells;
    sticky = xi < b || xi >= N - b || yi >= N - b || zi < b || zi >= N - b;
    floor_ = yi < b;
  } else {
    const float b = k.bfrac, hi = 1.0f - k.bfrac, fn = (float)N;
    const float xf = (float)xi / fn, yf = (float)yi / fn, zf = (float)zi / fn;
    sticky = xf < b || xf > hi || yf > hi || zf < b || zf > hi;
    floor_ = yf < b;
  }
}

__device__ __forceinline__ float4 grid_node_update(float4 c, int xi, int yi, int zi,
                                                   const StepConsts& k) {
  if (!(c.w > 0.0f)) return c;
  c.x = c.x / c.w;
  c.y = c.y / c.w;
  c.z = c.z / c.w;
  c.w = 1.0f;
  // cell[1] += dt * Gravity (mpm.cu:90), which nvcc contracts into one fused
  // multiply-add: written as one, so that every kernel that inlines this
  // update rounds it the same way (a compiler left to choose may hoist the
  // product of the two constants and round it first)
  c.y = __builtin_fmaf(k.dt, k.gravity, c.y);
  bool sticky, floor_;
  wall_flags(xi, yi, zi, k, sticky, floor_);
  if (sticky) { c.x = 0.f; c.y = 0.f; c.z = 0.f; }
  if (floor_) c.y = fmaxf(0.0f, c.y);
  return c;
}

// 3x3 helpers on row-major float[9]
__device__ __forceinline__ float det3(const float* m) {
  const float sub1 = m[3] * m[7] - m[4] * m[6];
  const float sub2 = m[3] * m[8] - m[5] * m[6];
  const float sub3 = m[4] * m[8] - m[5] * m[7];
  return m[0] * sub3 - m[1] * sub2 + m[2] * sub1;
}
__device__ __forceinline__ float det3_col0(const float* m) {
  return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
         m[6] * (m[1] * m[5] - m[4] * m[2]);
}
// out = a * b^T
__device__ __forceinline__ void mul_abt(const float* a, const float* b, float* out) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      out[3 * r + c] = a[3 * r] * b[3 * c] + (a[3 * r + 1] * b[3 * c + 1] + a[3 * r + 2] * b[3 * c + 2]);
}
// out = a * b
__device__ __forceinline__ void mul_ab(const float* a, const float* b, float* out) {
  float t[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      t[3 * r + c] = a[3 * r] * b[c] + (a[3 * r + 1] * b[3 + c] + a[3 * r + 2] * b[6 + c]);
#pragma unroll
  for (int i = 0; i < 9; ++i) out[i] = t[i];
}

// ---------------------------------------------------------------------------
// 3x3 SVD (McAdams-Selle-Tamstorf-Teran-Sifakis). A = U diag(s) V^T with U, V
// proper rotations, |s0| >= |s1| >= |s2|, s0, s1 >= 0 and s2 carrying the
// sign of det A. Four sweeps of approximate Jacobi rotations on A^T A,
// accumulated as a quaternion; then column sorting of A V and a Givens QR.
// gfx950 notes: every quantity lives in VGPRs (about 40 live), the branch-free
// selects become v_cndmask, and the reciprocal square roots use the hardware
// v_rsq_f32 refined by one Newton step where the method asks for it.
// ---------------------------------------------------------------------------
namespace svdk {
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float rsq_newton(float n2) {
  const float r = rsq(n2);
  const float h = r * 0.5f;
  return (r + h) - n2 * (r * (r * h));
}
__device__ __forceinline__ float f_from_bits(uint32_t b) { return __uint_as_float(b); }

// One approximate Jacobi conjugation on plane (p, q); r is the third axis.
__device__ __forceinline__ void conjugate(float& s_pp, float& s_qp, float& s_qq, float& s_rp,
                                          float& s_rq, float& s_rr, float& qs, float& qp,
                                          float& qq, float& qr) {
  const float half = s_qp * 0.5f;
  const float diff = s_pp - s_qq;
  const bool exact = half * half >= 1.e-20f;
  float sh = exact ? half : 0.0f;
  float ch = exact ? diff : 1.0f;
  const float sh2u = sh * sh, ch2u = ch * ch;
  const float w = rsq(sh2u + ch2u);
  sh *= w;
  ch *= w;
  const bool steep = ch2u <= 5.8284273147583007813f * sh2u;
  sh = steep ? f_from_bits(1053028117u) : sh;  // sin(pi/8)
  ch = steep ? f_from_bits(1064076127u) : ch;  // cos(pi/8) table value
  const float sh2 = sh * sh, ch2 = ch * ch;
  const float c = ch2 - sh2;
  const float s = 2.0f * (ch * sh);
  const float scale = sh2 + ch2;
  s_rr = (s_rr * scale) * scale;
  const float rp = s_rp * scale, rq = s_rq * scale;
  s_rp = c * rp + s * rq;
  s_rq = c * rq - s * rp;
  const float s2 = s * s, c2 = c * c, cs = c * s;
  const float npp = s_pp * c2 + s_qq * s2;
  const float nqq = s_qq * c2 + s_pp * s2;
  const float tw = (s_qp + s_qp) * cs;
  const float nqp = s_qp * (c2 - s2) - diff * cs;
  s_pp = npp + tw;
  s_qq = nqq - tw;
  s_qp = nqp;
  const float tp = sh * qp, tq = sh * qq, tr = sh * qr, ts = sh * qs;
  qs = ch * qs - tr;
  qp = ch * qp + tq;
  qq = ch * qq - tp;
  qr = ch * qr + ts;
}

__device__ __forceinline__ void givens_qr(float b_diag, float b_low, float& c, float& s) {
  float sh = (b_low * b_low >= 1.e-12f) ? b_low : 0.0f;
  float ch = fmaxf(fabsf(b_diag), 1.e-12f);
  const bool pos = b_diag >= 0.0f;
  float n2 = ch * ch + sh * sh;
  ch = ch + rsq_newton(n2) * n2;
  const float ch_ = pos ? ch : sh;
  const float sh_ = pos ? sh : ch;
  n2 = ch_ * ch_ + sh_ * sh_;
  const float r = rsq_newton(n2);
  const float chn = ch_ * r, shn = sh_ * r;
  c = chn * chn - shn * shn;
  s = 2.0f * (shn * chn);
}
}  // namespace svdk

// A, U, V row-major float[9]; s[3]
__device__ __forceinline__ void svd3(const float* A, float* U, float* s, float* V) {
  using namespace svdk;
  float s11 = A[0] * A[0] + (A[3] * A[3] + A[6] * A[6]);
  float s21 = A[1] * A[0] + (A[4] * A[3] + A[7] * A[6]);
  float s31 = A[2] * A[0] + (A[5] * A[3] + A[8] * A[6]);
  float s22 = A[1] * A[1] + (A[4] * A[4] + A[7] * A[7]);
  float s32 = A[2] * A[1] + (A[5] * A[4] + A[8] * A[7]);
  float s33 = A[2] * A[2] + (A[5] * A[5] + A[8] * A[8]);
  float qs = 1.0f, qx = 0.0f, qy = 0.0f, qz = 0.0f;
#pragma unroll
  for (int sweep = 0; sweep < 4; ++sweep) {
    conjugate(s11, s21, s22, s31, s32, s33, qs, qx, qy, qz);
    conjugate(s22, s32, s33, s21, s31, s11, qs, qy, qz, qx);
    conjugate(s33, s31, s11, s32, s21, s22, qs, qz, qx, qy);
  }
  {
    const float n2 = qs * qs + (qx * qx + (qy * qy + qz * qz));
    const float rn = rsq_newton(n2);
    qs *= rn; qx *= rn; qy *= rn; qz *= rn;
  }
  const float xx = qx * qx, yy = qy * qy, zz = qz * qz, ss = qs * qs;
  float v[9];
  v[0] = ((ss + xx) - yy) - zz;
  v[4] = ((ss - xx) + yy) - zz;
  v[8] = ((ss - xx) - yy) + zz;
  {
    const float x2 = qx + qx, y2 = qy + qy, z2 = qz + qz;
    const float sx = qs * x2, sy = qs * y2, sz = qs * z2;
    const float xy = qy * x2, yz = qz * y2, zx = qx * z2;
    v[1] = xy - sz; v[3] = xy + sz;
    v[5] = yz - sx; v[7] = yz + sx;
    v[6] = zx - sy; v[2] = zx + sy;
  }
  float b[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      b[3 * r + c] = A[3 * r] * v[c] + A[3 * r + 1] * v[3 + c] + A[3 * r + 2] * v[6 + c];
  float rho0 = b[0] * b[0] + b[3] * b[3] + b[6] * b[6];
  float rho1 = b[1] * b[1] + b[4] * b[4] + b[7] * b[7];
  float rho2 = b[2] * b[2] + b[5] * b[5] + b[8] * b[8];
  // column sort with sign fixes keeping V a rotation
#define MPM_SWAP_NEG(ca, cb, cn, ra, rb)                                          \
  {                                                                            \
    const bool sw = ra < rb;                                                   \
    _Pragma("unroll") for (int r = 0; r < 3; ++r) {                            \
      const float ba = b[3 * r + ca], bb = b[3 * r + cb];                      \
      const float va = v[3 * r + ca], vb = v[3 * r + cb];                      \
      b[3 * r + ca] = sw ? bb : ba; b[3 * r + cb] = sw ? ba : bb;              \
      v[3 * r + ca] = sw ? vb : va; v[3 * r + cb] = sw ? va : vb;              \
      b[3 * r + cn] = sw ? -b[3 * r + cn] : b[3 * r + cn];                     \
      v[3 * r + cn] = sw ? -v[3 * r + cn] : v[3 * r + cn];                     \
    }                                                                          \
    const float t_a = ra, t_b = rb;                                            \
    ra = sw ? t_b : t_a; rb = sw ? t_a : t_b;                                  \
  }
  MPM_SWAP_NEG(0, 1, 1, rho0, rho1)
  MPM_SWAP_NEG(0, 2, 0, rho0, rho2)
  MPM_SWAP_NEG(1, 2, 2, rho1, rho2)
#undef MPM_SWAP_NEG
  float u[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
#define MPM_QR_ROT(ra, rb)                                                        \
  {                                                                            \
    float c, sn;                                                               \
    givens_qr(b[3 * ra + ra], b[3 * rb + ra], c, sn);                          \
    _Pragma("unroll") for (int k = 0; k < 3; ++k) {                            \
      const float xa = b[3 * ra + k], xb = b[3 * rb + k];                      \
      b[3 * ra + k] = c * xa + sn * xb;                                        \
      b[3 * rb + k] = c * xb - sn * xa;                                        \
      const float ua = u[3 * k + ra], ub = u[3 * k + rb];                      \
      u[3 * k + ra] = c * ua + sn * ub;                                        \
      u[3 * k + rb] = c * ub - sn * ua;                                        \
    }                                                                          \
  }
  MPM_QR_ROT(0, 1)
  MPM_QR_ROT(0, 2)
  MPM_QR_ROT(1, 2)
#undef MPM_QR_ROT
#pragma unroll
  for (int i = 0; i < 9; ++i) { U[i] = u[i]; V[i] = v[i]; }
  s[0] = b[0];
  s[1] = b[4];
  s[2] = b[8];
}

// ---------------------------------------------------------------------------
// FP64 fallback SVD for near-degenerate F, taken when the FP32 result has
// |s2| < 1e-3 |s0| (or is not finite). Same scheme in double: Jacobi sweeps
// on A^T A with the half-angle taken from (s_pp - s_qq, s_qp / 2) (no pi/8
// substitution, iterated to convergence), quaternion-accumulated V, sorted
// columns of A V and a Givens QR with exact rotations. Kept out of line so
// its double registers do not count against the hot kernels' allocation.
// ---------------------------------------------------------------------------
namespace svd64 {
__device__ __forceinline__ void conjugate(double& s_pp, double& s_qp, double& s_qq, double& s_rp,
                                          double& s_rq, double& s_rr, double& qs, double& qp,
                                          double& qq, double& qr) {
  // exact Jacobi rotation: tan(2 theta) = 2 s_qp / (s_pp - s_qq), built from
  // t = tan(theta) without trigonometric functions
  const double diff = s_pp - s_qq;
  if (s_qp == 0.0) return;
  const double tau = diff / (2.0 * s_qp);
  const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
  const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
  // half-angle quaternion components: ch = cos(theta/2), sh = sin(theta/2)
  const double ch = sqrt(0.5 * (1.0 + c));
  const double sh = s / (2.0 * ch);
  const double rp = s_rp, rq = s_rq;
  s_rp = c * rp + s * rq;
  s_rq = c * rq - s * rp;
  const double c2 = c * c, s2 = s * s, cs = c * s;
  const double npp = s_pp * c2 + s_qq * s2 + 2.0 * s_qp * cs;
  const double nqq = s_qq * c2 + s_pp * s2 - 2.0 * s_qp * cs;
  s_qp = s_qp * (c2 - s2) - diff * cs;
  s_pp = npp;
  s_qq = nqq;
  (void)s_rr;
  const double tp = sh * qp, tq = sh * qq, tr = sh * qr, ts = sh * qs;
  qs = ch * qs - tr;
  qp = ch * qp + tq;
  qq = ch * qq - tp;
  qr = ch * qr + ts;
}
}  // namespace svd64

__device__ __noinline__ void svd3_f64(const float* Af, float* U, float* s, float* V) {
  double A[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) A[i] = (double)Af[i];
  double s11 = A[0] * A[0] + A[3] * A[3] + A[6] * A[6];
  double s21 = A[1] * A[0] + A[4] * A[3] + A[7] * A[6];
  double s31 = A[2] * A[0] + A[5] * A[3] + A[8] * A[6];
  double s22 = A[1] * A[1] + A[4] * A[4] + A[7] * A[7];
  double s32 = A[2] * A[1] + A[5] * A[4] + A[8] * A[7];
  double s33 = A[2] * A[2] + A[5] * A[5] + A[8] * A[8];
  double qs = 1, qx = 0, qy = 0, qz = 0;
  for (int sweep = 0; sweep < 12; ++sweep) {
    const double off = fabs(s21) + fabs(s32) + fabs(s31);
    const double dia = fabs(s11) + fabs(s22) + fabs(s33);
    if (!(off > 1e-17 * dia)) break;
    svd64::conjugate(s11, s21, s22, s31, s32, s33, qs, qx, qy, qz);
    svd64::conjugate(s22, s32, s33, s21, s31, s11, qs, qy, qz, qx);
    svd64::conjugate(s33, s31, s11, s32, s21, s22, qs, qz, qx, qy);
  }
  const double rn = 1.0 / sqrt(qs * qs + qx * qx + qy * qy + qz * qz);
  qs *= rn; qx *= rn; qy *= rn; qz *= rn;
  double v[9];
  v[0] = qs * qs + qx * qx - qy * qy - qz * qz;
  v[4] = qs * qs - qx * qx + qy * qy - qz * qz;
  v[8] = qs * qs - qx * qx - qy * qy + qz * qz;
  v[1] = 2 * (qx * qy - qs * qz); v[3] = 2 * (qx * qy + qs * qz);
  v[5] = 2 * (qy * qz - qs * qx); v[7] = 2 * (qy * qz + qs * qx);
  v[6] = 2 * (qx * qz - qs * qy); v[2] = 2 * (qx * qz + qs * qy);
  double b[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      b[3 * r + c] = A[3 * r] * v[c] + A[3 * r + 1] * v[3 + c] + A[3 * r + 2] * v[6 + c];
  double rho[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) rho[c] = b[c] * b[c] + b[3 + c] * b[3 + c] + b[6 + c] * b[6 + c];
#define MPM_SWAP64(ca, cb, cn)                                                           \
  if (rho[ca] < rho[cb]) {                                                               \
    _Pragma("unroll") for (int r = 0; r < 3; ++r) {                                      \
      const double tb = b[3 * r + ca]; b[3 * r + ca] = b[3 * r + cb]; b[3 * r + cb] = tb; \
      const double tv = v[3 * r + ca]; v[3 * r + ca] = v[3 * r + cb]; v[3 * r + cb] = tv; \
    }                                                                                    \
    const double tr = rho[ca]; rho[ca] = rho[cb]; rho[cb] = tr;                          \
    _Pragma("unroll") for (int r = 0; r < 3; ++r) {                                      \
      b[3 * r + cn] = -b[3 * r + cn]; v[3 * r + cn] = -v[3 * r + cn];                    \
    }                                                                                    \
  }
  MPM_SWAP64(0, 1, 1)
  MPM_SWAP64(0, 2, 0)
  MPM_SWAP64(1, 2, 2)
#undef MPM_SWAP64
  double u[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#define MPM_QR64(ra, rb)                                                   \
  {                                                                     \
    const double x = b[3 * ra + ra], y = b[3 * rb + ra];                \
    const double h = sqrt(x * x + y * y);                               \
    const double c = (h > 0.0) ? x / h : 1.0, sn = (h > 0.0) ? y / h : 0.0; \
    _Pragma("unroll") for (int k = 0; k < 3; ++k) {                     \
      const double xa = b[3 * ra + k], xb = b[3 * rb + k];              \
      b[3 * ra + k] = c * xa + sn * xb;                                 \
      b[3 * rb + k] = c * xb - sn * xa;                                 \
      const double ua = u[3 * k + ra], ub = u[3 * k + rb];              \
      u[3 * k + ra] = c * ua + sn * ub;                                 \
      u[3 * k + rb] = c * ub - sn * ua;                                 \
    }                                                                   \
  }
  MPM_QR64(0, 1)
  MPM_QR64(0, 2)
  MPM_QR64(1, 2)
#undef MPM_QR64
  // the exact-rotation QR leaves diagonal entries of either sign: make s0, s1
  // non-negative by flipping a column pair of U (U stays a rotation)
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    if (b[4 * t] < 0.0) {
      b[4 * t] = -b[4 * t];
      b[8] = -b[8];
#pragma unroll
      for (int k = 0; k < 3; ++k) { u[3 * k + t] = -u[3 * k + t]; u[3 * k + 2] = -u[3 * k + 2]; }
    }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) { U[i] = (float)u[i]; V[i] = (float)v[i]; }
  s[0] = (float)b[0]; s[1] = (float)b[4]; s[2] = (float)b[8];
}

__device__ __forceinline__ void svd3_robust(const float* A, float* U, float* s, float* V) {
  svd3(A, U, s, V);
  const bool bad = !(fabsf(s[2]) >= 1e-3f * fabsf(s[0])) || !isfinite(s[0]) || !isfinite(s[2]);
  if (bad) {
    // The fallback is a call: the arrays whose addresses it takes live in
    // private memory. These copies are those arrays, so that the caller's
    // (the particle state, U, s, V) stay in registers on the common path.
    float a[9], u[9], w[3], v[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) a[i] = A[i];
    svd3_f64(a, u, w, v);
#pragma unroll
    for (int i = 0; i < 9; ++i) { U[i] = u[i]; V[i] = v[i]; }
#pragma unroll
    for (int i = 0; i < 3; ++i) s[i] = w[i];
  }
}

// ---------------------------------------------------------------------------
// Material stress (Kirchhoff tau, what the reference's computePF returns) and
// end-of-step plasticity, dispatched on the material kind.
// ---------------------------------------------------------------------------
// R: polar rotation of F (unused for WC_LIQUID). Jp: plastic J (or J for
// WC_LIQUID). Writes tau into out[9].
__device__ __forceinline__ void kirchhoff_stress(const float* F, float Jp, const float* R,
                                                 const MaterialDev& m, int compat, float* out) {
  if (m.kind == MPM_MATERIAL_WC_LIQUID) {
    const float pr = m.bulk * ((Jp - 1.0f) * Jp);
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = (i % 4 == 0) ? pr : 0.0f;
    return;
  }
  float e = 1.0f;
  if (m.kind != MPM_MATERIAL_FIXED_COROTATED) e = __expf(m.hardening * (1.0f - Jp));
  const float mu = m.mu0 * e, lambda = m.lambda0 * e;
  const float J = (compat == MPM_COMPAT_FIXED) ? det3(F) : Jp;
  float fmr[9];
  const float two_mu = 2.0f * mu;
#pragma unroll
  for (int i = 0; i < 9; ++i) fmr[i] = two_mu * (F[i] - R[i]);
  mul_abt(fmr, F, out);
  const float vol = lambda * ((J - 1.0f) * J);
  out[0] += vol;
  out[4] += vol;
  out[8] += vol;
}

// Plasticity / state reset after the F update. For SNOW the SVD of the new F
// is computed here and its rotation U V^T is returned in R_out (the polar
// rotation of the returned F, reused by the fused P2G); has_R tells whether
// R_out was filled.
__device__ __forceinline__ void plasticity(float* F, float& Jp, const MaterialDev& m, float* R_out,
                                           bool& has_R) {
  has_R = false;
  if (m.kind == MPM_MATERIAL_FIXED_COROTATED) return;
  if (m.kind == MPM_MATERIAL_JELLY) {
    const float d = det3_col0(F);
    Jp = fmaxf(fminf(Jp * d / d, 20.0f), 0.6f);
    return;
  }
  if (m.kind == MPM_MATERIAL_WC_LIQUID) {
    const float J = det3(F);
    const float cj = cbrtf(J);
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = (i % 4 == 0) ? cj : 0.0f;
    Jp = J;
    return;
  }
  float U[9], s[3], V[9];
  svd3_robust(F, U, s, V);
#pragma unroll
  for (int a = 0; a < 3; ++a) s[a] = fmaxf(fminf(s[a], m.clamp_hi), m.clamp_lo);
  const float oldJ = det3(F);
  float us[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) us[3 * r + c] = U[3 * r + c] * s[c];
  mul_abt(us, V, F);
  const float Fdet = det3(F);
  Jp = fmaxf(fminf(Jp * oldJ / Fdet, 20.0f), 0.6f);
  mul_abt(U, V, R_out);
  has_R = true;
}

__device__ __forceinline__ void polar_rotation(const float* F, float* R) {
  float U[9], s[3], V[9];
  svd3_robust(F, U, s, V);
  mul_abt(U, V, R);
}

// ---------------------------------------------------------------------------
// Multi-scene launches: index g (a workgroup or a block of the launch) lies in
// scene j when prefix[j] <= g < prefix[j + 1]. Every lane of the wavefront
// calls this with the same g: lane l tests scene c + l of each chunk of 64
// scenes, so one round trip of loads finds the scene (a binary search takes
// log2(scenes) dependent ones). g must be below prefix[nscenes].
__device__ __forceinline__ int scene_of_wave(const uint32_t* __restrict__ prefix, int nscenes, uint32_t g) {
  const int lane = (int)(threadIdx.x & 63u);
  for (int c = 0; c < nscenes; c += 64) {
    const int j = c + lane;
    const bool hit = j < nscenes && prefix[j] <= g && g < prefix[j + 1];
    const uint64_t m = __ballot(hit);
    if (m != 0ull) return c + __ffsll((long long)m) - 1;
  }
  return nscenes - 1;
}

}  // namespace mpmdev
