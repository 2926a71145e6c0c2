// adjoint.hip -- a differentiable MLS-MPM substep and its hand-written adjoint
// (gfx950). No reference counterpart: the reference has no gradients.
//
// The substep is the engine's (P2G, grid update, G2P: reference src/mpm.cu:
// 14-178) for the fixed-corotated model (MMFixedCorotated, MaterialModel.cuh:
// 56-61) with J = det F, per-particle Lame parameters mu_p, lambda_p, no
// plasticity, and the stencil clipped at the grid faces (N - base). State
// planes (24 x n floats, plane-major): x(3) v(3) C(9) F(9), 3x3 row-major.
// The forward records a grid tape per node (velocity and mass after the
// update, velocity before the walls, wall flags; 8 floats) that the backward
// replays:
//   k_adj_g2p   G2P backward: adjoints of x', v', C', F' -> partial x, F
//               adjoints of the particle, and the grid velocity adjoint
//               (scattered with FP32 atomics);
//   k_adj_grid  grid update backward: momentum and mass adjoints per node;
//   k_adj_p2g   P2G and stress backward: completes the particle adjoints
//               (x, v, C, F) and accumulates the mu, lambda adjoints.
// The stress adjoint differentiates the polar rotation R of F = U S V^T:
// with Rbar the adjoint of R and M = V^T skew(R^T Rbar) V,
// Fbar += 2 U N V^T, N_ij = M_ij / (s_i + s_j).
//
// One thread per particle. The two scatters stage their sums in an LDS tile
// per workgroup (the box tile below); the forward and backward each take
// four launches per substep.
#include <hip/hip_runtime.h>

#include <mutex>

#include "device_math.hpp"
#include "engine_types.hpp"
#include "kernels.hpp"

using namespace mpmdev;

namespace {

enum : int { DX = 0, DV = 3, DC = 6, DF = 15 };

struct DiffK {
  StepConsts k;
  float mass, volume;
  int64_t n;
};

// quadratic B-spline weights and their derivatives (InterpolationKernel.cuh)
struct Stencil {
  int base[3];
  float w[3][3], dw[3][3];
  bool outside;  // no node of the stencil in the grid: the particle is left as it is
};

__device__ __forceinline__ Stencil make_stencil(const float x[3], const StepConsts& k) {
  Stencil s;
  s.outside = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float xg = x[a] * k.dx_inv;
    s.base[a] = (int)(xg - 0.5f);
    const float f = xg - (float)s.base[a];
    const float t0 = 1.5f - f, t1 = f - 1.0f, t2 = f - 0.5f;
    s.w[a][0] = 0.5f * (t0 * t0);
    s.w[a][1] = 0.75f - t1 * t1;
    s.w[a][2] = 0.5f * (t2 * t2);
    s.dw[a][0] = -t0 * k.dx_inv;
    s.dw[a][1] = -2.0f * t1 * k.dx_inv;
    s.dw[a][2] = t2 * k.dx_inv;
    s.outside |= s.base[a] + 3 < 0 || s.base[a] >= k.N;  // mpm.cu:31-35,129-133
  }
  return s;
}

__device__ __forceinline__ bool node_ok(const Stencil& s, int a, int b, int c, int N) {
  return (unsigned)(s.base[0] + a) < (unsigned)N && (unsigned)(s.base[1] + b) < (unsigned)N &&
         (unsigned)(s.base[2] + c) < (unsigned)N;
}

__device__ __forceinline__ size_t node_index(const Stencil& s, int a, int b, int c, int N) {
  return ((size_t)(s.base[0] + a) * N + (size_t)(s.base[1] + b)) * N + (size_t)(s.base[2] + c);
}

__device__ __forceinline__ float weight(const Stencil& s, int a, int b, int c) {
  return (s.w[0][a] * s.w[1][b]) * s.w[2][c];
}

__device__ __forceinline__ void weight_grad(const Stencil& s, int a, int b, int c, float gw[3]) {
  gw[0] = (s.dw[0][a] * s.w[1][b]) * s.w[2][c];
  gw[1] = (s.w[0][a] * s.dw[1][b]) * s.w[2][c];
  gw[2] = (s.w[0][a] * s.w[1][b]) * s.dw[2][c];
}

__device__ __forceinline__ void node_offset(const Stencil& s, int a, int b, int c, const float x[3],
                                            const StepConsts& k, float dp[3]) {
  dp[0] = (float)(s.base[0] + a) * k.dx - x[0];
  dp[1] = (float)(s.base[1] + b) * k.dx - x[1];
  dp[2] = (float)(s.base[2] + c) * k.dx - x[2];
}

__device__ __forceinline__ void load_state(const float* st, int64_t n, int64_t p, float x[3],
                                           float v[3], float C[9], float F[9]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    x[a] = st[(DX + a) * n + p];
    v[a] = st[(DV + a) * n + p];
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    C[a] = st[(DC + a) * n + p];
    F[a] = st[(DF + a) * n + p];
  }
}

__device__ __forceinline__ void store_state(float* st, int64_t n, int64_t p, const float x[3],
                                            const float v[3], const float C[9], const float F[9]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    st[(DX + a) * n + p] = x[a];
    st[(DV + a) * n + p] = v[a];
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    st[(DC + a) * n + p] = C[a];
    st[(DF + a) * n + p] = F[a];
  }
}

// tau = 2 mu (F - R) F^T + lambda (J - 1) J I, J = det F
__device__ __forceinline__ void fc_stress(const float F[9], float mu, float lam, float U[9],
                                          float s[3], float V[9], float R[9], float tau[9]) {
  svd3_robust(F, U, s, V);
  mul_abt(U, V, R);
  float FmR[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) FmR[i] = 2.0f * mu * (F[i] - R[i]);
  mul_abt(FmR, F, tau);
  const float J = det3(F);
  const float vol = lam * ((J - 1.0f) * J);
  tau[0] += vol;
  tau[4] += vol;
  tau[8] += vol;
}

// the affine momentum matrix of P2G is A = sc tau + m C, sc = -Dinv dt V
__device__ __forceinline__ float stress_scale(const DiffK& d) {
  return ((-d.k.dinv) * d.k.dt) * d.volume;
}

// ---------------------------------------------------------------- box tile
// The scatters (P2G, and the grid adjoint of G2P) stage a workgroup's
// contributions in LDS over the bounding box of its particles' stencils,
// then add each touched node to HBM once. Particles that neighbour in memory
// and in space (lattice or sorted order) share most nodes, so this turns
// ~1000 contended global atomics per node into a few. When the box exceeds
// the LDS tile (scattered particle order), the workgroup adds to HBM directly.
constexpr int kBoxFloats = 8192;  // 32 KiB

struct Box {
  int lo[3], ext[3];
  bool lds;  // the tile holds the box
};

// Every thread of the workgroup calls this (it synchronises).
__device__ __forceinline__ Box block_box(const Stencil& sn, bool active, int N, int nf, int* sh) {
  if (threadIdx.x < 6) sh[threadIdx.x] = threadIdx.x < 3 ? 0x7fffffff : -0x7fffffff;
  __syncthreads();
  if (active) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      atomicMin(&sh[a], max(sn.base[a], 0));
      atomicMax(&sh[3 + a], min(sn.base[a] + 2, N - 1));
    }
  }
  __syncthreads();
  Box b;
  long vol = 1;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = sh[a];
    b.ext[a] = sh[3 + a] - sh[a] + 1;
    vol *= b.ext[a] > 0 ? b.ext[a] : 0;
  }
  b.lds = vol > 0 && vol * nf <= kBoxFloats;
  return b;
}

__device__ __forceinline__ int box_slot(const Box& b, const Stencil& sn, int a, int c1, int c2) {
  return ((sn.base[0] + a - b.lo[0]) * b.ext[1] + (sn.base[1] + c1 - b.lo[1])) * b.ext[2] +
         (sn.base[2] + c2 - b.lo[2]);
}

__device__ __forceinline__ void box_clear(const Box& b, int nf, float* tile) {
  const int n = b.ext[0] * b.ext[1] * b.ext[2] * nf;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tile[i] = 0.0f;
}

// adds the tile's nonzero entries to the grid (4 floats per node)
__device__ __forceinline__ void box_flush(const Box& b, int nf, const float* tile, float* grid, int N) {
  const int n = b.ext[0] * b.ext[1] * b.ext[2];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int z = i % b.ext[2], y = (i / b.ext[2]) % b.ext[1], x = i / (b.ext[2] * b.ext[1]);
    float* g = grid + 4 * (((size_t)(b.lo[0] + x) * N + (size_t)(b.lo[1] + y)) * N + (size_t)(b.lo[2] + z));
    for (int r = 0; r < nf; ++r) {
      const float v = tile[nf * i + r];
      if (v != 0.0f) unsafeAtomicAdd(g + r, v);
    }
  }
}

// ---------------------------------------------------------------- forward
__global__ void __launch_bounds__(256)
    k_diff_p2g(const float* __restrict__ st, const float* __restrict__ mu,
               const float* __restrict__ lam, DiffK d, float* __restrict__ grid) {
  __shared__ float tile[kBoxFloats];
  __shared__ int sh[6];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const StepConsts& k = d.k;
  float x[3] = {0.f, 0.f, 0.f}, v[3], C[9], F[9];
  Stencil sn;
  bool active = p < d.n;
  if (active) {
    load_state(st, d.n, p, x, v, C, F);
    sn = make_stencil(x, k);
    active = !sn.outside;
  }
  const Box box = block_box(sn, active, k.N, 4, sh);
  if (box.lds) {
    box_clear(box, 4, tile);
    __syncthreads();
  }
  if (active) {
    float U[9], s[3], V[9], R[9], tau[9];
    fc_stress(F, mu[p], lam[p], U, s, V, R, tau);
    const float sc = stress_scale(d);
    float A[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i] = sc * tau[i] + d.mass * C[i];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (!node_ok(sn, a, b, c, k.N)) continue;
          const float wt = weight(sn, a, b, c);
          float dp[3];
          node_offset(sn, a, b, c, x, k, dp);
          float q[4];
#pragma unroll
          for (int r = 0; r < 3; ++r)
            q[r] = wt * (d.mass * v[r] + (A[3 * r] * dp[0] + A[3 * r + 1] * dp[1] + A[3 * r + 2] * dp[2]));
          q[3] = wt * d.mass;
          float* g = box.lds ? tile + 4 * box_slot(box, sn, a, b, c)
                             : grid + 4 * node_index(sn, a, b, c, k.N);
#pragma unroll
          for (int r = 0; r < 4; ++r) unsafeAtomicAdd(g + r, q[r]);
        }
  }
  if (box.lds) {
    __syncthreads();
    box_flush(box, 4, tile, grid, k.N);
  }
}

// grid update with tape: tape[8 node + 0..3] = (velocity, mass) after the
// update, [4..6] = velocity before the walls, [7] = wall flags (1 sticky,
// 2 floor)
__global__ void __launch_bounds__(256)
    k_diff_grid(const float4* __restrict__ grid, float* __restrict__ tape, StepConsts k) {
  const int N = k.N;
  const size_t nodes = (size_t)N * N * N;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nodes) return;
  const float4 c = grid[t];
  float out[8] = {0.f, 0.f, 0.f, c.w, 0.f, 0.f, 0.f, 0.f};
  if (c.w > 0.0f) {
    const int zi = (int)(t % N), yi = (int)((t / N) % N), xi = (int)(t / ((size_t)N * N));
    out[4] = c.x / c.w;
    out[5] = c.y / c.w + k.dt * k.gravity;
    out[6] = c.z / c.w;
    const float4 u = grid_node_update(c, xi, yi, zi, k);
    bool sticky, floor_;
    wall_flags(xi, yi, zi, k, sticky, floor_);
    out[0] = u.x;
    out[1] = u.y;
    out[2] = u.z;
    out[7] = (float)((sticky ? 1 : 0) | (floor_ ? 2 : 0));
  }
  float4* o = reinterpret_cast<float4*>(tape + 8 * t);
  o[0] = make_float4(out[0], out[1], out[2], out[3]);
  o[1] = make_float4(out[4], out[5], out[6], out[7]);
}

// v' = sum_i w_i v_i and C' = Dinv sum_i w_i v_i (x_i - x_p)^T over the stencil
__device__ __forceinline__ void g2p_affine(const Stencil& sn, const float x[3], const float* tape,
                                           const StepConsts& k, float vn[3], float Cn[9]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) vn[a] = 0.f;
#pragma unroll
  for (int i = 0; i < 9; ++i) Cn[i] = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!node_ok(sn, a, b, c, k.N)) continue;
        const float wt = weight(sn, a, b, c);
        const float4 gv = *reinterpret_cast<const float4*>(tape + 8 * node_index(sn, a, b, c, k.N));
        float dp[3];
        node_offset(sn, a, b, c, x, k, dp);
        const float g[3] = {gv.x, gv.y, gv.z};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          vn[r] += wt * g[r];
#pragma unroll
          for (int q = 0; q < 3; ++q) Cn[3 * r + q] += (k.dinv * wt) * g[r] * dp[q];
        }
      }
}

__global__ void __launch_bounds__(256)
    k_diff_g2p(const float* __restrict__ st, float* __restrict__ out, const float* __restrict__ tape,
               DiffK d) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= d.n) return;
  const StepConsts& k = d.k;
  float x[3], v[3], C[9], F[9];
  load_state(st, d.n, p, x, v, C, F);
  const Stencil sn = make_stencil(x, k);
  if (!sn.outside) {
    float vn[3], Cn[9];
    g2p_affine(sn, x, tape, k, vn, Cn);
    float A[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i] = ((i % 4 == 0) ? 1.0f : 0.0f) + k.dt * Cn[i];
    mul_ab(A, F, F);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      v[a] = vn[a];
      x[a] = x[a] + k.dt * vn[a];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) C[i] = Cn[i];
  }
  store_state(out, d.n, p, x, v, C, F);
}

// ---------------------------------------------------------------- backward
// Adjoint planes use the state layout. adj_in receives the partial x and F
// adjoints here (k_adj_p2g completes them); grid_adj (4 floats per node)
// receives the adjoint of the grid velocity after the update.
__global__ void __launch_bounds__(256)
    k_adj_g2p(const float* __restrict__ st, const float* __restrict__ tape,
              const float* __restrict__ adj_out, float* __restrict__ adj_in,
              float* __restrict__ grid_adj, DiffK d) {
  __shared__ float tile[kBoxFloats];
  __shared__ int sh[6];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const StepConsts& k = d.k;
  const int64_t n = d.n;
  float x[3] = {0.f, 0.f, 0.f}, v[3], C[9], F[9];
  Stencil sn;
  bool active = p < n;
  if (active) {
    load_state(st, n, p, x, v, C, F);
    sn = make_stencil(x, k);
    active = !sn.outside;
  }
  const Box box = block_box(sn, active, k.N, 3, sh);
  if (box.lds) {
    box_clear(box, 3, tile);
    __syncthreads();
  }
  if (p < n) {
    float xb2[3], vb2[3], Cb2[9], Fb2[9];
    load_state(adj_out, n, p, xb2, vb2, Cb2, Fb2);
    if (!active) {
      store_state(adj_in, n, p, xb2, vb2, Cb2, Fb2);  // the identity map
    } else {
      float vn[3], Cn[9];
      g2p_affine(sn, x, tape, k, vn, Cn);
      // x' = x + dt v', F' = (I + dt C') F: adjoints of v' and C' through them
      float vbt[3], Cbt[9], FbFt[9];
#pragma unroll
      for (int a = 0; a < 3; ++a) vbt[a] = vb2[a] + k.dt * xb2[a];
      mul_abt(Fb2, F, FbFt);
#pragma unroll
      for (int i = 0; i < 9; ++i) Cbt[i] = Cb2[i] + k.dt * FbFt[i];
      float xb[3] = {xb2[0], xb2[1], xb2[2]}, Fb[9];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          // ((I + dt C')^T Fb')_rc
          float acc = 0.f;
#pragma unroll
          for (int q = 0; q < 3; ++q)
            acc += (((q == r) ? 1.0f : 0.0f) + k.dt * Cn[3 * q + r]) * Fb2[3 * q + c];
          Fb[3 * r + c] = acc;
        }
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            if (!node_ok(sn, a, b, c, k.N)) continue;
            const float wt = weight(sn, a, b, c);
            float gw[3], dp[3];
            weight_grad(sn, a, b, c, gw);
            node_offset(sn, a, b, c, x, k, dp);
            const size_t ni = node_index(sn, a, b, c, k.N);
            const float4 gv4 = *reinterpret_cast<const float4*>(tape + 8 * ni);
            const float gv[3] = {gv4.x, gv4.y, gv4.z};
            float Cdp[3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
              Cdp[r] = Cbt[3 * r] * dp[0] + Cbt[3 * r + 1] * dp[1] + Cbt[3 * r + 2] * dp[2];
            const float s1 = gv[0] * vbt[0] + gv[1] * vbt[1] + gv[2] * vbt[2];
            const float s2 = gv[0] * Cdp[0] + gv[1] * Cdp[1] + gv[2] * Cdp[2];
            const float coef = s1 + k.dinv * s2;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              const float ctg = Cbt[q] * gv[0] + Cbt[3 + q] * gv[1] + Cbt[6 + q] * gv[2];  // (C'^T g)_q
              xb[q] += gw[q] * coef - (k.dinv * wt) * ctg;
            }
            float* ga = box.lds ? tile + 3 * box_slot(box, sn, a, b, c) : grid_adj + 4 * ni;
#pragma unroll
            for (int r = 0; r < 3; ++r) unsafeAtomicAdd(ga + r, wt * (vbt[r] + k.dinv * Cdp[r]));
          }
#pragma unroll
      for (int a = 0; a < 3; ++a) adj_in[(DX + a) * n + p] = xb[a];
#pragma unroll
      for (int i = 0; i < 9; ++i) adj_in[(DF + i) * n + p] = Fb[i];
    }
  }
  if (box.lds) {
    __syncthreads();
    box_flush(box, 3, tile, grid_adj, k.N);
  }
}

// adjoint of the updated velocity -> adjoints of the P2G momentum and mass
__global__ void __launch_bounds__(256)
    k_adj_grid(float4* __restrict__ grid_adj, const float* __restrict__ tape, StepConsts k) {
  const size_t nodes = (size_t)k.N * k.N * k.N;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nodes) return;
  const float4 t0 = *reinterpret_cast<const float4*>(tape + 8 * t);
  const float4 t1 = *reinterpret_cast<const float4*>(tape + 8 * t + 4);
  const float m = t0.w;
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  if (m > 0.0f) {
    const float4 g = grid_adj[t];
    float b0 = g.x, b1 = g.y, b2 = g.z;
    const int flags = (int)t1.w;
    if (flags & 1) { b0 = 0.f; b1 = 0.f; b2 = 0.f; }
    if ((flags & 2) && !(t1.y > 0.0f)) b1 = 0.f;  // max(0, v_y) is flat where it clamps
    // v = P / m + dt g e_y: dP = b / m, dm = -b . (P / m) / m
    const float py = t1.y - k.dt * k.gravity;
    out.x = b0 / m;
    out.y = b1 / m;
    out.z = b2 / m;
    out.w = -(b0 * t1.x + b1 * py + b2 * t1.z) / m;
  }
  grid_adj[t] = out;
}

__global__ void __launch_bounds__(256)
    k_adj_p2g(const float* __restrict__ st, const float* __restrict__ mu,
              const float* __restrict__ lam, const float4* __restrict__ grid_adj,
              float* __restrict__ adj_in, float* __restrict__ adj_mu, float* __restrict__ adj_lam,
              DiffK d) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= d.n) return;
  const StepConsts& k = d.k;
  const int64_t n = d.n;
  float x[3], v[3], C[9], F[9];
  load_state(st, n, p, x, v, C, F);
  const Stencil sn = make_stencil(x, k);
  if (sn.outside) return;  // k_adj_g2p wrote the identity adjoints
  const float mup = mu[p], lamp = lam[p];
  float U[9], s[3], V[9], R[9], tau[9];
  fc_stress(F, mup, lamp, U, s, V, R, tau);
  const float sc = stress_scale(d);
  float A[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) A[i] = sc * tau[i] + d.mass * C[i];
  float xb[3], Fb[9];
#pragma unroll
  for (int a = 0; a < 3; ++a) xb[a] = adj_in[(DX + a) * n + p];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fb[i] = adj_in[(DF + i) * n + p];
  float vb[3] = {0.f, 0.f, 0.f}, Ab[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Ab[i] = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!node_ok(sn, a, b, c, k.N)) continue;
        const float wt = weight(sn, a, b, c);
        float gw[3], dp[3];
        weight_grad(sn, a, b, c, gw);
        node_offset(sn, a, b, c, x, k, dp);
        const float4 g = grid_adj[node_index(sn, a, b, c, k.N)];
        const float pb[3] = {g.x, g.y, g.z};
        float val = g.w * d.mass;  // pb . (m v + A dp) + mb m
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          vb[r] += d.mass * wt * pb[r];
          val += pb[r] * (d.mass * v[r] + (A[3 * r] * dp[0] + A[3 * r + 1] * dp[1] + A[3 * r + 2] * dp[2]));
#pragma unroll
          for (int q = 0; q < 3; ++q) Ab[3 * r + q] += wt * pb[r] * dp[q];
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const float atp = A[q] * pb[0] + A[3 + q] * pb[1] + A[6 + q] * pb[2];  // (A^T pb)_q
          xb[q] += gw[q] * val - wt * atp;
        }
      }
  // stress adjoint tb = sc Ab; mu and lambda adjoints
  float tb[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) tb[i] = sc * Ab[i];
  const float J = det3(F);
  float FmR[9], FmRFt[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) FmR[i] = F[i] - R[i];
  mul_abt(FmR, F, FmRFt);
  float dmu = 0.f;
#pragma unroll
  for (int i = 0; i < 9; ++i) dmu += FmRFt[i] * tb[i];
  const float trt = tb[0] + tb[4] + tb[8];
  adj_mu[p] += 2.0f * dmu;
  adj_lam[p] += ((J - 1.0f) * J) * trt;
  // with R held: d/dF of 2 mu (F - R) F^T contracted with tb = 2 mu (tb F + tb^T (F - R))
  float tbF[9], tbtFmR[9];
  mul_ab(tb, F, tbF);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      tbtFmR[3 * r + c] = tb[r] * FmR[c] + tb[3 + r] * FmR[3 + c] + tb[6 + r] * FmR[6 + c];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fb[i] += 2.0f * mup * (tbF[i] + tbtFmR[i]);
  // through R: Rbar = -2 mu tb F, M = R^T Rbar
  float M[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      M[3 * r + c] = -2.0f * mup * (R[r] * tbF[c] + R[3 + r] * tbF[3 + c] + R[6 + r] * tbF[6 + c]);
  float Sk[9], T1[9], Mt[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Sk[3 * r + c] = 0.5f * (M[3 * r + c] - M[3 * c + r]);
  mul_ab(Sk, V, T1);  // skew(M) V
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Mt[3 * r + c] = V[r] * T1[c] + V[3 + r] * T1[3 + c] + V[6 + r] * T1[6 + c];
  float Nm[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float den = s[r] + s[c];
      Nm[3 * r + c] = (r == c || den == 0.0f) ? 0.0f : Mt[3 * r + c] / den;
    }
  float UN[9], UNVt[9];
  mul_ab(U, Nm, UN);
  mul_abt(UN, V, UNVt);
#pragma unroll
  for (int i = 0; i < 9; ++i) Fb[i] += 2.0f * UNVt[i];
  // volumetric: lambda (2J - 1) tr(tb) dJ/dF, dJ/dF = cof(F)
  const float cv = lamp * (2.0f * J - 1.0f) * trt;
  const float cof[9] = {F[4] * F[8] - F[5] * F[7], F[5] * F[6] - F[3] * F[8], F[3] * F[7] - F[4] * F[6],
                        F[2] * F[7] - F[1] * F[8], F[0] * F[8] - F[2] * F[6], F[1] * F[6] - F[0] * F[7],
                        F[1] * F[5] - F[2] * F[4], F[2] * F[3] - F[0] * F[5], F[0] * F[4] - F[1] * F[3]};
#pragma unroll
  for (int i = 0; i < 9; ++i) Fb[i] += cv * cof[i];
  float Cb[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Cb[i] = d.mass * Ab[i];
  store_state(adj_in, n, p, xb, vb, Cb, Fb);
}

__global__ void __launch_bounds__(256) k_zero4(float4* __restrict__ g, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) g[t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

DiffK diff_consts(const StepConsts& k, float mass, float volume, int64_t n) {
  DiffK d;
  d.k = k;
  d.mass = mass;
  d.volume = volume;
  d.n = n;
  return d;
}

}  // namespace

// ---------------------------------------------------------------- no-allocation forward
// The scatter grid of the forward (4 floats per node) for callers that cannot
// allocate: one array per device, sized for grids up to kNoAllocMaxN^3 nodes,
// held until the process ends. Host code never uses the array's name (a
// host-side stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, once per device with that device current.
static __device__ __align__(256) float4 scratch[(size_t)mpmk::kNoAllocMaxN * mpmk::kNoAllocMaxN *
                                                mpmk::kNoAllocMaxN];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float4* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's copy of `scratch`.
hipError_t scratch_address(float4** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = static_cast<float4*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}
}  // namespace

namespace mpmk {

hipError_t diff_forward_noalloc(const StepConsts& k, float mass, float volume, int64_t n,
                                const float* st, float* out, const float* mu, const float* lam,
                                float* tape, hipStream_t s) {
  if (k.N < 1 || k.N > kNoAllocMaxN) return hipErrorInvalidValue;
  float4* grid = nullptr;
  const hipError_t e = scratch_address(&grid);
  if (e != hipSuccess) return e;
  // diff_forward clears the first 4 N^3 floats with a kernel before the
  // scatter, whatever the last call left there
  return diff_forward(k, mass, volume, n, st, out, mu, lam, grid, tape, s);
}

hipError_t diff_forward(const StepConsts& k, float mass, float volume, int64_t n, const float* st,
                        float* out, const float* mu, const float* lam, float4* grid, float* tape,
                        hipStream_t s) {
  const size_t nodes = (size_t)k.N * k.N * k.N;
  const DiffK d = diff_consts(k, mass, volume, n);
  const unsigned gp = (unsigned)((n + 255) / 256), gn = (unsigned)((nodes + 255) / 256);
  hipLaunchKernelGGL(k_zero4, dim3(gn), dim3(256), 0, s, grid, nodes);
  if (n > 0) hipLaunchKernelGGL(k_diff_p2g, dim3(gp), dim3(256), 0, s, st, mu, lam, d, (float*)grid);
  hipLaunchKernelGGL(k_diff_grid, dim3(gn), dim3(256), 0, s, (const float4*)grid, tape, k);
  if (n > 0) hipLaunchKernelGGL(k_diff_g2p, dim3(gp), dim3(256), 0, s, st, out, tape, d);
  return hipGetLastError();
}

hipError_t diff_backward(const StepConsts& k, float mass, float volume, int64_t n, const float* st,
                         const float* mu, const float* lam, const float* tape, const float* adj_out,
                         float* adj_in, float* adj_mu, float* adj_lam, float4* grid_adj,
                         hipStream_t s) {
  const size_t nodes = (size_t)k.N * k.N * k.N;
  const DiffK d = diff_consts(k, mass, volume, n);
  const unsigned gp = (unsigned)((n + 255) / 256), gn = (unsigned)((nodes + 255) / 256);
  hipLaunchKernelGGL(k_zero4, dim3(gn), dim3(256), 0, s, grid_adj, nodes);
  if (n > 0)
    hipLaunchKernelGGL(k_adj_g2p, dim3(gp), dim3(256), 0, s, st, tape, adj_out, adj_in, (float*)grid_adj, d);
  hipLaunchKernelGGL(k_adj_grid, dim3(gn), dim3(256), 0, s, grid_adj, tape, k);
  if (n > 0)
    hipLaunchKernelGGL(k_adj_p2g, dim3(gp), dim3(256), 0, s, st, mu, lam, (const float4*)grid_adj, adj_in,
                       adj_mu, adj_lam, d);
  return hipGetLastError();
}

}  // namespace mpmk
