// adjoint.hip -- a differentiable MLS-MPM substep and its hand-written adjoint
// (gfx950). No reference counterpart: the reference has no gradients.
//
// The substep is the engine's (P2G, grid update, G2P: reference src/mpm.cu:
// 14-178) for the fixed-corotated model (MMFixedCorotated, MaterialModel.cuh:
// 56-61) with J = det F, per-particle Lame parameters mu_p, lambda_p, no
// plasticity, and the stencil clipped at the grid faces (N - base). State
// planes (24 x n floats, plane-major): x(3) v(3) C(9) F(9), 3x3 row-major.
// The forward records a grid tape per node that the backward replays: the
// velocity and mass after the update (N^3 float4, read by G2P), then the
// velocity before the walls and the wall flags (N^3 float4).
//
// Kernels, one particle per lane, one wavefront per workgroup:
//   k_diff_step<G2P, P2G>   G2P of substep n (gather of the tape velocities,
//                           F and x update) fused with the P2G of substep n+1
//                           (stress through the polar rotation of F, scatter)
//   k_diff_grid             grid update into the tape; clears the grid
//   k_adj_step<P2G', G2P'>  the adjoint of the P2G of substep s (gather of the
//                           momentum and mass adjoints, stress adjoint through
//                           the SVD, mu and lambda adjoints) fused with the
//                           adjoint of the G2P of substep s-1 (scatter of the
//                           grid velocity adjoint)
//   k_adj_grid              grid update backward: momentum and mass adjoints
// so a substep is two launches forward and two backward, and a run of
// substeps is enqueued from one host call (diff_forward_steps /
// diff_backward_steps).
//
// The stress adjoint differentiates the polar rotation R of F = U S V^T:
// with Rbar the adjoint of R and M = V^T skew(R^T Rbar) V,
// Fbar += 2 U N V^T, N_ij = M_ij / (s_i + s_j).
//
// Scatters (P2G: 27 nodes x 4 values; the grid adjoint of G2P: 27 x 3): each
// lane writes its particle's values into a row of an LDS table; then each
// lane takes a column and walks the rows in lane order, summing runs -- a run
// is a stretch of lanes whose particles share a base cell (particles sorted by
// cell make long runs: ~37 per cell for a dense lattice) -- and adding each
// run's sum to the grid with one FP32 atomic. A dense scene issues ~3 runs x
// 108 atomics per wavefront instead of 64 x 108; unsorted particles only give
// shorter runs.
#include <hip/hip_runtime.h>

#include <mutex>

#include "device_math.hpp"
#include "engine_types.hpp"
#include "kernels.hpp"

using namespace mpmdev;

namespace {

enum : int { DX = 0, DV = 3, DC = 6, DF = 15 };

constexpr int kWave = 64;        // lanes per workgroup (one wavefront)
constexpr int kP2GVals = 108;    // 27 nodes x (momentum, mass)
constexpr int kP2GStride = 109;  // odd LDS row stride: a column spreads over the banks
constexpr int kAdjVals = 81;     // 27 nodes x grid velocity adjoint
constexpr int kAdjStride = 83;

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

// the base cell of a particle inside the domain (base in [-3, N) per axis)
// as one non-negative integer
__device__ __forceinline__ int cell_key(const Stencil& s, int N) {
  const int M = N + 3;
  return ((s.base[0] + 3) * M + (s.base[1] + 3)) * M + (s.base[2] + 3);
}

__device__ __forceinline__ void load_xF(const float* st, int64_t n, int64_t p, float x[3], float F[9]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) x[a] = st[(DX + a) * n + p];
#pragma unroll
  for (int a = 0; a < 9; ++a) F[a] = st[(DF + a) * n + p];
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

// ---------------------------------------------------------------- run scatter
// Adds the LDS table rows (NV values per lane, row stride STRIDE; VPN values
// per node, stencil node o = 9a + 3b + c at columns VPN o .. VPN o + VPN - 1)
// to `grid` (4 floats per node): one sum per (run, column), in lane order,
// and one FP32 atomic per non-zero sum of a node inside the grid. A run is a
// stretch of lanes with the same key (key < 0: the lane adds nothing). Each
// lane takes a column and walks the rows in lane order, 16 rows loaded at a
// time; the run ends are the same for every column, so the walk branches
// uniformly and a run's atomics for 64 columns are one instruction. Every
// lane of the (one-wavefront) workgroup calls it.
template <int NV, int STRIDE, int VPN>
__device__ __forceinline__ void run_scatter(const float* table, int key, const int base[3], int N,
                                            float* __restrict__ grid) {
  const int lane = threadIdx.x;
  const int next = __shfl_down(key, 1);
  const uint64_t ends = __ballot(lane == kWave - 1 || key != next);
  const uint64_t live = __ballot(key >= 0);
  __syncthreads();  // the table rows are visible to every lane
  for (int c0 = 0; c0 < NV; c0 += kWave) {
    const int q = c0 + lane;
    const bool qok = q < NV;
    const int qc = qok ? q : 0;  // (a column past the table reads column 0, unused)
    const int o = qc / VPN, comp = qc - (qc / VPN) * VPN;
    const int da = o / 9, db = (o / 3) % 3, dc = o % 3;
    float sum = 0.0f;
#pragma unroll
    for (int p0 = 0; p0 < kWave; p0 += 16) {
      float v[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = table[(p0 + i) * STRIDE + qc];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int p = p0 + i;
        if ((live >> p) & 1ull) sum += v[i];  // (a row of no particle is never read)
        if ((ends >> p) & 1ull) {
          if (((live >> p) & 1ull) && qok && sum != 0.0f) {
            const int gi = __builtin_amdgcn_readlane(base[0], p) + da;
            const int gj = __builtin_amdgcn_readlane(base[1], p) + db;
            const int gk = __builtin_amdgcn_readlane(base[2], p) + dc;
            if ((unsigned)gi < (unsigned)N && (unsigned)gj < (unsigned)N && (unsigned)gk < (unsigned)N)
              unsafeAtomicAdd(grid + 4 * (((size_t)gi * N + (size_t)gj) * N + (size_t)gk) + comp, sum);
          }
          sum = 0.0f;
        }
      }
    }
  }
}

// v' = sum_i w_i v_i and C' = Dinv sum_i w_i v_i (x_i - x_p)^T over the stencil
__device__ __forceinline__ void g2p_affine(const Stencil& sn, const float x[3], const float4* vel,
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
        const float4 gv = vel[node_index(sn, a, b, c, k.N)];
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

// ---------------------------------------------------------------- forward
// G2P of substep n from `st` with the tape velocities `vel` into `out`
// (particles outside the grid are copied), fused with the P2G of substep n+1
// of the new state into `grid`. Without G2P: the P2G of `st`. Without P2G:
// the G2P only.
template <bool G2P, bool P2G>
__global__ void __launch_bounds__(kWave)
    k_diff_step(const float* __restrict__ st, float* __restrict__ out, const float4* __restrict__ vel,
                const float* __restrict__ mu, const float* __restrict__ lam, DiffK d,
                float* __restrict__ grid) {
  __shared__ float table[P2G ? kWave * kP2GStride : 1];
  const int lane = threadIdx.x;
  const int64_t p = (int64_t)blockIdx.x * kWave + lane;
  const StepConsts& k = d.k;
  const bool valid = p < d.n;
  float x[3] = {0.f, 0.f, 0.f}, v[3] = {0.f, 0.f, 0.f}, C[9], F[9];
  Stencil sn;
  sn.base[0] = sn.base[1] = sn.base[2] = 0;
  bool inside = false;
  if (valid) {
    load_state(st, d.n, p, x, v, C, F);
    sn = make_stencil(x, k);
    inside = !sn.outside;
  }
  if (G2P) {
    if (inside) {
      float vn[3], Cn[9];
      g2p_affine(sn, x, vel, k, vn, Cn);
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
    if (valid) store_state(out, d.n, p, x, v, C, F);
    if (P2G && valid) {
      sn = make_stencil(x, k);
      inside = !sn.outside;
    }
  }
  if (P2G) {
    int key = -1;
    if (inside) {
      float U[9], s[3], V[9], R[9], tau[9];
      fc_stress(F, mu[p], lam[p], U, s, V, R, tau);
      const float sc = stress_scale(d);
      float A[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) A[i] = sc * tau[i] + d.mass * C[i];
      float* row = table + lane * kP2GStride;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int o = 4 * (9 * a + 3 * b + c);
            const float wt = weight(sn, a, b, c);
            float dp[3];
            node_offset(sn, a, b, c, x, k, dp);
#pragma unroll
            for (int r = 0; r < 3; ++r)
              row[o + r] = wt * (d.mass * v[r] + (A[3 * r] * dp[0] + A[3 * r + 1] * dp[1] + A[3 * r + 2] * dp[2]));
            row[o + 3] = wt * d.mass;
          }
      key = cell_key(sn, k.N);
    }
    run_scatter<kP2GVals, kP2GStride, 4>(table, key, sn.base, k.N, grid);
  }
}

// grid update into the tape: tv = (velocity, mass) after the update, tw =
// (velocity before the walls, wall flags: 1 sticky, 2 floor); clears the
// grid for the next scatter
__global__ void __launch_bounds__(256)
    k_diff_grid(float4* __restrict__ grid, float4* __restrict__ tv, float4* __restrict__ tw, StepConsts k) {
  const int N = k.N;
  const size_t nodes = (size_t)N * N * N;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nodes) return;
  const float4 c = grid[t];
  grid[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 o0 = make_float4(0.f, 0.f, 0.f, c.w), o1 = make_float4(0.f, 0.f, 0.f, 0.f);
  if (c.w > 0.0f) {
    // (32-bit: N <= 1024, so a node index is below 2^30)
    const uint32_t ti = (uint32_t)t, n = (uint32_t)N;
    const int zi = (int)(ti % n), yi = (int)((ti / n) % n), xi = (int)(ti / (n * n));
    o1.x = c.x / c.w;
    o1.y = __builtin_fmaf(k.dt, k.gravity, c.y / c.w);  // (as grid_node_update rounds it)
    o1.z = c.z / c.w;
    const float4 u = grid_node_update(c, xi, yi, zi, k);
    bool sticky, floor_;
    wall_flags(xi, yi, zi, k, sticky, floor_);
    o0 = make_float4(u.x, u.y, u.z, c.w);  // the mass, not the update's w (1)
    o1.w = (float)((sticky ? 1 : 0) | (floor_ ? 2 : 0));
  }
  tv[t] = o0;
  tw[t] = o1;
}

// ---------------------------------------------------------------- backward
// The adjoint of the P2G of substep s (P2GA: the momentum and mass adjoints
// `gm` of that substep, gathered) fused with the adjoint of the G2P of
// substep s-1 (G2PA: the grid velocity adjoint scattered into `gv_out`).
//   adj_src   P2GA: the partial adjoint of X_s the previous G2PA left (x and F
//             parts; the whole adjoint for particles outside the grid);
//             otherwise the full adjoint of X_s.
//   adj_dst   G2PA: the partial adjoint of X_{s-1} (may be adj_src);
//             otherwise the full adjoint of X_s.
// X_s = st_s (state before substep s), X_{s-1} = st_prev with tape velocities
// vel_prev. The mu and lambda adjoints are added into adj_mu, adj_lam.
template <bool P2GA, bool G2PA>
__global__ void __launch_bounds__(kWave)
    k_adj_step(const float* __restrict__ st_s, const float* __restrict__ mu, const float* __restrict__ lam,
               const float4* __restrict__ gm, const float* adj_src, const float* __restrict__ st_prev,
               const float4* __restrict__ vel_prev, float* adj_dst, float* __restrict__ adj_mu,
               float* __restrict__ adj_lam, float* __restrict__ gv_out, DiffK d) {
  __shared__ float table[G2PA ? kWave * kAdjStride : 1];
  const int lane = threadIdx.x;
  const int64_t p = (int64_t)blockIdx.x * kWave + lane;
  const StepConsts& k = d.k;
  const int64_t n = d.n;
  const bool valid = p < n;
  // adjoint of X_s
  float xb[3] = {0.f, 0.f, 0.f}, vb[3] = {0.f, 0.f, 0.f}, Cb[9], Fb[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Cb[i] = Fb[i] = 0.f;
  if (valid) load_state(adj_src, n, p, xb, vb, Cb, Fb);

  if (P2GA && valid) {
    float x[3], v[3], C[9], F[9];
    load_state(st_s, n, p, x, v, C, F);
    const Stencil sn = make_stencil(x, k);
    if (!sn.outside) {  // outside: the partial already holds the identity map
      const float mup = mu[p], lamp = lam[p];
      float U[9], s[3], V[9], R[9], tau[9];
      fc_stress(F, mup, lamp, U, s, V, R, tau);
      const float sc = stress_scale(d);
      float A[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) A[i] = sc * tau[i] + d.mass * C[i];
      float Ab[9];
#pragma unroll
      for (int i = 0; i < 3; ++i) vb[i] = 0.f;
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
            const float4 g = gm[node_index(sn, a, b, c, k.N)];
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
#pragma unroll
      for (int i = 0; i < 9; ++i) Cb[i] = d.mass * Ab[i];
    }
  }
  // (xb, vb, Cb, Fb) is the adjoint of X_s
  if (!G2PA) {
    if (valid) store_state(adj_dst, n, p, xb, vb, Cb, Fb);
    return;
  }
  // G2P of substep s-1: X_s = G2P(X_{s-1}); the adjoint of X_s flows back
  // to x and F of X_{s-1} and to the grid velocities of substep s-1
  int key = -1;
  Stencil sn;
  sn.base[0] = sn.base[1] = sn.base[2] = 0;
  if (valid) {
    float x[3], F[9];
    load_xF(st_prev, n, p, x, F);
    sn = make_stencil(x, k);
    if (sn.outside) {
      store_state(adj_dst, n, p, xb, vb, Cb, Fb);  // the identity map
    } else {
      float vn[3], Cn[9];
      g2p_affine(sn, x, vel_prev, k, vn, Cn);
      // x' = x + dt v', F' = (I + dt C') F: adjoints of v' and C' through them
      float vbt[3], Cbt[9], FbFt[9];
#pragma unroll
      for (int a = 0; a < 3; ++a) vbt[a] = vb[a] + k.dt * xb[a];
      mul_abt(Fb, F, FbFt);
#pragma unroll
      for (int i = 0; i < 9; ++i) Cbt[i] = Cb[i] + k.dt * FbFt[i];
      float xo[3] = {xb[0], xb[1], xb[2]}, Fo[9];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          // ((I + dt C')^T Fb')_rc
          float acc = 0.f;
#pragma unroll
          for (int q = 0; q < 3; ++q) acc += (((q == r) ? 1.0f : 0.0f) + k.dt * Cn[3 * q + r]) * Fb[3 * q + c];
          Fo[3 * r + c] = acc;
        }
      float* row = table + lane * kAdjStride;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int o = 3 * (9 * a + 3 * b + c);
            if (!node_ok(sn, a, b, c, k.N)) {
              row[o] = row[o + 1] = row[o + 2] = 0.0f;
              continue;
            }
            const float wt = weight(sn, a, b, c);
            float gw[3], dp[3];
            weight_grad(sn, a, b, c, gw);
            node_offset(sn, a, b, c, x, k, dp);
            const float4 gv4 = vel_prev[node_index(sn, a, b, c, k.N)];
            const float gv[3] = {gv4.x, gv4.y, gv4.z};
            float Cdp[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) Cdp[r] = Cbt[3 * r] * dp[0] + Cbt[3 * r + 1] * dp[1] + Cbt[3 * r + 2] * dp[2];
            const float s1 = gv[0] * vbt[0] + gv[1] * vbt[1] + gv[2] * vbt[2];
            const float s2 = gv[0] * Cdp[0] + gv[1] * Cdp[1] + gv[2] * Cdp[2];
            const float coef = s1 + k.dinv * s2;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              const float ctg = Cbt[q] * gv[0] + Cbt[3 + q] * gv[1] + Cbt[6 + q] * gv[2];  // (C'^T g)_q
              xo[q] += gw[q] * coef - (k.dinv * wt) * ctg;
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) row[o + r] = wt * (vbt[r] + k.dinv * Cdp[r]);
          }
#pragma unroll
      for (int a = 0; a < 3; ++a) adj_dst[(DX + a) * n + p] = xo[a];
#pragma unroll
      for (int i = 0; i < 9; ++i) adj_dst[(DF + i) * n + p] = Fo[i];
      key = cell_key(sn, k.N);
    }
  }
  run_scatter<kAdjVals, kAdjStride, 3>(table, key, sn.base, k.N, gv_out);
}

// adjoint of the updated velocity (gadj.xyz) -> adjoints of the P2G momentum
// and mass (gadj, in place); clears `clear` (the next scatter's target)
__global__ void __launch_bounds__(256)
    k_adj_grid(float4* __restrict__ gadj, float4* __restrict__ clear, const float4* __restrict__ tv,
               const float4* __restrict__ tw, StepConsts k) {
  const size_t nodes = (size_t)k.N * k.N * k.N;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nodes) return;
  if (clear) clear[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 t0 = tv[t];
  const float4 t1 = tw[t];
  const float m = t0.w;
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  if (m > 0.0f) {
    const float4 g = gadj[t];
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
  gadj[t] = out;
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

size_t nodes_of(const StepConsts& k) { return (size_t)k.N * k.N * k.N; }
unsigned node_blocks(const StepConsts& k) { return (unsigned)((nodes_of(k) + 255) / 256); }
unsigned particle_blocks(int64_t n) { return (unsigned)((n + kWave - 1) / kWave); }

// the two halves of a tape slot: velocities (read by G2P), then the rest
const float4* tape_v(const float* tape) { return reinterpret_cast<const float4*>(tape); }
const float4* tape_w(const float* tape, const StepConsts& k) {
  return reinterpret_cast<const float4*>(tape) + nodes_of(k);
}
float4* tape_v(float* tape) { return reinterpret_cast<float4*>(tape); }
float4* tape_w(float* tape, const StepConsts& k) { return reinterpret_cast<float4*>(tape) + nodes_of(k); }

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


hipError_t diff_forward_steps(const StepConsts& k, float mass, float volume, int64_t n, int32_t steps,
                              const float* st0, float* states, int64_t state_stride, const float* mu,
                              const float* lam, float4* grid, float* tapes, int64_t tape_stride,
                              hipStream_t s) {
  const DiffK d = diff_consts(k, mass, volume, n);
  const unsigned gp = particle_blocks(n), gn = node_blocks(k);
  hipLaunchKernelGGL(k_zero4, dim3(gn), dim3(256), 0, s, grid, nodes_of(k));
  if (n > 0)
    hipLaunchKernelGGL((k_diff_step<false, true>), dim3(gp), dim3(kWave), 0, s, st0, (float*)nullptr,
                       (const float4*)nullptr, mu, lam, d, (float*)grid);
  for (int32_t j = 0; j < steps; ++j) {
    float* tape = tapes + (size_t)j * (size_t)tape_stride;
    hipLaunchKernelGGL(k_diff_grid, dim3(gn), dim3(256), 0, s, grid, tape_v(tape), tape_w(tape, k), k);
    if (n == 0) continue;
    const float* in = j == 0 ? st0 : states + (size_t)(j - 1) * (size_t)state_stride;
    float* out = states + (size_t)j * (size_t)state_stride;
    if (j + 1 < steps)
      hipLaunchKernelGGL((k_diff_step<true, true>), dim3(gp), dim3(kWave), 0, s, in, out,
                         (const float4*)tape_v(tape), mu, lam, d, (float*)grid);
    else
      hipLaunchKernelGGL((k_diff_step<true, false>), dim3(gp), dim3(kWave), 0, s, in, out,
                         (const float4*)tape_v(tape), mu, lam, d, (float*)nullptr);
  }
  return hipGetLastError();
}

hipError_t diff_forward(const StepConsts& k, float mass, float volume, int64_t n, const float* st,
                        float* out, const float* mu, const float* lam, float4* grid, float* tape,
                        hipStream_t s) {
  return diff_forward_steps(k, mass, volume, n, 1, st, out, 0, mu, lam, grid, tape, 0, s);
}

hipError_t diff_backward_steps(const StepConsts& k, float mass, float volume, int64_t n, int32_t steps,
                               const float* states, int64_t state_stride, const float* mu,
                               const float* lam, const float* tapes, int64_t tape_stride,
                               const float* adj_out, float* adj_in, float* adj_mu, float* adj_lam,
                               float4* grid2, float* work, hipStream_t s) {
  const DiffK d = diff_consts(k, mass, volume, n);
  const unsigned gp = particle_blocks(n), gn = node_blocks(k);
  const size_t nodes = nodes_of(k);
  float4* ga = grid2;
  float4* gb = steps > 1 ? grid2 + nodes : nullptr;  // a single substep needs one grid
  auto state = [&](int32_t j) { return states + (size_t)j * (size_t)state_stride; };
  auto tape = [&](int32_t j) { return tapes + (size_t)j * (size_t)tape_stride; };
  hipLaunchKernelGGL(k_zero4, dim3(gn), dim3(256), 0, s, ga, nodes);
  // the G2P adjoint of the last substep, from the caller's adjoint
  if (n > 0)
    hipLaunchKernelGGL((k_adj_step<false, true>), dim3(gp), dim3(kWave), 0, s, (const float*)nullptr,
                       (const float*)nullptr, (const float*)nullptr, (const float4*)nullptr, adj_out,
                       state(steps - 1), tape_v(tape(steps - 1)), work, (float*)nullptr, (float*)nullptr,
                       (float*)ga, d);
  for (int32_t j = steps - 1; j >= 0; --j) {
    hipLaunchKernelGGL(k_adj_grid, dim3(gn), dim3(256), 0, s, ga, j > 0 ? gb : (float4*)nullptr,
                       tape_v(tape(j)), tape_w(tape(j), k), k);
    if (n == 0) continue;
    if (j > 0) {
      hipLaunchKernelGGL((k_adj_step<true, true>), dim3(gp), dim3(kWave), 0, s, state(j), mu, lam,
                         (const float4*)ga, (const float*)work, state(j - 1), tape_v(tape(j - 1)), work, adj_mu,
                         adj_lam, (float*)gb, d);
      float4* t = ga;
      ga = gb;
      gb = t;
    } else {
      hipLaunchKernelGGL((k_adj_step<true, false>), dim3(gp), dim3(kWave), 0, s, state(0), mu, lam,
                         (const float4*)ga, (const float*)work, (const float*)nullptr, (const float4*)nullptr,
                         adj_in, adj_mu, adj_lam, (float*)nullptr, d);
    }
  }
  return hipGetLastError();
}

hipError_t diff_backward(const StepConsts& k, float mass, float volume, int64_t n, const float* st,
                         const float* mu, const float* lam, const float* tape, const float* adj_out,
                         float* adj_in, float* adj_mu, float* adj_lam, float4* grid_adj, hipStream_t s) {
  // one substep: one grid, and adj_in holds the partial adjoint in between
  return diff_backward_steps(k, mass, volume, n, 1, st, 0, mu, lam, tape, 0, adj_out, adj_in, adj_mu,
                             adj_lam, grid_adj, adj_in, s);
}

}  // namespace mpmk
