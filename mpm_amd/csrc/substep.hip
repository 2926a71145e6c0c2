// substep.hip -- the MLS-MPM substep kernels for gfx950.
//
// One substep of the reference (Simulation::advance, src/mpm.cu:323-329) is
// clear grid -> particleToGrid (mpm.cu:14-74) -> gridOpKernel (:76-107) ->
// gridToParticle (:109-178). The engine runs the same arithmetic in this
// pipeline (steady state, two grids A and B):
//
//   k_grid_update(A -> velocities, B := 0)    streaming pass, 32 B/node + 16 B/node clear
//   k_g2p2g(A -> particles -> B)              G2P of substep n fused with the P2G of
//                                             substep n+1: one read and one write of
//                                             the particle state, one SVD per particle
//
// so the P2G scatter of substep n+1 is already in B when substep n ends. The
// unfused reference order (k_p2g, k_grid_update, k_g2p) is kept for the first
// substep after an upload and as an option.
//
// P2G scatter variants (option "p2g_mode"):
//   0 direct   27 nodes x 4 FP32 global atomics per particle.
//   1 tiled    particles are processed in cell-sorted order, one workgroup per
//              work item of one grid block (p2g_block.hip): G2P from an LDS
//              velocity tile, P2G as an in-tile gather, one global atomic per
//              float of each non-empty tile node.
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "engine_types.hpp"
#include "kernels.hpp"
#include "plugins.hpp"

using namespace mpmdev;
using mpmplug::CubicScheme;
using mpmplug::ReferenceScheme;

namespace {

__device__ __forceinline__ void load_state(const ParticlesDev& P, int i, float x[3], float v[3],
                                           float C[9], float F[9], float& Jp) {
  const float* f = P.f;
  const size_t s = P.stride;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    x[a] = f[(PL_X + a) * s + i];
    v[a] = f[(PL_V + a) * s + i];
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    C[a] = f[(PL_C + a) * s + i];
    F[a] = f[(PL_F + a) * s + i];
  }
  Jp = f[PL_JP * s + i];
}

__device__ __forceinline__ void store_state(const ParticlesDev& P, int i, const float x[3],
                                            const float v[3], const float C[9], const float F[9],
                                            float Jp) {
  float* f = P.f;
  const size_t s = P.stride;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    f[(PL_X + a) * s + i] = x[a];
    f[(PL_V + a) * s + i] = v[a];
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    f[(PL_C + a) * s + i] = C[a];
    f[(PL_F + a) * s + i] = F[a];
  }
  f[PL_JP * s + i] = Jp;
}

__device__ __forceinline__ void atomic_add_f32(float* p, float v) {
  // gfx950 has a native FP32 global atomic add; the unsafe form maps to it
  // without a compare-and-swap loop.
  unsafeAtomicAdd(p, v);
}

// Scatter one particle's stencil contributions (p2g_node_contribution,
// TransferScheme.h:88-100) with global atomics. Reference node bounds
// (mpm.cu:39-44): begin max(0, -base); end min(S, N+base) in compat REFERENCE
// (j, k >= N alias into the next row of the linear index; indices >= N^3 are
// dropped), min(S, N-base) in compat FIXED.
template <class Sch>
__device__ __forceinline__ void scatter_direct(const float x[3], const float v[3],
                                               const float aff[9], float mass, const Sch& sch,
                                               const StepConsts& k, float* __restrict__ grid) {
  constexpr int S = Sch::S;
  const int* base = sch.base;
  const auto& w = sch.w;
  const int N = k.N;
  int beg[3], end[3];
  float d[3][S];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    beg[a] = max(0, -base[a]);
    end[a] = min(S, (k.compat == MPM_COMPAT_FIXED) ? N - base[a] : N + base[a]);
#pragma unroll
    for (int o = 0; o < S; ++o) d[a][o] = (float)(uint32_t)(base[a] + o) * k.dx - x[a];
  }
  const float mv0 = v[0] * mass, mv1 = v[1] * mass, mv2 = v[2] * mass;
  const uint64_t n3 = (uint64_t)N * N * N;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    if (i < beg[0] || i >= end[0]) continue;
#pragma unroll
    for (int j = 0; j < S; ++j) {
      if (j < beg[1] || j >= end[1]) continue;
      const float wij = w[0][i] * w[1][j];
#pragma unroll
      for (int kk = 0; kk < S; ++kk) {
        if (kk < beg[2] || kk >= end[2]) continue;
        const uint32_t idx = (uint32_t)N * (uint32_t)N * (uint32_t)(base[0] + i) +
                             (uint32_t)N * (uint32_t)(base[1] + j) + (uint32_t)(base[2] + kk);
        if ((uint64_t)idx >= n3) continue;
        const float wt = wij * w[2][kk];
        const float ad0 = aff[0] * d[0][i] + (aff[1] * d[1][j] + aff[2] * d[2][kk]);
        const float ad1 = aff[3] * d[0][i] + (aff[4] * d[1][j] + aff[5] * d[2][kk]);
        const float ad2 = aff[6] * d[0][i] + (aff[7] * d[1][j] + aff[8] * d[2][kk]);
        float* cell = grid + 4 * (size_t)idx;
        atomic_add_f32(cell + 0, wt * (mv0 + ad0));
        atomic_add_f32(cell + 1, wt * (mv1 + ad1));
        atomic_add_f32(cell + 2, wt * (mv2 + ad2));
        atomic_add_f32(cell + 3, wt * mass);
      }
    }
  }
}

// G2P gather (mpm.cu:135-167; g2p_node_contribution, TransferScheme.h:120-132)
// into v and C: v += w v_i, C += (w v_i)(d^T D^-1).
template <class Sch>
__device__ __forceinline__ void gather(const float x[3], const Sch& sch, const StepConsts& k,
                                       const float4* __restrict__ grid, float v[3], float C[9]) {
  constexpr int S = Sch::S;
  const int* base = sch.base;
  const auto& w = sch.w;
  const int N = k.N;
  float dd[3][S];  // constant D^-1: (d * dinv) per axis; generic: the offset d
  bool ok[3][S];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = 0; o < S; ++o) {
      const int g = base[a] + o;
      ok[a][o] = g >= 0 && g < N;
      if constexpr (Sch::Kernel::d_is_const())
        dd[a][o] = ((float)(uint32_t)g * k.dx - x[a]) * k.dinv;
      else
        dd[a][o] = (float)(uint32_t)g * k.dx - x[a];
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) v[a] = 0.0f;
#pragma unroll
  for (int a = 0; a < 9; ++a) C[a] = 0.0f;
#pragma unroll
  for (int i = 0; i < S; ++i) {
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const float wij = w[0][i] * w[1][j];
      const size_t row = ((size_t)(base[0] + i) * N + (size_t)(base[1] + j)) * N;
#pragma unroll
      for (int kk = 0; kk < S; ++kk) {
        if (!(ok[0][i] && ok[1][j] && ok[2][kk])) continue;
        const float4 g = grid[row + (size_t)(base[2] + kk)];
        const float wt = wij * w[2][kk];
        const float wv0 = wt * g.x, wv1 = wt * g.y, wv2 = wt * g.z;
        v[0] += wv0;
        v[1] += wv1;
        v[2] += wv2;
        float r0, r1, r2;  // d^T D^-1
        if constexpr (Sch::Kernel::d_is_const()) {
          r0 = dd[0][i];
          r1 = dd[1][j];
          r2 = dd[2][kk];
        } else {
          float r[3];
          sch.dinv_rows(dd[0][i], dd[1][j], dd[2][kk], k, r);
          r0 = r[0];
          r1 = r[1];
          r2 = r[2];
        }
        C[0] += wv0 * r0; C[1] += wv0 * r1; C[2] += wv0 * r2;
        C[3] += wv1 * r0; C[4] += wv1 * r1; C[5] += wv1 * r2;
        C[6] += wv2 * r0; C[7] += wv2 * r1; C[8] += wv2 * r2;
      }
    }
  }
}

// G2P of one particle already loaded: gather, F update (g2p_finish_particle),
// plasticity, advection. Returns false when the particle is outside the
// domain (left untouched).
template <class Sch>
__device__ __forceinline__ bool g2p_update(float x[3], float v[3], float C[9], float F[9],
                                           float& Jp, const MaterialDev& m, const StepConsts& k,
                                           const float4* __restrict__ grid, float R[9],
                                           bool& has_R) {
  Sch sch;
  has_R = false;
  if (!sch.prepare(x, k)) return false;
  gather(x, sch, k, grid, v, C);
  // F = (I + dt C) F (TransferScheme.h:141)
  float A[9];
#pragma unroll
  for (int a = 0; a < 9; ++a) A[a] = ((a % 4 == 0) ? 1.0f : 0.0f) + k.dt * C[a];
  mul_ab(A, F, F);
  plasticity(F, Jp, m, R, has_R);
#pragma unroll
  for (int a = 0; a < 3; ++a) x[a] = x[a] + k.dt * v[a];
  return true;
}

// P2G of one particle (p2g_prepare_particle + node loop).
template <class Sch>
__device__ __forceinline__ void p2g_particle(const float x[3], const float v[3], const float C[9],
                                             const float F[9], float Jp, const MaterialDev& m,
                                             const StepConsts& k, const float* Rin, bool has_R,
                                             float* __restrict__ grid) {
  Sch sch;
  if (!sch.prepare(x, k)) return;
  float R[9];
  if (m.kind != MPM_MATERIAL_WC_LIQUID) {
    if (has_R) {
#pragma unroll
      for (int a = 0; a < 9; ++a) R[a] = Rin[a];
    } else {
      polar_rotation(F, R);
    }
  }
  float tau[9], aff[9];
  kirchhoff_stress(F, Jp, R, m, k.compat, tau);
  sch.affine(tau, C, m, k, aff);
  scatter_direct(x, v, aff, m.mass, sch, k, grid);
}

}  // namespace

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fill_zero(float4* __restrict__ g, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t t = i; t < n; t += step) g[t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

template <class Sch>
__global__ void __launch_bounds__(256)
    k_p2g_direct(ParticlesDev P, const MaterialDev* __restrict__ mats, StepConsts k,
                 float* __restrict__ grid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  float x[3], v[3], C[9], F[9], Jp;
  load_state(P, i, x, v, C, F, Jp);
  const MaterialDev m = mats[P.mat[i]];
  p2g_particle<Sch>(x, v, C, F, Jp, m, k, nullptr, false, grid);
}

// gridOpKernel: normalise, gravity, walls. Optionally zeroes `clear` (the
// other grid) in the same pass.
__global__ void __launch_bounds__(256)
    k_grid_update(float4* __restrict__ g, float4* __restrict__ clear, StepConsts k) {
  const int N = k.N;
  const size_t n = (size_t)N * N * N;
  const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t t = t0; t < n; t += step) {
    if (clear) clear[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 c = g[t];
    if (!(c.w > 0.0f)) continue;
    const uint32_t u = (uint32_t)t, nu = (uint32_t)N;  // (N <= 1024: a node index is below 2^30)
    const int zi = (int)(u % nu);
    const int yi = (int)((u / nu) % nu);
    const int xi = (int)(u / (nu * nu));
    g[t] = grid_node_update(c, xi, yi, zi, k);
  }
}

template <class Sch>
__global__ void __launch_bounds__(256)
    k_g2p(ParticlesDev P, const MaterialDev* __restrict__ mats, StepConsts k,
          const float4* __restrict__ grid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  if (P.id[i] > k.guard) return;  // mpm.cu:114 (compat REFERENCE)
  float x[3], v[3], C[9], F[9], Jp;
  load_state(P, i, x, v, C, F, Jp);
  const MaterialDev m = mats[P.mat[i]];
  float R[9];
  bool has_R;
  if (!g2p_update<Sch>(x, v, C, F, Jp, m, k, grid, R, has_R)) return;
  store_state(P, i, x, v, C, F, Jp);
}

// Fused G2P(n) + P2G(n+1) with direct global atomics.
template <class Sch>
__global__ void __launch_bounds__(256)
    k_g2p2g_direct(ParticlesDev P, const MaterialDev* __restrict__ mats, StepConsts k,
                   const float4* __restrict__ grid_in, float* __restrict__ grid_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  float x[3], v[3], C[9], F[9], Jp;
  load_state(P, i, x, v, C, F, Jp);
  const MaterialDev m = mats[P.mat[i]];
  float R[9];
  bool has_R = false;
  if (P.id[i] <= k.guard) {
    if (g2p_update<Sch>(x, v, C, F, Jp, m, k, grid_in, R, has_R)) store_state(P, i, x, v, C, F, Jp);
  }
  p2g_particle<Sch>(x, v, C, F, Jp, m, k, R, has_R, grid_out);
}

// ---------------------------------------------------------------------------
// launchers (host)
// ---------------------------------------------------------------------------
namespace mpmk {

static inline unsigned blocks_for(size_t n, unsigned threads, unsigned cap) {
  size_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

hipError_t fill_zero(float4* g, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, s, g, n);
  return hipGetLastError();
}

hipError_t p2g_direct(const ParticlesDev& P, const MaterialDev* mats, const StepConsts& k,
                      float4* grid, hipStream_t s, int interp) {
  if (P.n <= 0) return hipSuccess;
  if (interp == MPM_INTERP_CUBIC)
    hipLaunchKernelGGL(k_p2g_direct<CubicScheme>, dim3((P.n + 255) / 256), dim3(256), 0, s, P, mats, k,
                       (float*)grid);
  else
    hipLaunchKernelGGL(k_p2g_direct<ReferenceScheme>, dim3((P.n + 255) / 256), dim3(256), 0, s, P, mats, k,
                       (float*)grid);
  return hipGetLastError();
}

hipError_t grid_update(float4* g, float4* clear, const StepConsts& k, hipStream_t s) {
  const size_t n = (size_t)k.N * k.N * k.N;
  hipLaunchKernelGGL(k_grid_update, dim3(blocks_for(n, 256, 16384)), dim3(256), 0, s, g, clear, k);
  return hipGetLastError();
}

hipError_t g2p(const ParticlesDev& P, const MaterialDev* mats, const StepConsts& k,
               const float4* grid, hipStream_t s, int interp) {
  if (P.n <= 0) return hipSuccess;
  if (interp == MPM_INTERP_CUBIC)
    hipLaunchKernelGGL(k_g2p<CubicScheme>, dim3((P.n + 255) / 256), dim3(256), 0, s, P, mats, k, grid);
  else
    hipLaunchKernelGGL(k_g2p<ReferenceScheme>, dim3((P.n + 255) / 256), dim3(256), 0, s, P, mats, k, grid);
  return hipGetLastError();
}

hipError_t g2p2g_direct(const ParticlesDev& P, const MaterialDev* mats, const StepConsts& k,
                        const float4* grid_in, float4* grid_out, hipStream_t s, int interp) {
  if (P.n <= 0) return hipSuccess;
  if (interp == MPM_INTERP_CUBIC)
    hipLaunchKernelGGL(k_g2p2g_direct<CubicScheme>, dim3((P.n + 255) / 256), dim3(256), 0, s, P, mats, k,
                       grid_in, (float*)grid_out);
  else
    hipLaunchKernelGGL(k_g2p2g_direct<ReferenceScheme>, dim3((P.n + 255) / 256), dim3(256), 0, s, P, mats,
                       k, grid_in, (float*)grid_out);
  return hipGetLastError();
}

}  // namespace mpmk
