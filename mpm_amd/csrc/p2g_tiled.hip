// p2g_tiled.hip -- fused G2P + LDS-tiled P2G over cell-sorted particles (gfx950).
//
// One workgroup of B threads takes B consecutive particles of the cell-sorted
// order (sort.hip), one per thread:
//   1. G2P of substep n (gather from grid_in, C/F update, plasticity with the
//      SVD in registers, advection) and the store of the new state
//      (gridToParticle, reference src/mpm.cu:109-178);
//   2. the P2G of substep n+1 for the new state (particleToGrid, mpm.cu:14-74):
//      the workgroup reduces the bounding box of its particles' base cells,
//      zeroes an LDS tile covering the 3x3x3 stencils of that box, and adds
//      each particle's 27 node contributions into the tile with LDS FP32
//      atomics; contributions of particles sharing a cell are combined before
//      the LDS adds either by a wavefront segmented reduction (seg_reduce=1)
//      or spread over distinct nodes by a per-lane rotation of the stencil
//      order (seg_reduce=0) so equal-cell lanes never hit one LDS address in
//      the same instruction;
//   3. the non-empty tile nodes are flushed to the global grid with FP32
//      atomics (one per node component instead of one per particle-node).
// Particles whose stencil leaves the tile (grid edges, where the reference's
// node bounds and index aliasing apply) and workgroups whose box exceeds the
// LDS budget use direct global atomics with the reference's index rules.
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "engine_types.hpp"
#include "kernels.hpp"

using namespace mpmdev;

namespace {


__device__ __forceinline__ void atomic_add_global(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add_lds(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

struct PState {
  float x[3], v[3], C[9], F[9], Jp;
};

__device__ __forceinline__ void load_state(const ParticlesDev& P, int i, PState& s) {
  const float* f = P.f;
  const size_t st = P.stride;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s.x[a] = f[(PL_X + a) * st + i];
    s.v[a] = f[(PL_V + a) * st + i];
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    s.C[a] = f[(PL_C + a) * st + i];
    s.F[a] = f[(PL_F + a) * st + i];
  }
  s.Jp = f[PL_JP * st + i];
}

__device__ __forceinline__ void store_state(const ParticlesDev& P, int i, const PState& s) {
  float* f = P.f;
  const size_t st = P.stride;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    f[(PL_X + a) * st + i] = s.x[a];
    f[(PL_V + a) * st + i] = s.v[a];
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    f[(PL_C + a) * st + i] = s.C[a];
    f[(PL_F + a) * st + i] = s.F[a];
  }
  f[PL_JP * st + i] = s.Jp;
}

__device__ __forceinline__ void gather(const PState& s, const int base[3], const float w[3][3],
                                       const StepConsts& k, const float4* __restrict__ grid,
                                       float v[3], float C[9]) {
  const int N = k.N;
  float dd[3][3];
  bool ok[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      const int g = base[a] + o;
      ok[a][o] = g >= 0 && g < N;
      dd[a][o] = ((float)(uint32_t)g * k.dx - s.x[a]) * k.dinv;
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) v[a] = 0.0f;
#pragma unroll
  for (int a = 0; a < 9; ++a) C[a] = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float wij = w[0][i] * w[1][j];
      const size_t row = ((size_t)(base[0] + i) * N + (size_t)(base[1] + j)) * N;
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        if (!(ok[0][i] && ok[1][j] && ok[2][kk])) continue;
        const float4 g = grid[row + (size_t)(base[2] + kk)];
        const float wt = wij * w[2][kk];
        const float wv0 = wt * g.x, wv1 = wt * g.y, wv2 = wt * g.z;
        v[0] += wv0;
        v[1] += wv1;
        v[2] += wv2;
        C[0] += wv0 * dd[0][i]; C[1] += wv0 * dd[1][j]; C[2] += wv0 * dd[2][kk];
        C[3] += wv1 * dd[0][i]; C[4] += wv1 * dd[1][j]; C[5] += wv1 * dd[2][kk];
        C[6] += wv2 * dd[0][i]; C[7] += wv2 * dd[1][j]; C[8] += wv2 * dd[2][kk];
      }
    }
}

// Direct scatter with the reference's node bounds and linear-index rules
// (mpm.cu:39-73): compat REFERENCE bounds the upper end by N + base and lets
// j, k >= N alias into the next row; indices >= N^3 are dropped.
__device__ __forceinline__ void scatter_global(const float q0[3][3], const float q1[3][3],
                                               const float q2[3][3], float mass,
                                               const int base[3], const float w[3][3],
                                               const StepConsts& k, float* __restrict__ grid) {
  const int N = k.N;
  int beg[3], end[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    beg[a] = max(0, -base[a]);
    end[a] = min(3, (k.compat == MPM_COMPAT_FIXED) ? N - base[a] : N + base[a]);
  }
  const uint64_t n3 = (uint64_t)N * N * N;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (i < beg[0] || i >= end[0]) continue;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (j < beg[1] || j >= end[1]) continue;
      const float wij = w[0][i] * w[1][j];
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        if (kk < beg[2] || kk >= end[2]) continue;
        const uint32_t idx = (uint32_t)N * (uint32_t)N * (uint32_t)(base[0] + i) +
                             (uint32_t)N * (uint32_t)(base[1] + j) + (uint32_t)(base[2] + kk);
        if ((uint64_t)idx >= n3) continue;
        const float wt = wij * w[2][kk];
        float* cell = grid + 4 * (size_t)idx;
        atomic_add_global(cell + 0, wt * (q0[0][i] + (q1[0][j] + q2[0][kk])));
        atomic_add_global(cell + 1, wt * (q0[1][i] + (q1[1][j] + q2[1][kk])));
        atomic_add_global(cell + 2, wt * (q0[2][i] + (q1[2][j] + q2[2][kk])));
        atomic_add_global(cell + 3, wt * mass);
      }
    }
  }
}

__device__ __forceinline__ float pick3(const float a[3], int r) {
  return r == 0 ? a[0] : (r == 1 ? a[1] : a[2]);
}
__device__ __forceinline__ int pick3i(const int a[3], int r) {
  return r == 0 ? a[0] : (r == 1 ? a[1] : a[2]);
}

}  // namespace

template <int B>
__global__ void __launch_bounds__(B)
    k_g2p2g_tiled(ParticlesDev P, const MaterialDev* __restrict__ mats, int nmats, StepConsts k,
                  const float4* __restrict__ grid_in, float* __restrict__ grid_out, int do_g2p,
                  int seg_reduce) {
  constexpr int W = B / 64;
  // LDS tile budget: 4 nodes per thread (B = 256: 4 planes x 1024 x 4 B = 16 KiB),
  // capped so tile + material table stay within 64 KiB per workgroup
  constexpr int kMaxTileNodes = (4 * B > 3584) ? 3584 : 4 * B;
  __shared__ float tile[4 * kMaxTileNodes];
  __shared__ MaterialDev smat[MPM_MAX_MATERIALS];
  __shared__ int red[6][W];
  __shared__ int tinfo[8];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int i = blockIdx.x * B + tid;
  const bool valid = i < P.n;
  {
    const int words = nmats * (int)(sizeof(MaterialDev) / 4);
    for (int t = tid; t < words; t += B) ((float*)smat)[t] = ((const float*)mats)[t];
  }
  __syncthreads();

  PState s;
  MaterialDev m;
  float R[9];
  bool has_R = false;
  int base[3] = {0, 0, 0};
  float w[3][3];
  bool inside = false;
  if (valid) {
    load_state(P, i, s);
    m = smat[P.mat[i]];
    quad_weights(s.x, k.dx_inv, base, w);
    bool moved = false;
    if (do_g2p && P.id[i] <= k.guard && !outside_domain(base, k.N)) {
      // ---- G2P of substep n
      gather(s, base, w, k, grid_in, s.v, s.C);
      float A[9];
#pragma unroll
      for (int a = 0; a < 9; ++a) A[a] = ((a % 4 == 0) ? 1.0f : 0.0f) + k.dt * s.C[a];
      mul_ab(A, s.F, s.F);
      plasticity(s.F, s.Jp, m, R, has_R);
#pragma unroll
      for (int a = 0; a < 3; ++a) s.x[a] = s.x[a] + k.dt * s.v[a];
      store_state(P, i, s);
      moved = true;
    }
    if (moved) quad_weights(s.x, k.dx_inv, base, w);
    inside = !outside_domain(base, k.N);
  }

  // ---- tile box of the workgroup's base cells
  {
    int mn[3], mx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      mn[a] = wave_min(inside ? base[a] : 0x7fffffff);
      mx[a] = wave_max(inside ? base[a] : -0x7fffffff);
    }
    if (lane == 0) {
#pragma unroll
      for (int a = 0; a < 3; ++a) { red[a][wave] = mn[a]; red[3 + a][wave] = mx[a]; }
    }
  }
  __syncthreads();
  if (tid == 0) {
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      int mn = red[a][0], mx = red[3 + a][0];
      for (int q = 1; q < W; ++q) { mn = min(mn, red[a][q]); mx = max(mx, red[3 + a][q]); }
      // in-domain bases lie in [-3, N-1], so these stay far from overflow;
      // an empty workgroup (mn = INT_MAX, mx = -INT_MAX) gives lo > hi
      lo[a] = max(mn, 0);
      hi[a] = (mx < -3) ? -1 : min(mx + 2, k.N - 1);
    }
    int dims[3];
    long long nodes = 1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      // no in-domain particle in the workgroup: lo > hi, empty tile
      dims[a] = (lo[a] <= hi[a]) ? hi[a] - lo[a] + 1 : 0;
      nodes *= dims[a];
    }
    const bool use = nodes > 0 && nodes <= kMaxTileNodes;
    tinfo[0] = lo[0]; tinfo[1] = lo[1]; tinfo[2] = lo[2];
    tinfo[3] = use ? dims[0] : 0; tinfo[4] = use ? dims[1] : 0; tinfo[5] = use ? dims[2] : 0;
    tinfo[6] = use ? (int)nodes : 0;
  }
  __syncthreads();
  const int o0 = tinfo[0], o1 = tinfo[1], o2 = tinfo[2];
  const int t0 = tinfo[3], t1 = tinfo[4], t2 = tinfo[5];
  const int tn = tinfo[6];
  for (int t = tid; t < 4 * tn; t += B) tile[t] = 0.0f;
  __syncthreads();

  // ---- P2G preparation (TransferScheme.h:66-86): stress and affine momentum.
  // Node value_c(a,b,c) = wt * (q0[c][a] + (q1[c][b] + q2[c][c'])), with
  // q0 = m v + aff(:,0) d0, q1 = aff(:,1) d1, q2 = aff(:,2) d2.
  float q0[3][3], q1[3][3], q2[3][3];
  bool tiled_lane = false;
  float pmass = 0.0f;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int o = 0; o < 3; ++o) { q0[c][o] = 0.0f; q1[c][o] = 0.0f; q2[c][o] = 0.0f; }
  if (inside) {
    float tau[9];
    if (m.kind != MPM_MATERIAL_WC_LIQUID && !has_R) polar_rotation(s.F, R);
    kirchhoff_stress(s.F, s.Jp, R, m, k.compat, tau);
    float aff[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) aff[a] = m.stress_scale * tau[a] + m.mass * s.C[a];
    float d[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int o = 0; o < 3; ++o) d[a][o] = (float)(uint32_t)(base[a] + o) * k.dx - s.x[a];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int o = 0; o < 3; ++o) {
        q0[c][o] = s.v[c] * m.mass + aff[3 * c + 0] * d[0][o];
        q1[c][o] = aff[3 * c + 1] * d[1][o];
        q2[c][o] = aff[3 * c + 2] * d[2][o];
      }
    const bool in_tile = tn > 0 && base[0] >= o0 && base[1] >= o1 && base[2] >= o2 &&
                         base[0] + 2 < o0 + t0 && base[1] + 2 < o1 + t1 && base[2] + 2 < o2 + t2;
    if (in_tile) {
      tiled_lane = true;
      pmass = m.mass;
    } else {
      scatter_global(q0, q1, q2, m.mass, base, w, k, grid_out);
    }
  }

  if (!seg_reduce) {
    if (tiled_lane) {
      // stencil order rotated per lane: lanes sharing a cell visit distinct
      // nodes in each step of the 27-node loop
      const int r = lane % 27;
      const int r0 = r % 3, r1 = (r / 3) % 3, r2 = r / 9;
      int off0[3], off1[3], off2[3];
#pragma unroll
      for (int o = 0; o < 3; ++o) {
        off0[o] = (base[0] + o - o0) * t1 * t2;
        off1[o] = (base[1] + o - o1) * t2;
        off2[o] = base[2] + o - o2;
      }
      float wr0[3], wr1[3], wr2[3];
      int ofr0[3], ofr1[3], ofr2[3];
      float qr0[3][3], qr1[3][3], qr2[3][3];
#pragma unroll
      for (int o = 0; o < 3; ++o) {
        const int s0 = (o + r0) % 3, s1 = (o + r1) % 3, s2 = (o + r2) % 3;
        wr0[o] = pick3(w[0], s0); wr1[o] = pick3(w[1], s1); wr2[o] = pick3(w[2], s2);
        ofr0[o] = pick3i(off0, s0); ofr1[o] = pick3i(off1, s1); ofr2[o] = pick3i(off2, s2);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          qr0[c][o] = pick3(q0[c], s0);
          qr1[c][o] = pick3(q1[c], s1);
          qr2[c][o] = pick3(q2[c], s2);
        }
      }
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const float wab = wr0[a] * wr1[b];
          const int ob = ofr0[a] + ofr1[b];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const float wt = wab * wr2[c];
            const int n = ob + ofr2[c];
            atomic_add_lds(&tile[n], wt * (qr0[0][a] + (qr1[0][b] + qr2[0][c])));
            atomic_add_lds(&tile[tn + n], wt * (qr0[1][a] + (qr1[1][b] + qr2[1][c])));
            atomic_add_lds(&tile[2 * tn + n], wt * (qr0[2][a] + (qr1[2][b] + qr2[2][c])));
            atomic_add_lds(&tile[3 * tn + n], wt * pmass);
          }
        }
    }
  } else if (tn > 0) {
    // ---- wavefront segmented reduction: lanes of a run of equal base cells
    // (adjacent after the cell sort) sum their 27 x 4 node values into the
    // run's first lane, which alone issues the LDS adds. The number of
    // halving steps is set by the longest run in the wavefront.
    const int key = tiled_lane ? ((base[0] - o0) * t1 + (base[1] - o1)) * t2 + (base[2] - o2)
                               : -1 - lane;
    const int prev = __shfl_up(key, 1);
    const bool head = lane == 0 || prev != key;
    const uint64_t heads = __ballot(head);
    const uint64_t above = (lane == 63) ? 0ull : (heads & (~0ull << (lane + 1)));
    const int run_end = above ? (__ffsll((unsigned long long)above) - 1) : 64;
    const int len = wave_max(head ? run_end - lane : 0);
    int steps = 0;
    while ((1 << steps) < len) ++steps;
    float ww[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int o = 0; o < 3; ++o) ww[a][o] = tiled_lane ? w[a][o] : 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const float wab = ww[0][a] * ww[1][b];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float wt = wab * ww[2][c];
          float val[4];
          val[0] = wt * (q0[0][a] + (q1[0][b] + q2[0][c]));
          val[1] = wt * (q0[1][a] + (q1[1][b] + q2[1][c]));
          val[2] = wt * (q0[2][a] + (q1[2][b] + q2[2][c]));
          val[3] = wt * pmass;
          for (int st = 0; st < steps; ++st) {
            const int off = 1 << st;
            const bool take = lane + off < run_end;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const float o = __shfl_down(val[q], off);
              val[q] += take ? o : 0.0f;
            }
          }
          if (head && tiled_lane) {
            const int n = key + (a * t1 + b) * t2 + c;
            atomic_add_lds(&tile[n], val[0]);
            atomic_add_lds(&tile[tn + n], val[1]);
            atomic_add_lds(&tile[2 * tn + n], val[2]);
            atomic_add_lds(&tile[3 * tn + n], val[3]);
          }
        }
      }
  }
  __syncthreads();

  // ---- flush non-empty tile nodes to the global grid
  for (int t = tid; t < tn; t += B) {
    const float mass = tile[3 * tn + t];
    if (mass == 0.0f) continue;
    const int a = t / (t1 * t2), rem = t - a * t1 * t2;
    const int b = rem / t2, c = rem - b * t2;
    const size_t idx = ((size_t)(o0 + a) * k.N + (size_t)(o1 + b)) * k.N + (size_t)(o2 + c);
    float* cell = grid_out + 4 * idx;
    atomic_add_global(cell + 0, tile[t]);
    atomic_add_global(cell + 1, tile[tn + t]);
    atomic_add_global(cell + 2, tile[2 * tn + t]);
    atomic_add_global(cell + 3, mass);
  }
}

namespace mpmk {

hipError_t g2p2g_tiled(const ParticlesDev& P, const MaterialDev* mats, int nmats,
                       const StepConsts& k, const float4* grid_in, float4* grid_out, int do_g2p,
                       int chunk, int seg_reduce, hipStream_t s) {
  if (P.n <= 0) return hipSuccess;
  if (nmats < 1 || nmats > MPM_MAX_MATERIALS) return hipErrorInvalidValue;
  if (do_g2p && !grid_in) return hipErrorInvalidValue;
  switch (chunk) {
    case 256:
      hipLaunchKernelGGL(k_g2p2g_tiled<256>, dim3((P.n + 255) / 256), dim3(256), 0, s, P, mats,
                         nmats, k, grid_in, (float*)grid_out, do_g2p, seg_reduce);
      break;
    case 512:
      hipLaunchKernelGGL(k_g2p2g_tiled<512>, dim3((P.n + 511) / 512), dim3(512), 0, s, P, mats,
                         nmats, k, grid_in, (float*)grid_out, do_g2p, seg_reduce);
      break;
    case 1024:
      hipLaunchKernelGGL(k_g2p2g_tiled<1024>, dim3((P.n + 1023) / 1024), dim3(1024), 0, s, P,
                         mats, nmats, k, grid_in, (float*)grid_out, do_g2p, seg_reduce);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mpmk
