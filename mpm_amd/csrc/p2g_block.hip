// p2g_block.hip -- fused G2P + block-tiled P2G (gfx950).
//
// The particles are cell-sorted (sort.hip) and cut into work items: runs of
// at most 256 particles that share one home block of BE^3 grid cells. One
// workgroup of 256 threads runs one item, one particle per thread:
//   1. stage the grid velocities of the block's node region in LDS (the block,
//      its 2-node stencil reach and a 1-cell drift margin on each side:
//      W = BE + 4 nodes per axis);
//   2. G2P of substep n (gridToParticle, reference src/mpm.cu:109-178)
//      gathering the 27 nodes from LDS -- C/F update, plasticity with the SVD
//      in registers, advection -- and the store of the new state;
//   3. P2G of substep n+1 (particleToGrid, mpm.cu:14-74) as a gather: each
//      particle writes its per-axis weights and momentum terms to LDS, slotted
//      by a counting sort on its new cell (LDS integer counters); then each
//      thread owns tile nodes and sums, in registers, the contributions of the
//      particles in the 27 cells around each node. No floating-point atomics
//      touch LDS (gfx950 retires LDS FP32 atomics at a small fraction of the
//      rate of plain LDS reads), and the order of the sum is fixed;
//   4. add the non-empty tile nodes to the global grid with FP32 atomics,
//      consecutive lanes on consecutive nodes of a row.
// Particles whose stencil leaves the tile (drift beyond the margin between
// sorts) or the grid (where the reference's node bounds and index aliasing
// apply) take the direct path: global gathers and global atomics with the
// reference's index rules.
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "engine_types.hpp"
#include "kernels.hpp"

using namespace mpmdev;

namespace {


__device__ __forceinline__ void atomic_add_global(float* p, float v) { unsafeAtomicAdd(p, v); }
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



__device__ __forceinline__ uint32_t compact3(uint32_t v) {
  // inverse of the 3-way bit interleave: keep bits 0, 3, 6, ...
  v &= 0x09249249u;
  v = (v | (v >> 2)) & 0x030c30c3u;
  v = (v | (v >> 4)) & 0x0300f00fu;
  v = (v | (v >> 8)) & 0x030000ffu;
  v = (v | (v >> 16)) & 0x000003ffu;
  return v;
}

}  // namespace

// Material-kind sets compiled into an instantiation.
enum KindSet : int { KINDS_ALL = 0, KINDS_LIQUID = 1 };

template <int LB, int KS>
__global__ void __launch_bounds__(256)
    k_g2p2g_block(ParticlesDev P, const MaterialDev* __restrict__ mats, StepConsts k,
                  const uint32_t* __restrict__ item_start, const uint32_t* __restrict__ item_block,
                  const uint32_t* __restrict__ item_count, const float4* __restrict__ grid_in,
                  float* __restrict__ grid_out, int do_g2p) {
  constexpr int BE = 1 << LB;
  constexpr int MARGIN = 1;
  constexpr int W = BE + 2 + 2 * MARGIN;  // tile width in nodes
  constexpr int TN = W * W * W;           // tile nodes
  constexpr int NCE = W - 2;              // cells whose 3x3x3 stencil fits the tile, per axis
  constexpr int NC = NCE * NCE * NCE;
  // The velocity tile (read by G2P) and the particle packets (written after
  // the first barrier that follows every G2P read) share one LDS buffer.
  constexpr int PK_F4 = 9 * 256, U_F4 = PK_F4 > TN ? PK_F4 : TN;
  __shared__ float4 ubuf[U_F4];
  float4* const vt = ubuf;                // staged grid velocities (G2P)
  float4 (*const pk)[256] = reinterpret_cast<float4 (*)[256]>(ubuf);  // [axis*3+offset][slot]: (q, w)
  __shared__ float pm[256];               // particle mass, by slot
  __shared__ int cstart[NC + 1];          // per-cell counts, then first slot
  __shared__ int red[8];

  const int item = blockIdx.x;
  if ((uint32_t)item >= *item_count) return;  // uniform for the workgroup
  const int tid = threadIdx.x;
  const uint32_t start = item_start[item], end = item_start[item + 1];  // end - start <= 256
  const uint32_t blk = item_block[item];
  const int N = k.N;
  const int o0 = (int)compact3(blk >> 2) * BE - MARGIN;
  const int o1 = (int)compact3(blk >> 1) * BE - MARGIN;
  const int o2 = (int)compact3(blk) * BE - MARGIN;

  if (do_g2p) {
    for (int n = tid; n < TN; n += 256) {
      const int a = n / (W * W), r = n - a * W * W, b = r / W, c = r - b * W;
      const int gi = o0 + a, gj = o1 + b, gk = o2 + c;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gi >= 0 && gi < N && gj >= 0 && gj < N && gk >= 0 && gk < N)
        v = grid_in[((size_t)gi * N + (size_t)gj) * N + (size_t)gk];
      vt[n] = v;
    }
  }
  for (int t = tid; t <= NC; t += 256) cstart[t] = 0;
  __syncthreads();

  const uint32_t i = start + (uint32_t)tid;
  const bool valid = i < end;
  PState s;
  MaterialDev m;
  float R[9];
  bool has_R = false;
  int base[3] = {0, 0, 0};
  float w[3][3];
  bool inside = false;
  if (valid) {
    load_state(P, (int)i, s);
    m = mats[P.mat[i]];
    quad_weights(s.x, k.dx_inv, base, w);
    if (do_g2p && P.id[i] <= k.guard && !outside_domain(base, N)) {
      // ---- G2P of substep n, gathering the 27 nodes from the LDS tile
      const int l0 = base[0] - o0, l1 = base[1] - o1, l2 = base[2] - o2;
      if (l0 >= 0 && l1 >= 0 && l2 >= 0 && l0 < NCE && l1 < NCE && l2 < NCE) {
        float dd[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int o = 0; o < 3; ++o) dd[a][o] = ((float)(uint32_t)(base[a] + o) * k.dx - s.x[a]) * k.dinv;
#pragma unroll
        for (int a = 0; a < 3; ++a) s.v[a] = 0.0f;
#pragma unroll
        for (int a = 0; a < 9; ++a) s.C[a] = 0.0f;
        const int nb = (l0 * W + l1) * W + l2;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            const float wab = w[0][a] * w[1][b];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const float4 g = vt[nb + (a * W + b) * W + c];
              const float wt = wab * w[2][c];
              const float wv0 = wt * g.x, wv1 = wt * g.y, wv2 = wt * g.z;
              s.v[0] += wv0; s.v[1] += wv1; s.v[2] += wv2;
              s.C[0] += wv0 * dd[0][a]; s.C[1] += wv0 * dd[1][b]; s.C[2] += wv0 * dd[2][c];
              s.C[3] += wv1 * dd[0][a]; s.C[4] += wv1 * dd[1][b]; s.C[5] += wv1 * dd[2][c];
              s.C[6] += wv2 * dd[0][a]; s.C[7] += wv2 * dd[1][b]; s.C[8] += wv2 * dd[2][c];
            }
          }
      } else {
        gather(s, base, w, k, grid_in, s.v, s.C);
      }
      float A[9];
#pragma unroll
      for (int a = 0; a < 9; ++a) A[a] = ((a % 4 == 0) ? 1.0f : 0.0f) + k.dt * s.C[a];
      mul_ab(A, s.F, s.F);
      if (KS == KINDS_LIQUID) {
        // F <- J^(1/3) I, Jp <- J (the only kind this instantiation handles)
        const float J = det3(s.F);
        const float cj = cbrtf(J);
#pragma unroll
        for (int a = 0; a < 9; ++a) s.F[a] = (a % 4 == 0) ? cj : 0.0f;
        s.Jp = J;
      } else {
        plasticity(s.F, s.Jp, m, R, has_R);
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) s.x[a] = s.x[a] + k.dt * s.v[a];
      store_state(P, (int)i, s);
      quad_weights(s.x, k.dx_inv, base, w);
    }
    inside = !outside_domain(base, N);
  }

  // ---- P2G preparation (TransferScheme.h:66-86): stress and affine momentum.
  // Node value_c(a,b,c) = wt * (q0[c][a] + (q1[c][b] + q2[c][c'])), with
  // q0 = m v + aff(:,0) d0, q1 = aff(:,1) d1, q2 = aff(:,2) d2.
  int cell = -1, rank = 0;
  float q0[3][3], q1[3][3], q2[3][3];
  if (inside) {
    float tau[9];
    if (KS == KINDS_LIQUID) {
      const float pr = m.bulk * ((s.Jp - 1.0f) * s.Jp);
#pragma unroll
      for (int a = 0; a < 9; ++a) tau[a] = (a % 4 == 0) ? pr : 0.0f;
    } else {
      if (m.kind != MPM_MATERIAL_WC_LIQUID && !has_R) polar_rotation(s.F, R);
      kirchhoff_stress(s.F, s.Jp, R, m, k.compat, tau);
    }
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
    const int l0 = base[0] - o0, l1 = base[1] - o1, l2 = base[2] - o2;
    // tile path: the stencil lies in the tile and inside the grid (at the grid
    // faces the reference's node bounds and aliasing rules apply: direct path)
    const bool in_tile = l0 >= 0 && l1 >= 0 && l2 >= 0 && l0 < NCE && l1 < NCE && l2 < NCE &&
                         base[0] >= 0 && base[1] >= 0 && base[2] >= 0 && base[0] + 2 < N &&
                         base[1] + 2 < N && base[2] + 2 < N;
    if (in_tile) {
      cell = (l0 * NCE + l1) * NCE + l2;
      rank = atomicAdd(&cstart[cell], 1);  // LDS integer counter: slot within the cell
    } else {
      scatter_global(q0, q1, q2, m.mass, base, w, k, grid_out);
    }
  }
  __syncthreads();

  // ---- cell ranges: exclusive scan of the NC counts (each thread owns a run)
  {
    constexpr int PER = (NC + 255) / 256;
    int loc[PER];
    int sum = 0;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int c = tid * PER + t;
      loc[t] = (c < NC) ? cstart[c] : 0;
      sum += loc[t];
    }
    // workgroup exclusive scan of `sum` (wave scans, then wave totals)
    int incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(incl, off);
      if ((tid & 63) >= off) incl += y;
    }
    if ((tid & 63) == 63) red[tid >> 6] = incl;
    __syncthreads();
    int wave_base = 0;
    for (int q = 0; q < (tid >> 6); ++q) wave_base += red[q];
    int run = wave_base + incl - sum;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int c = tid * PER + t;
      if (c < NC) cstart[c] = run;
      run += loc[t];
    }
    if (tid == 255) cstart[NC] = run;
  }
  __syncthreads();
  if (cell >= 0) {
    const int slot = cstart[cell] + rank;
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      pk[0 * 3 + o][slot] = make_float4(q0[0][o], q0[1][o], q0[2][o], w[0][o]);
      pk[1 * 3 + o][slot] = make_float4(q1[0][o], q1[1][o], q1[2][o], w[1][o]);
      pk[2 * 3 + o][slot] = make_float4(q2[0][o], q2[1][o], q2[2][o], w[2][o]);
    }
    pm[slot] = m.mass;
  }
  __syncthreads();

  // ---- P2G as a gather over the tile: each thread sums, in registers, the
  // contributions of the particles in the 27 cells around each of its nodes,
  // then adds the node to the global grid (coalesced FP32 atomics)
  for (int n = tid; n < TN; n += 256) {
    const int a = n / (W * W), r = n - a * W * W, b = r / W, c = r - b * W;
    float ax = 0.f, ay = 0.f, az = 0.f, am = 0.f;
    bool touched = false;
#pragma unroll
    for (int oa = 0; oa < 3; ++oa) {
      const int ca = a - oa;
      if (ca < 0 || ca >= NCE) continue;
#pragma unroll
      for (int ob = 0; ob < 3; ++ob) {
        const int cb = b - ob;
        if (cb < 0 || cb >= NCE) continue;
#pragma unroll
        for (int oc = 0; oc < 3; ++oc) {
          const int cc = c - oc;
          if (cc < 0 || cc >= NCE) continue;
          const int cl = (ca * NCE + cb) * NCE + cc;
          const int j0 = cstart[cl], j1 = cstart[cl + 1];
          touched |= j1 > j0;
          for (int j = j0; j < j1; ++j) {
            const float4 A = pk[oa][j];
            const float4 B = pk[3 + ob][j];
            const float4 Cq = pk[6 + oc][j];
            const float wt = (A.w * B.w) * Cq.w;
            ax += wt * (A.x + (B.x + Cq.x));
            ay += wt * (A.y + (B.y + Cq.y));
            az += wt * (A.z + (B.z + Cq.z));
            am += wt * pm[j];
          }
        }
      }
    }
    if (touched) {
      const size_t idx = ((size_t)(o0 + a) * N + (size_t)(o1 + b)) * N + (size_t)(o2 + c);
      float* cellp = grid_out + 4 * idx;
      atomic_add_global(cellp + 0, ax);
      atomic_add_global(cellp + 1, ay);
      atomic_add_global(cellp + 2, az);
      atomic_add_global(cellp + 3, am);
    }
  }
}

namespace mpmk {

template <int LB, int KS>
static void launch_block(const ParticlesDev& P, const MaterialDev* mats, int nmats,
                         const StepConsts& k, const WorkItems& w, const float4* gin, float4* gout,
                         int do_g2p, hipStream_t s) {
  hipLaunchKernelGGL((k_g2p2g_block<LB, KS>), dim3((unsigned)w.max_items), dim3(256), 0, s, P, mats,
                     k, w.start, w.block, w.count, gin, (float*)gout, do_g2p);
}

hipError_t g2p2g_block(const ParticlesDev& P, const MaterialDev* mats, int nmats,
                       const StepConsts& k, const WorkItems& w, const float4* grid_in,
                       float4* grid_out, int do_g2p, int kinds, hipStream_t s) {
  if (P.n <= 0 || w.max_items <= 0) return hipSuccess;
  if (nmats < 1 || nmats > MPM_MAX_MATERIALS) return hipErrorInvalidValue;
  if (do_g2p && !grid_in) return hipErrorInvalidValue;
  if (w.max_items > 0x7fffffff) return hipErrorInvalidValue;
  const bool liquid = kinds == KINDS_LIQUID;
  if (w.lb == 2) {
    if (liquid) launch_block<2, KINDS_LIQUID>(P, mats, nmats, k, w, grid_in, grid_out, do_g2p, s);
    else launch_block<2, KINDS_ALL>(P, mats, nmats, k, w, grid_in, grid_out, do_g2p, s);
  } else if (w.lb == 3) {
    if (liquid) launch_block<3, KINDS_LIQUID>(P, mats, nmats, k, w, grid_in, grid_out, do_g2p, s);
    else launch_block<3, KINDS_ALL>(P, mats, nmats, k, w, grid_in, grid_out, do_g2p, s);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mpmk
