// p2g_block.hip -- fused G2P + tiled P2G over work items, and the tile
// reduction that turns the P2G tiles into the updated grid (gfx950).
//
// The particles are cell-sorted (sort.hip) and cut into work items: at most
// 256 particles of one home block of 4^3 cells, arranged over the 4
// wavefronts of a workgroup so that particles sharing a cell sit in different
// wavefronts (k_item_arrange). One workgroup runs one item:
//   1. stage the grid velocities of the item's node region in LDS: the block,
//      its 2-node stencil reach and a 1-cell drift margin on each side
//      (8^3 nodes);
//   2. G2P of substep n (gridToParticle, reference src/mpm.cu:109-178), one
//      particle per thread, gathering the 27 nodes from LDS -- C/F update,
//      plasticity with the SVD in registers, advection -- and the store of
//      the new state;
//   3. P2G of substep n+1 (particleToGrid, mpm.cu:14-74) into a private LDS
//      tile of each wavefront with plain read-modify-writes (gfx950 retires
//      LDS FP32 atomics at a small fraction of the rate of plain LDS
//      accesses). Within one instruction the lanes of a wavefront write 27
//      distinct nodes when their cells are distinct; lanes that share a cell
//      take turns (a per-wavefront claim byte per cell decides who goes);
//   4. store the sum of the 4 wavefront tiles as the item's tile in HBM
//      (plain coalesced stores, no atomics).
// k_tile_reduce then builds every grid node from the tiles that cover it
// (the items of the 27 home blocks around its node block), in a fixed order,
// and applies the grid update (gridOpKernel, mpm.cu:76-107) in the same pass.
// The sum order is a function of the sorted particle order only, so a run is
// bit-reproducible. Particles whose stencil leaves their tile (drift beyond
// the margin between sorts) or the grid (where the reference's node bounds
// and index aliasing apply) take the direct path: global gathers, and global
// atomics into a side grid that k_tile_reduce adds and clears.
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
// With `dirty`, every touched node also flags its 4^3 node block.
__device__ __forceinline__ void scatter_global(const float q0[3][3], const float q1[3][3],
                                               const float q2[3][3], float mass,
                                               const int base[3], const float w[3][3],
                                               const StepConsts& k, float* __restrict__ grid,
                                               uint32_t* __restrict__ dirty = nullptr) {
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
        if (dirty) {
          const uint32_t nbx = ((uint32_t)N + 3u) >> 2;
          const uint32_t gx = idx / ((uint32_t)N * (uint32_t)N), gy = (idx / (uint32_t)N) % (uint32_t)N,
                         gz = idx % (uint32_t)N;
          dirty[((gx >> 2) * nbx + (gy >> 2)) * nbx + (gz >> 2)] = 1u;
        }
        atomic_add_global(cell + 0, wt * (q0[0][i] + (q1[0][j] + q2[0][kk])));
        atomic_add_global(cell + 1, wt * (q0[1][i] + (q1[1][j] + q2[1][kk])));
        atomic_add_global(cell + 2, wt * (q0[2][i] + (q1[2][j] + q2[2][kk])));
        atomic_add_global(cell + 3, wt * mass);
      }
    }
  }
}

}  // namespace

// Material-kind sets compiled into an instantiation.
enum KindSet : int { KINDS_ALL = 0, KINDS_LIQUID = 1 };

template <int KS>
__global__ void __launch_bounds__(256)
    k_g2p2g_tiled(ParticlesDev P, const MaterialDev* __restrict__ mats, StepConsts k,
                  const uint32_t* __restrict__ item_start, const uint32_t* __restrict__ item_block,
                  const uint32_t* __restrict__ item_waves, const uint32_t* __restrict__ item_count,
                  const float4* __restrict__ grid_v, float4* __restrict__ tiles,
                  float* __restrict__ grid_atomic, uint32_t* __restrict__ dirty, int do_g2p) {
  constexpr int BE = 1 << mpmk::kBlockLog2;
  constexpr int W = mpmk::kTileW;        // BE + 2 (stencil reach) + 2 (drift margin)
  constexpr int TN = mpmk::kTileNodes;
  constexpr int NCE = W - 2;             // cells per axis whose stencil fits the tile
  static_assert(NCE * NCE * NCE <= 256, "claim bytes per wavefront");
  __shared__ float4 vt[TN];              // grid velocities; claim bytes once G2P is done
  __shared__ float4 acc[4][TN];          // P2G tile of each wavefront

  const int item = blockIdx.x;
  if ((uint32_t)item >= *item_count) return;  // uniform for the workgroup
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t start = item_start[item];
  const uint32_t wc = item_waves[item];
  const uint32_t blk = item_block[item];
  int off_w = 0;
  for (int q = 0; q < wave; ++q) off_w += (int)((wc >> (8 * q)) & 255u);
  const int cnt_w = (int)((wc >> (8 * wave)) & 255u);
  const int N = k.N;
  const int o0 = (int)compact3(blk >> 2) * BE - 1;
  const int o1 = (int)compact3(blk >> 1) * BE - 1;
  const int o2 = (int)compact3(blk) * BE - 1;

  if (do_g2p) {
    for (int n = tid; n < TN; n += 256) {
      const int a = n / (W * W), r = n - a * W * W, b = r / W, c = r - b * W;
      const int gi = o0 + a, gj = o1 + b, gk = o2 + c;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gi >= 0 && gi < N && gj >= 0 && gj < N && gk >= 0 && gk < N)
        v = grid_v[((size_t)gi * N + (size_t)gj) * N + (size_t)gk];
      vt[n] = v;
    }
  }
  {
    float4* const flat = &acc[0][0];
    for (int n = tid; n < 4 * TN; n += 256) flat[n] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  __syncthreads();

  const bool valid = lane < cnt_w;
  const uint32_t i = start + (uint32_t)(off_w + lane);
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
        gather(s, base, w, k, grid_v, s.v, s.C);
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
  // Node value (a,b,c) = wt * (q0[.][a] + (q1[.][b] + q2[.][c])), with
  // q0 = m v + aff(:,0) d0, q1 = aff(:,1) d1, q2 = aff(:,2) d2.
  int cell = 0, nb = 0;
  bool pending = false;
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
      nb = (l0 * W + l1) * W + l2;
      pending = true;
    } else {
      scatter_global(q0, q1, q2, m.mass, base, w, k, grid_atomic, dirty);
    }
  }
  __syncthreads();  // every G2P read of vt is done: its bytes become claim bytes

  // ---- P2G into the wavefront's tile. Per round, every pending lane claims
  // its cell; the lane whose claim stands adds its 27 nodes. The claimed
  // cells are distinct, so the 64 lanes of one read-modify-write address
  // distinct nodes; the compiler barrier keeps the 27 read-modify-writes in
  // program order (a lane's next read may address a node another lane is
  // writing in the same round).
  typedef __attribute__((address_space(3))) volatile uint8_t lds_byte;
  lds_byte* const claim = (lds_byte*)(reinterpret_cast<uint8_t*>(vt) + wave * 256);
  float4* const my = acc[wave];
  while (__ballot(pending) != 0ull) {
    if (pending) claim[cell] = (uint8_t)lane;
    if (pending && claim[cell] == (uint8_t)lane) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const float wab = w[0][a] * w[1][b];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const float wt = wab * w[2][c];
            float4* const p = my + nb + (a * W + b) * W + c;
            float4 t = *p;
            t.x += wt * (q0[0][a] + (q1[0][b] + q2[0][c]));
            t.y += wt * (q0[1][a] + (q1[1][b] + q2[1][c]));
            t.z += wt * (q0[2][a] + (q1[2][b] + q2[2][c]));
            t.w += wt * m.mass;
            *p = t;
            __asm__ __volatile__("" ::: "memory");
          }
        }
      pending = false;
    }
  }
  __syncthreads();

  // ---- the item's tile: sum of the wavefront tiles, fixed order
  float4* const out = tiles + (size_t)item * TN;
  for (int n = tid; n < TN; n += 256) {
    const float4 a = acc[0][n], b = acc[1][n], c = acc[2][n], e = acc[3][n];
    out[n] = make_float4(((a.x + b.x) + c.x) + e.x, ((a.y + b.y) + c.y) + e.y,
                         ((a.z + b.z) + c.z) + e.z, ((a.w + b.w) + c.w) + e.w);
  }
}

// One wavefront per 4^3 node block: each lane sums its node over the tiles of
// the items of the home blocks around it, adds the side grid where particles
// left their tiles (and clears it), and writes the node: grid update
// (update=1) or raw sums. Home block h covers nodes 4h-1 .. 4h+6 per axis, so
// a node with in-block coordinate u is covered by home blocks h-1 and h
// (u < 3) or h and h+1 (u = 3) on each axis: 8 home blocks per node. Lane
// j < 27 fetches the item range of neighbour j of the node block; each lane
// picks the ranges of its 8 blocks with lane shuffles and reads the first
// item of each in one batch of independent loads, further items (blocks
// holding more than one item) after it. The order of the sum is fixed.
__global__ void __launch_bounds__(256)
    k_tile_reduce(const float4* __restrict__ tiles, const uint32_t* __restrict__ block_items,
                  float4* __restrict__ grid_atomic, uint32_t* __restrict__ dirty,
                  float4* __restrict__ grid, StepConsts k, int nbx, int update) {
  constexpr int W = mpmk::kTileW, TN = mpmk::kTileNodes;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t nbk = (int64_t)blockIdx.x * 4 + wave;
  if (nbk >= (int64_t)nbx * nbx * nbx) return;  // uniform for the wavefront
  const int bz = (int)(nbk % nbx), by = (int)((nbk / nbx) % nbx), bx = (int)(nbk / ((int64_t)nbx * nbx));
  const int u0 = lane >> 4, u1 = (lane >> 2) & 3, u2 = lane & 3;
  uint32_t r0 = 0, r1 = 0;
  if (lane < 27) {
    const int hx = bx + lane / 9 - 1, hy = by + (lane / 3) % 3 - 1, hz = bz + lane % 3 - 1;
    if (hx >= 0 && hx < nbx && hy >= 0 && hy < nbx && hz >= 0 && hz < nbx) {
      const uint32_t mb = morton3((uint32_t)hx, (uint32_t)hy, (uint32_t)hz);
      r0 = block_items[2 * (size_t)mb];
      r1 = block_items[2 * (size_t)mb + 1];
    }
  }
  const int N = k.N;
  const int x = bx * 4 + u0, y = by * 4 + u1, z = bz * 4 + u2;
  const bool in_grid = x < N && y < N && z < N;
  const size_t node = ((size_t)x * N + (size_t)y) * N + (size_t)z;
  const bool is_dirty = dirty[nbk] != 0u;
  float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
  if (__ballot(r0 < r1) != 0ull) {  // some home block around holds items
    const int e0 = (u0 == 3) ? 0 : -1, e1 = (u1 == 3) ? 0 : -1, e2 = (u2 == 3) ? 0 : -1;
    float4 v[8];
    uint32_t a0[8], a1[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int f0 = e0 + (q >> 2), f1 = e1 + ((q >> 1) & 1), f2 = e2 + (q & 1);
      const int j = 9 * (f0 + 1) + 3 * (f1 + 1) + (f2 + 1);
      a0[q] = (uint32_t)__shfl((int)r0, j);
      a1[q] = (uint32_t)__shfl((int)r1, j);
      const int t = ((u0 + 1 - 4 * f0) * W + (u1 + 1 - 4 * f1)) * W + (u2 + 1 - 4 * f2);
      a1[q] = a0[q] < a1[q] ? a1[q] : a0[q];
      v[q] = a0[q] < a1[q] ? tiles[(size_t)a0[q] * TN + t] : make_float4(0.f, 0.f, 0.f, 0.f);
      a0[q] = (uint32_t)t;  // tile node from here on
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      sum.x += v[q].x; sum.y += v[q].y; sum.z += v[q].z; sum.w += v[q].w;
    }
    // further items of blocks holding more than one
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int f0 = e0 + (q >> 2), f1 = e1 + ((q >> 1) & 1), f2 = e2 + (q & 1);
      const int j = 9 * (f0 + 1) + 3 * (f1 + 1) + (f2 + 1);
      const uint32_t first = (uint32_t)__shfl((int)r0, j);
      for (uint32_t it = first + 1; it < a1[q]; ++it) {
        const float4 w4 = tiles[(size_t)it * TN + a0[q]];
        sum.x += w4.x; sum.y += w4.y; sum.z += w4.z; sum.w += w4.w;
      }
    }
  }
  if (is_dirty) {
    if (in_grid) {
      const float4 g = grid_atomic[node];
      sum.x += g.x; sum.y += g.y; sum.z += g.z; sum.w += g.w;
      grid_atomic[node] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (lane == 0) dirty[nbk] = 0u;
  }
  if (in_grid) grid[node] = update ? grid_node_update(sum, x, y, z, k) : sum;
}

namespace mpmk {

template <int KS>
static void launch_tiled(const ParticlesDev& P, const MaterialDev* mats, const StepConsts& k,
                         const WorkItems& w, const float4* gv, const TileSums& o, int do_g2p,
                         hipStream_t s) {
  hipLaunchKernelGGL((k_g2p2g_tiled<KS>), dim3((unsigned)w.max_items), dim3(256), 0, s, P, mats, k,
                     w.start, w.block, w.waves, w.count, gv, o.tiles, (float*)o.atomic, o.dirty,
                     do_g2p);
}

hipError_t g2p2g_block(const ParticlesDev& P, const MaterialDev* mats, int nmats,
                       const StepConsts& k, const WorkItems& w, const float4* grid_v,
                       const TileSums& out, int do_g2p, int kinds, hipStream_t s) {
  if (P.n <= 0 || w.max_items <= 0) return hipSuccess;
  if (nmats < 1 || nmats > MPM_MAX_MATERIALS) return hipErrorInvalidValue;
  if (do_g2p && !grid_v) return hipErrorInvalidValue;
  if (!out.tiles || !out.atomic || !out.dirty) return hipErrorInvalidValue;
  if (w.max_items > 0x7fffffff) return hipErrorInvalidValue;
  if (kinds == KINDS_LIQUID) launch_tiled<KINDS_LIQUID>(P, mats, k, w, grid_v, out, do_g2p, s);
  else launch_tiled<KINDS_ALL>(P, mats, k, w, grid_v, out, do_g2p, s);
  return hipGetLastError();
}

hipError_t tile_reduce(const TileSums& in, const StepConsts& k, float4* grid, int update,
                       hipStream_t s) {
  if (!in.tiles || !in.atomic || !in.dirty || !in.block_items || !grid) return hipErrorInvalidValue;
  const int nbx = (k.N + 3) / 4;
  const int64_t nblocks = (int64_t)nbx * nbx * nbx;
  hipLaunchKernelGGL(k_tile_reduce, dim3((unsigned)((nblocks + 3) / 4)), dim3(256), 0, s, in.tiles,
                     in.block_items, in.atomic, in.dirty, grid, k, nbx, update);
  return hipGetLastError();
}

}  // namespace mpmk
