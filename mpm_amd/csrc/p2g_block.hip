// p2g_block.hip -- fused G2P + tiled P2G over work items, and the tile
// reduction that turns the P2G tiles into the updated grid (gfx950).
//
// The particles are cell-sorted (sort.hip) and cut into work items: at most
// 64 kItemWaves particles of one home block of 4^3 cells, arranged over the
// kItemWaves wavefronts of a workgroup so that particles sharing a cell sit in
// different wavefronts (k_item_arrange). One workgroup runs one item:
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
//   4. store the sum of the wavefront tiles as the item's tile in HBM
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

#include <algorithm>

#include "device_math.hpp"
#include "engine_types.hpp"
#include "kernels.hpp"

using namespace mpmdev;

namespace {


__device__ __forceinline__ void atomic_add_global(float* p, float v) { unsafeAtomicAdd(p, v); }
struct PState {
  float x[3], v[3], C[9], F[9], Jp;
};

// The particle planes through one buffer resource: the lane's byte offset in
// a VGPR, the plane's offset in an SGPR, so the 25 loads and stores of a
// particle need no per-plane 64-bit address arithmetic. The planes must span
// less than 4 GiB (the engine checks its capacity against that).
struct PlaneIO {
  __amdgpu_buffer_rsrc_t rsrc;
  uint32_t plane_bytes;
  __device__ __forceinline__ explicit PlaneIO(const ParticlesDev& P)
      : rsrc(__builtin_amdgcn_make_buffer_rsrc(
            (void*)P.f, (short)0, (int)(uint32_t)((size_t)NUM_PLANES * P.stride * sizeof(float)),
            0x00020000 /* gfx9 buffer format: 32-bit data */)),
        plane_bytes((uint32_t)(P.stride * sizeof(float))) {}
  __device__ __forceinline__ float ld(int plane, uint32_t voff) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)voff,
                                                                          (int)(plane * plane_bytes), 0));
  }
  __device__ __forceinline__ void st(int plane, uint32_t voff, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), rsrc, (int)voff,
                                          (int)(plane * plane_bytes), 0);
  }
};

// x, F, Jp: what G2P reads (v and C are recomputed by it)
__device__ __forceinline__ void load_xfj(const PlaneIO& io, uint32_t i, PState& s) {
  const uint32_t o = i * 4u;
#pragma unroll
  for (int a = 0; a < 3; ++a) s.x[a] = io.ld(PL_X + a, o);
#pragma unroll
  for (int a = 0; a < 9; ++a) s.F[a] = io.ld(PL_F + a, o);
  s.Jp = io.ld(PL_JP, o);
}
// x, Jp: what G2P reads when F is derived from J (liquid, see k_g2p2g_tiled)
__device__ __forceinline__ void load_xj(const PlaneIO& io, uint32_t i, PState& s) {
  const uint32_t o = i * 4u;
#pragma unroll
  for (int a = 0; a < 3; ++a) s.x[a] = io.ld(PL_X + a, o);
  s.Jp = io.ld(PL_JP, o);
}
// v, C: needed only for particles G2P leaves as they are
__device__ __forceinline__ void load_vc(const PlaneIO& io, uint32_t i, PState& s) {
  const uint32_t o = i * 4u;
#pragma unroll
  for (int a = 0; a < 3; ++a) s.v[a] = io.ld(PL_V + a, o);
#pragma unroll
  for (int a = 0; a < 9; ++a) s.C[a] = io.ld(PL_C + a, o);
}

// with_f = false leaves the F planes unwritten (F derived from J)
__device__ __forceinline__ void store_state(const PlaneIO& io, uint32_t i, const PState& s,
                                            bool with_f = true) {
  const uint32_t o = i * 4u;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    io.st(PL_X + a, o, s.x[a]);
    io.st(PL_V + a, o, s.v[a]);
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) io.st(PL_C + a, o, s.C[a]);
  if (with_f) {
#pragma unroll
    for (int a = 0; a < 9; ++a) io.st(PL_F + a, o, s.F[a]);
  }
  io.st(PL_JP, o, s.Jp);
}

__device__ __forceinline__ void gather_global(const PState& s, const int base[3], const float w[3][3],
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
  // A stencil inside the grid writes node (base + offset) itself: its node
  // blocks are flagged once, at most 2 per axis. Elsewhere (the reference's
  // aliasing of j, k >= N) each node's block comes from its linear index.
  const uint32_t nbx = ((uint32_t)N + 3u) >> 2;
  const bool plain = base[0] >= 0 && base[1] >= 0 && base[2] >= 0 && base[0] + 2 < N &&
                     base[1] + 2 < N && base[2] + 2 < N;
  if (dirty && plain) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
          dirty[(((uint32_t)(base[0] + 2 * i) >> 2) * nbx + ((uint32_t)(base[1] + 2 * j) >> 2)) * nbx +
                ((uint32_t)(base[2] + 2 * kk) >> 2)] = 1u;
  }
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
        if (dirty && !plain) {
          // idx's own coordinates: 0 <= base + offset <= N + 1 on every axis
          // (base < N inside the domain), so a component at or past N carries
          // once into the next axis; idx < N^3 leaves gx < N
          uint32_t gz = (uint32_t)(base[2] + kk), gy = (uint32_t)(base[1] + j), gx = (uint32_t)(base[0] + i);
          if (gz >= (uint32_t)N) { gz -= (uint32_t)N; ++gy; }
          if (gy >= (uint32_t)N) { gy -= (uint32_t)N; ++gx; }
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

// Wavefront lane shifts and an inclusive max-scan with DPP (no LDS traffic).
__device__ __forceinline__ int dpp_wave_shr1(int v) {  // lane l gets lane l-1; lane 0 gets 0
  return __builtin_amdgcn_update_dpp(0, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, true);
}
__device__ __forceinline__ int dpp_wave_shl1(int v) {  // lane l gets lane l+1; lane 63 gets 0
  return __builtin_amdgcn_update_dpp(0, v, 0x130 /* wave_shl:1 */, 0xf, 0xf, true);
}
__device__ __forceinline__ int wave_max_scan(int v) {  // v >= 0
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111 /* row_shr:1 */, 0xf, 0xf, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112 /* row_shr:2 */, 0xf, 0xf, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114 /* row_shr:4 */, 0xf, 0xf, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118 /* row_shr:8 */, 0xf, 0xf, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142 /* row_bcast:15 */, 0xa, 0xf, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143 /* row_bcast:31 */, 0xc, 0xf, true));
  return v;
}

// LDS slot of tile node (x, y, z) in an 8^3 tile: z is permuted within its
// row (bit 2 flipped when x + y is odd), so the runs of 4 neighbouring nodes
// that the lanes of one access touch in adjacent rows or planes land in
// different halves of the 32 LDS banks instead of the same ones.
__device__ __forceinline__ int tile_slot(int x, int y, int z) {
  return (x * mpmk::kTileW + y) * mpmk::kTileW + (z ^ (((x ^ y) & 1) << 2));
}

// The 27 stencil slots of a cell at tile coordinates (l0, l1, l2):
// slot(a, b, c) = row + (a W + b) W + z[(a ^ b) & 1][c].
struct StencilSlots {
  int row;
  int z[2][3];
  __device__ __forceinline__ StencilSlots(int l0, int l1, int l2) {
    constexpr int W = mpmk::kTileW;
    row = (l0 * W + l1) * W;
    const int f = (l0 ^ l1) & 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      z[0][c] = (l2 + c) ^ (f << 2);
      z[1][c] = (l2 + c) ^ ((f ^ 1) << 2);
    }
  }
  __device__ __forceinline__ int operator()(int a, int b, int c) const {
    constexpr int W = mpmk::kTileW;
    return row + (a * W + b) * W + z[(a ^ b) & 1][c];
  }
};

// One work item of the fused pass (one workgroup of 64 kItemWaves threads).
template <int KS>
__device__ __forceinline__ void g2p2g_item(const mpmk::ItemDesc& d, const int item) {
  const ParticlesDev P = d.P, Pdst = d.Pdst;
  const uint32_t* __restrict__ const gather = d.gather;
  const MaterialDev* __restrict__ const mats = d.mats;
  const StepConsts k = d.k;
  const uint32_t* __restrict__ const item_start = d.item_start;
  const uint32_t* __restrict__ const item_block = d.item_block;
  const uint32_t* __restrict__ const item_waves = d.item_waves;
  const float4* __restrict__ const grid_v = d.grid_v;
  float4* __restrict__ const tiles = d.tiles;
  float* __restrict__ const grid_atomic = d.grid_atomic;
  uint32_t* __restrict__ const dirty = d.dirty;
  const int flags = d.flags;
  constexpr int BE = 1 << mpmk::kBlockLog2;
  constexpr int W = mpmk::kTileW;        // BE + 2 (stencil reach) + 2 (drift margin)
  constexpr int TN = mpmk::kTileNodes;
  constexpr int NCE = W - 2;             // cells per axis whose stencil fits the tile
  constexpr int NW = mpmk::kItemWaves, NT = 64 * NW;
  static_assert(NCE * NCE * NCE <= 256, "claim bytes per wavefront");
  __shared__ float4 acc[NW][TN];         // P2G tile of each wavefront
  __shared__ uint8_t claims[NW][256];    // claim byte per cell, per wavefront
  // The grid velocities for G2P share wavefront 0's tile: wavefront 0 clears
  // it after the barrier that ends every G2P read.
  float4* const vt = acc[0];

  if ((uint32_t)item >= *d.item_count) return;  // uniform for the workgroup
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

  const bool do_g2p = (flags & mpmk::TILED_G2P) != 0;
  // Liquid with F derived from J: every particle's F is exactly cbrt(Jp) I
  // (the state F = I, Jp = 1 at upload, and what this kernel computes), so
  // the F planes are neither read nor written; F_old = cbrt(Jp) I is rebuilt
  // in registers with the same bits and the engine writes the planes out
  // before anything else reads them.
  const bool df = KS == KINDS_LIQUID && (flags & mpmk::TILED_DERIVE_F) != 0;
  // the original-index guard of G2P (compat REFERENCE) can bind only when
  // some particle index exceeds it; otherwise the ids are read only to move
  // them with a reorder
  const bool guard = (flags & mpmk::TILED_GUARD) != 0;
  const bool valid = lane < cnt_w;
  const uint32_t i = start + (uint32_t)(off_w + lane);
  PState s;
  MaterialDev m;
  float R[9];
  bool has_R = false;
  int base[3] = {0, 0, 0};
  float w[3][3];
  bool inside = false;
  // After a sort the particles still sit in the order before it: thread i
  // reads particle gather[i] of P and writes its state to slot i of Pdst.
  // The particle loads are issued before the tile staging so that both are
  // in flight together.
  const PlaneIO io(P), io_dst(Pdst);
  uint32_t src = 0, pid = 0;
  uint8_t mat = 0;
  if (valid) {
    src = gather ? gather[i] : i;
    if (df) load_xj(io, src, s);
    else load_xfj(io, src, s);
    mat = P.mat[src];
    if (gather || guard) pid = P.id[src];
    m = mats[mat];
  }

  // Tile node n = (a W + tb) W + tc: thread tid takes the nodes with
  // n mod 64 = tid mod 64 (its tb, tc) and a = tid / 64 + 3 j.
  static_assert(NT % 64 == 0 && W * W == 64, "tile rows per wavefront");
  const int tb = (lane >> 3) & 7, tc = lane & 7;
  const int slot_even = tb * W + (tc ^ ((tb & 1) << 2)), slot_odd = tb * W + (tc ^ (((tb & 1) ^ 1) << 2));
  if (do_g2p) {
    const int gj = o1 + tb, gk = o2 + tc;
    const bool jk_in = gj >= 0 && gj < N && gk >= 0 && gk < N;
    const uint32_t jk = (uint32_t)gj * (uint32_t)N + (uint32_t)gk;
    for (int a = wave; a < W; a += NT / 64) {
      const int gi = o0 + a;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (jk_in && gi >= 0 && gi < N) v = grid_v[(size_t)(uint32_t)gi * ((size_t)N * N) + jk];
      vt[a * W * W + ((a & 1) ? slot_odd : slot_even)] = v;
    }
  }
  {
    float4* const flat = &acc[1][0];
    for (int n = tid; n < (NW - 1) * TN; n += NT) flat[n] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  __syncthreads();
  // The material arrives with the particle loads; take it into registers here,
  // before any particle store: vector-memory waits count stores too, so a
  // first use after the stores would wait for all of them to complete.
  if (valid) {
    // (through an opaque move, so the compiler keeps the registers instead of
    // loading the read-only table again at a later use)
    __asm__ __volatile__("" : "+v"(m.kind), "+v"(m.volume), "+v"(m.mass), "+v"(m.mu0),
                         "+v"(m.lambda0), "+v"(m.hardening), "+v"(m.clamp_lo), "+v"(m.clamp_hi),
                         "+v"(m.bulk), "+v"(m.stress_scale));
  }

  if (valid) {
    if (gather) {
      Pdst.mat[i] = mat;
      Pdst.id[i] = pid;
    }
    quad_weights(s.x, k.dx_inv, base, w);
    const bool moved = do_g2p && (!guard || pid <= k.guard) && !outside_domain(base, N);
    if (!moved) {
      load_vc(io, src, s);
      // complete these loads here: left in flight they would meet the
      // particle stores of the other lanes at the next vector-memory wait
      __asm__ __volatile__("" : "+v"(s.v[0]), "+v"(s.v[1]), "+v"(s.v[2]), "+v"(s.C[0]), "+v"(s.C[1]),
                           "+v"(s.C[2]), "+v"(s.C[3]), "+v"(s.C[4]), "+v"(s.C[5]), "+v"(s.C[6]),
                           "+v"(s.C[7]), "+v"(s.C[8]));
    }
    if (moved) {
      // ---- G2P of substep n, gathering the 27 nodes from the LDS tile
      const int l0 = base[0] - o0, l1 = base[1] - o1, l2 = base[2] - o2;
      if (l0 >= 0 && l1 >= 0 && l2 >= 0 && l0 < NCE && l1 < NCE && l2 < NCE) {
        // v = sum w g, C = sum w g dd^T (g2p_node_contribution, TransferScheme.h:124-132),
        // summed axis by axis: c innermost, then b, then a
        float wd[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int o = 0; o < 3; ++o)
            wd[a][o] = w[a][o] * (((float)(uint32_t)(base[a] + o) * k.dx - s.x[a]) * k.dinv);
        const StencilSlots slot(l0, l1, l2);
        float V[3][3], X[3][3], Y[3][3];  // [a][component]
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
          for (int i = 0; i < 3; ++i) { V[a][i] = 0.f; X[a][i] = 0.f; Y[a][i] = 0.f; }
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            float T[3] = {0.f, 0.f, 0.f}, U[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const float4 g = vt[slot(a, b, c)];
              T[0] += w[2][c] * g.x; T[1] += w[2][c] * g.y; T[2] += w[2][c] * g.z;
              U[0] += wd[2][c] * g.x; U[1] += wd[2][c] * g.y; U[2] += wd[2][c] * g.z;
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              V[a][i] += w[1][b] * T[i];
              X[a][i] += wd[1][b] * T[i];
              Y[a][i] += w[1][b] * U[i];
            }
          }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          float vi = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            vi += w[0][a] * V[a][i];
            c0 += wd[0][a] * V[a][i];
            c1 += w[0][a] * X[a][i];
            c2 += w[0][a] * Y[a][i];
          }
          s.v[i] = vi;
          s.C[3 * i + 0] = c0;
          s.C[3 * i + 1] = c1;
          s.C[3 * i + 2] = c2;
        }
      } else {
        gather_global(s, base, w, k, grid_v, s.v, s.C);
      }
      float A[9];
#pragma unroll
      for (int a = 0; a < 9; ++a) A[a] = ((a % 4 == 0) ? 1.0f : 0.0f) + k.dt * s.C[a];
      if (df) {
        // F_old = c0 I with c0 = cbrt(Jp) (the stored F_old, bit for bit), so
        // F_new = A F_old is c0 A elementwise: the full product's bits, its
        // other terms being products with zero
        const float c0 = cbrtf(s.Jp);
#pragma unroll
        for (int a = 0; a < 9; ++a) s.F[a] = A[a] * c0;
      } else {
        mul_ab(A, s.F, s.F);
      }
      if (KS == KINDS_LIQUID) {
        // F <- J^(1/3) I, Jp <- J (the only kind this instantiation handles).
        // With F derived from J the new F is neither stored nor used (the
        // liquid's stress reads J): its cube root is left to the next pass,
        // which rebuilds F_old = cbrt(Jp) I with the same bits.
        const float J = det3(s.F);
        if (!df) {
          const float cj = cbrtf(J);
#pragma unroll
          for (int a = 0; a < 9; ++a) s.F[a] = (a % 4 == 0) ? cj : 0.0f;
        }
        s.Jp = J;
      } else {
        plasticity(s.F, s.Jp, m, R, has_R);
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) s.x[a] = s.x[a] + k.dt * s.v[a];
      quad_weights(s.x, k.dx_inv, base, w);
    }
    if (moved || gather) store_state(io_dst, i, s, !df);
    inside = !outside_domain(base, N);
  }

  // ---- P2G preparation (TransferScheme.h:66-86): stress and affine momentum.
  // Node value (a,b,c) = wt * ((q0[.][a] + q1[.][b]) + q2[.][c]), with
  // q0 = m v + aff(:,0) d0, q1 = aff(:,1) d1, q2 = aff(:,2) d2
  // (p2g_node_contribution, TransferScheme.h:88-100, regrouped by axis).
  int cell = 0, l0 = 0, l1 = 0, l2 = 0;
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
    for (int a = 0; a < 9; ++a)  // (the liquid's tau is diagonal: its zeros are left out)
      aff[a] = (KS == KINDS_LIQUID && a % 4 != 0) ? m.mass * s.C[a]
                                                  : m.stress_scale * tau[a] + m.mass * s.C[a];
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
    l0 = base[0] - o0;
    l1 = base[1] - o1;
    l2 = base[2] - o2;
    // tile path: the stencil lies in the tile and inside the grid (at the grid
    // faces the reference's node bounds and aliasing rules apply: direct path)
    const bool in_tile = l0 >= 0 && l1 >= 0 && l2 >= 0 && l0 < NCE && l1 < NCE && l2 < NCE &&
                         base[0] >= 0 && base[1] >= 0 && base[2] >= 0 && base[0] + 2 < N &&
                         base[1] + 2 < N && base[2] + 2 < N;
    if (in_tile) {
      cell = (l0 * NCE + l1) * NCE + l2;
      pending = true;
    } else {
      scatter_global(q0, q1, q2, m.mass, base, w, k, grid_atomic, dirty);
    }
  }
  __syncthreads();  // every G2P read of vt is done
  if (wave == 0) {
#pragma unroll
    for (int n = lane; n < TN; n += 64) acc[0][n] = make_float4(0.f, 0.f, 0.f, 0.f);
  }

  // ---- P2G into the wavefront's tile.
  // Per slab of 9 nodes (x offset a):
  // 1. the lane's 36 node contributions (9 nodes x momentum, mass);
  // 2. lanes holding the same cell form contiguous runs (the particles are in
  //    cell order within a wavefront); a segmented inclusive scan over the
  //    wavefront with DPP lane shifts leaves each run's total in its last
  //    lane: shifts of 1, 2, 4, 8 within rows of 16 and the two
  //    row-broadcast steps that carry sums across rows. Done only when some
  //    run is longer than 4. Shorter runs merge pairs of lanes with one
  //    whole-wavefront shift (odd ranks add their left neighbour) and go
  //    through step 3 with at most two writers per run.
  // 3. the writers (run tails after a reduction, odd ranks and even tails
  //    after a pair merge, else every lane with a tile particle) add their
  //    9 nodes with plain read-modify-writes, in rounds:
  //    every round a writer claims its cell (one byte per cell) and only the
  //    lane whose claim stands writes (particles that moved since the last
  //    sort can also put one cell in two runs). Within
  //    one read-modify-write the writing lanes address distinct nodes; the
  //    compiler barrier keeps the read-modify-writes in program order (a
  //    lane's next read may address a node another lane writes this round).
  // A wavefront without tile particles (an item needs ceil(n / 64) of the
  // workgroup's wavefronts; ~20% of them idle on a liquid scene's boundary)
  // has nothing to add: its tile stays as cleared.
  if (__ballot(pending) != 0ull) {
    // runs of equal cells (DPP lane shifts; lanes without a tile particle are
    // runs of their own)
    const int cid1 = (pending ? cell : 1024 + lane) + 1;
    const int prev = dpp_wave_shr1(cid1), next = dpp_wave_shl1(cid1);  // 0 beyond the ends
    const int head = wave_max_scan(prev != cid1 ? lane : 0);
    const int rank = lane - head;  // position within the run
    const bool tail = next != cid1;
    const int maxrun = __builtin_amdgcn_readlane(wave_max_scan(rank + 1), 63);
    const bool reduce = maxrun > 4;
    // runs of 2 to 4: one wavefront shift merges pairs, the odd ranks taking
    // their left neighbour's sums; a run then has at most 2 writers
    const bool pair = !reduce && maxrun > 1;
    const bool odd = (rank & 1) != 0;
    const float take = odd ? 1.0f : 0.0f;
    const int lr = lane & 15;

    typedef __attribute__((address_space(3))) volatile uint8_t lds_byte;
    lds_byte* const claim = (lds_byte*)(&claims[wave][0]);
    float4* const my = acc[wave];
    const StencilSlots slot(l0, l1, l2);
    // Writers (see step 3 above) claim their cell; the lanes whose claim stands add their nodes,
    // one slab of 9 nodes (offset a along x) at a time so that 36 sums sit in
    // registers. Writers that lost (a cell with more than one writer in the
    // wavefront) retry slab by slab in further claim rounds.
    const bool writer = pending && (reduce ? tail : (!pair || odd || tail));
    if (writer) claim[cell] = (uint8_t)lane;
    const bool mine = writer && claim[cell] == (uint8_t)lane;
    const bool lost = writer && !mine;
    const bool any_lost = __ballot(lost) != 0ull;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float cv[9][4];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const float wab = w[0][a] * w[1][b];
        const float s0 = q0[0][a] + q1[0][b], s1 = q0[1][a] + q1[1][b], s2 = q0[2][a] + q1[2][b];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int o = b * 3 + c;
          const float wt = pending ? wab * w[2][c] : 0.0f;
          cv[o][0] = wt * (s0 + q2[0][c]);
          cv[o][1] = wt * (s1 + q2[1][c]);
          cv[o][2] = wt * (s2 + q2[2][c]);
          cv[o][3] = wt * m.mass;
          if (pair) {  // odd ranks add their left neighbour (whole-wavefront shift)
            // as cv += shifted cv * (0 or 1): one multiply-add reading its
            // first operand through the lane shift (exact: t * 1 = t, t * 0 =
            // 0). Every lane of the wavefront is active here. The s_nop gives
            // the two wait states a DPP read needs after the VALU writes of
            // these registers (the compiler does not pad inside inline asm).
            __asm__(
                "s_nop 1\n\t"
                "v_fmac_f32_dpp %0, %0, %4 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                "v_fmac_f32_dpp %1, %1, %4 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                "v_fmac_f32_dpp %2, %2, %4 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                "v_fmac_f32_dpp %3, %3, %4 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                : "+v"(cv[o][0]), "+v"(cv[o][1]), "+v"(cv[o][2]), "+v"(cv[o][3])
                : "v"(take));
          }
        }
      }
      // one step of the segmented scan: cv += (lane shifted by CTRL) * f, f = 1
      // where COND holds, else 0 (exact as in the pair merge: t * 1 = t, and
      // t * 0 adds a zero, which leaves every tile sum's bits as they are),
      // one v_fmac_f32_dpp per value; ROWS masks the 16-lane rows written
#define MPM_SEG_NODE(o, CTRL, ROWS)                                                          \
      __asm__("s_nop 1\n\t"                                                                   \
              "v_fmac_f32_dpp %0, %0, %4 " CTRL " row_mask:" ROWS " bank_mask:0xf bound_ctrl:1\n\t" \
              "v_fmac_f32_dpp %1, %1, %4 " CTRL " row_mask:" ROWS " bank_mask:0xf bound_ctrl:1\n\t" \
              "v_fmac_f32_dpp %2, %2, %4 " CTRL " row_mask:" ROWS " bank_mask:0xf bound_ctrl:1\n\t" \
              "v_fmac_f32_dpp %3, %3, %4 " CTRL " row_mask:" ROWS " bank_mask:0xf bound_ctrl:1"       \
              : "+v"(cv[o][0]), "+v"(cv[o][1]), "+v"(cv[o][2]), "+v"(cv[o][3])                  \
              : "v"(f_));
#define MPM_SEG_STEP(CTRL, ROWS, COND)                                                      \
      {                                                                                      \
        const float f_ = (COND) ? 1.0f : 0.0f;                                               \
        _Pragma("unroll") for (int o = 0; o < 9; ++o) { MPM_SEG_NODE(o, CTRL, ROWS) }        \
      }
      if (reduce) {
        MPM_SEG_STEP("row_shr:1", "0xf", rank >= 1 && lr >= 1)
        MPM_SEG_STEP("row_shr:2", "0xf", rank >= 2 && lr >= 2)
        MPM_SEG_STEP("row_shr:4", "0xf", rank >= 4 && lr >= 4)
        if (maxrun > 8) MPM_SEG_STEP("row_shr:8", "0xf", rank >= 8 && lr >= 8)
        // carry across rows: rows 1, 3 take lane 15 of the row below, rows 2, 3 lane 31
        MPM_SEG_STEP("row_bcast:15", "0xa", ((lane >> 4) & 1) && rank > lr)
        MPM_SEG_STEP("row_bcast:31", "0xc", lane >= 32 && rank > lane - 32)
      }
#undef MPM_SEG_STEP
#undef MPM_SEG_NODE
#define MPM_RMW_SLAB                                                                       \
      _Pragma("unroll") for (int b = 0; b < 3; ++b) _Pragma("unroll") for (int c = 0; c < 3; ++c) { \
        const int o = b * 3 + c;                                                             \
        float4* const p = my + slot(a, b, c);                                                \
        float4 t = *p;                                                                       \
        t.x += cv[o][0];                                                                     \
        t.y += cv[o][1];                                                                     \
        t.z += cv[o][2];                                                                     \
        t.w += cv[o][3];                                                                     \
        *p = t;                                                                              \
        __asm__ __volatile__("" ::: "memory");                                               \
      }
      if (mine) { MPM_RMW_SLAB }
      if (any_lost) {
        bool todo = lost;
        while (__ballot(todo) != 0ull) {
          if (todo) claim[cell] = (uint8_t)lane;
          if (todo && claim[cell] == (uint8_t)lane) {
            MPM_RMW_SLAB
            todo = false;
          }
        }
      }
#undef MPM_RMW_SLAB
    }
  }
  __syncthreads();

  // ---- the item's tile: sum of the wavefront tiles, fixed order
  float4* const out = tiles + (size_t)item * TN;
  for (int a = wave; a < W; a += NT / 64) {
    const int n = a * W * W + lane;
    const int sl = a * W * W + ((a & 1) ? slot_odd : slot_even);
    float4 t = acc[0][sl];
#pragma unroll
    for (int q = 1; q < NW; ++q) {
      const float4 u = acc[q][sl];
      t.x += u.x; t.y += u.y; t.z += u.z; t.w += u.w;
    }
    out[n] = t;
  }
}

// Workgroup b of a launch of g workgroups over n units, XCD-contiguous: the
// units are cut into kXcds runs of ceil(n / kXcds) and XCD b % kXcds takes
// its run in order. Returns n (no unit) for the workgroups beyond, and b
// itself when the launch is too small to remap.
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t g, uint32_t n) {
  constexpr uint32_t X = mpmk::kXcds;
  const uint32_t per = (n + X - 1) / X;
  if (X * per > g) return b;
  const uint32_t j = b / X;
  return j < per ? (b % X) * per + j : n;
}

// Items first, first + B, ... of a scene below its item count, in turn (a
// batch scene that outgrew its workgroups). Out of line: the loop would hold
// its state in registers through the whole item body, and the first item,
// the common case, keeps the registers of a single pass.
template <int KS>
__device__ __noinline__ void g2p2g_further_items(const mpmk::ItemDesc& d, uint32_t first, uint32_t B) {
  for (uint32_t item = first; item < *d.item_count; item += B) {
    __syncthreads();  // the previous item's flush read the tiles this one stages
    g2p2g_item<KS>(d, (int)item);
  }
}

template <int KS>
__global__ void __launch_bounds__(64 * mpmk::kItemWaves)
    __attribute__((amdgpu_waves_per_eu(KS == KINDS_LIQUID ? 5 : 4)))
    k_g2p2g_tiled(mpmk::ItemDesc d) {
  const uint32_t b = blockIdx.x;
  g2p2g_item<KS>(d, (int)((d.flags & mpmk::TILED_XCD) ? xcd_order(b, gridDim.x, *d.item_count) : b));
}

template <int KS>
__global__ void __launch_bounds__(64 * mpmk::kItemWaves)
    __attribute__((amdgpu_waves_per_eu(KS == KINDS_LIQUID ? 5 : 4)))
    k_g2p2g_multi(const mpmk::ItemDesc* __restrict__ descs, const uint32_t* __restrict__ prefix,
                  int nscenes) {
  // workgroup g takes item l = g - prefix[j] of its scene j; the scene's B
  // workgroups are sized from an estimate of its items, and when the scene
  // has more, every workgroup takes every B-th further one in turn
  const uint32_t g = blockIdx.x;
  const int j = scene_of_wave(prefix, nscenes, g);
  const mpmk::ItemDesc& d = descs[j];
  const uint32_t l = g - prefix[j];
  g2p2g_item<KS>(d, (int)l);
  const uint32_t B = prefix[j + 1] - prefix[j];
  if (l + B < *d.item_count) g2p2g_further_items<KS>(d, l + B, B);  // uniform for the workgroup
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
template <bool MLP>
__device__ __forceinline__ void reduce_block(const mpmk::ReduceDesc& d, const int64_t nbk, const int lane) {
  constexpr int W = mpmk::kTileW, TN = mpmk::kTileNodes;
  const float4* __restrict__ const tiles = d.tiles;
  const uint32_t* __restrict__ const block_items = d.block_items;
  const uint32_t* __restrict__ const covered = d.covered;
  float4* __restrict__ const grid_atomic = d.grid_atomic;
  uint32_t* __restrict__ const dirty = d.dirty;
  float4* __restrict__ const grid = d.grid;
  const StepConsts k = d.k;
  const int nbx = d.nbx, update = d.update;
  // (32-bit: at most 256^3 node blocks)
  const uint32_t nb32 = (uint32_t)nbk, nbxu = (uint32_t)nbx;
  const int bz = (int)(nb32 % nbxu), by = (int)((nb32 / nbxu) % nbxu), bx = (int)(nb32 / (nbxu * nbxu));
  const int u0 = lane >> 4, u1 = (lane >> 2) & 3, u2 = lane & 3;
  const int N = k.N;
  const int x = bx * 4 + u0, y = by * 4 + u1, z = bz * 4 + u2;
  const bool in_grid = x < N && y < N && z < N;
  const size_t node = ((size_t)x * N + (size_t)y) * N + (size_t)z;
  // the item ranges of the 27 home blocks around (lane j < 27: neighbour j),
  // read in the same round trip as the block's flags
  uint32_t r0 = 0, r1 = 0;
  if (lane < 27) {
    const int hx = bx + lane / 9 - 1, hy = by + (lane / 3) % 3 - 1, hz = bz + lane % 3 - 1;
    if (hx >= 0 && hx < nbx && hy >= 0 && hy < nbx && hz >= 0 && hz < nbx) {
      const uint32_t mb = morton3((uint32_t)hx, (uint32_t)hy, (uint32_t)hz);
      const uint2 r = *reinterpret_cast<const uint2*>(block_items + 2 * (size_t)mb);
      r0 = r.x;
      r1 = r.y;
    }
  }
  const bool is_dirty = dirty[nbk] != 0u;
  const bool live = covered[nbk] || is_dirty;
  if (d.live && lane == 0) d.live[nbk] = live ? 1u : 0u;
  if (!live) {  // no tile and no side-grid scatter reaches it
    if (!d.skip_dead && in_grid) grid[node] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
  if (__ballot(r0 < r1) != 0ull) {  // some home block around holds items
    const int e0 = (u0 == 3) ? 0 : -1, e1 = (u1 == 3) ? 0 : -1, e2 = (u2 == 3) ? 0 : -1;
    // per home block q the sum of its items in item order (what the combine
    // pass leaves in the block's first tile), then the 8 block sums in order.
    // The first items of the 8 blocks are read in one batch of independent
    // loads; blocks holding more than one item (dense scenes) add the rest
    // after it: MLP = false block by block, one load in flight (fewer
    // registers, more waves resident), MLP = true round by round, item r of
    // every block in one batch of 8 loads. The sums are the same bits.
    float4 v[8];
    uint32_t a0[8], a1[8];
    uint32_t more = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int f0 = e0 + (q >> 2), f1 = e1 + ((q >> 1) & 1), f2 = e2 + (q & 1);
      const int j = 9 * (f0 + 1) + 3 * (f1 + 1) + (f2 + 1);
      const uint32_t b0 = (uint32_t)__shfl((int)r0, j), b1 = (uint32_t)__shfl((int)r1, j);
      const int t = ((u0 + 1 - 4 * f0) * W + (u1 + 1 - 4 * f1)) * W + (u2 + 1 - 4 * f2);
      v[q] = b0 < b1 ? tiles[(size_t)b0 * TN + t] : make_float4(0.f, 0.f, 0.f, 0.f);
      more = max(more, b1 > b0 ? b1 - b0 : 0u);
      if (MLP) {
        a0[q] = b0 + 1;  // next item
        a1[q] = b1;
      }
    }
    if (!d.combined && __ballot(more > 1) != 0ull) {  // (combined: the first tile holds the block's sum)
      if (MLP) {
        for (uint32_t r = 1; r < more; ++r) {
          float4 u[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int f0 = e0 + (q >> 2), f1 = e1 + ((q >> 1) & 1), f2 = e2 + (q & 1);
            const int t = ((u0 + 1 - 4 * f0) * W + (u1 + 1 - 4 * f1)) * W + (u2 + 1 - 4 * f2);
            u[q] = a0[q] < a1[q] ? tiles[(size_t)a0[q] * TN + t] : make_float4(0.f, 0.f, 0.f, 0.f);
          }
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            if (a0[q] < a1[q]) { v[q].x += u[q].x; v[q].y += u[q].y; v[q].z += u[q].z; v[q].w += u[q].w; }
            ++a0[q];
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int f0 = e0 + (q >> 2), f1 = e1 + ((q >> 1) & 1), f2 = e2 + (q & 1);
          const int j = 9 * (f0 + 1) + 3 * (f1 + 1) + (f2 + 1);
          const uint32_t b0 = (uint32_t)__shfl((int)r0, j), b1 = (uint32_t)__shfl((int)r1, j);
          const int t = ((u0 + 1 - 4 * f0) * W + (u1 + 1 - 4 * f1)) * W + (u2 + 1 - 4 * f2);
          for (uint32_t it = b0 + 1; it < b1; ++it) {
            const float4 u = tiles[(size_t)it * TN + t];
            v[q].x += u.x; v[q].y += u.y; v[q].z += u.z; v[q].w += u.w;
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      sum.x += v[q].x; sum.y += v[q].y; sum.z += v[q].z; sum.w += v[q].w;
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

template <bool MLP>
__global__ void __launch_bounds__(256) k_tile_reduce(mpmk::ReduceDesc d) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t nbk = (int64_t)blockIdx.x * 4 + wave;
  if (nbk >= (int64_t)d.nbx * d.nbx * d.nbx) return;  // uniform for the wavefront
  reduce_block<MLP>(d, nbk, lane);
}

// threads of a combine workgroup: two tile nodes each
constexpr int kCombineThreads = mpmk::kTileNodes / 2;
// home block b's tiles: tiles[first] += tiles[first + 1 .. end), in item order
__device__ __forceinline__ void combine_block(const mpmk::ReduceDesc& d, const int64_t b) {
  constexpr int TN = mpmk::kTileNodes;
  const uint32_t r0 = d.block_items[2 * (size_t)b], r1 = d.block_items[2 * (size_t)b + 1];
  if (r1 <= r0 + 1) return;  // uniform for the workgroup
  float4* const t0 = const_cast<float4*>(d.tiles) + (size_t)r0 * TN;  // the tiles are the engine's scratch
  // each thread sums its two nodes over the block's items in item order:
  // rounds of 16 items of both nodes in flight (the densest blocks hold
  // dozens), then the rest (most crowded blocks hold a few) in rounds of 4
  // whose loads are clamped to the block's last item, so that all 4 are
  // issued at once. A load past the block is added times 0 (fma(v, 0, a) =
  // a + (+-0) = a: a tile node is never -0, its sums start from +0), one of
  // the block times 1 (fma(v, 1, a) = a + v in one rounding): the bits of
  // the plain sum. An add under a branch would let the compiler sink each
  // load into it, one load in flight at a time.
  static_assert(TN == 2 * kCombineThreads, "two tile nodes per thread");
  const int n0 = threadIdx.x, n1 = threadIdx.x + kCombineThreads;
  float4 a0 = t0[n0], a1 = t0[n1];
  const float4* __restrict__ const tiles = d.tiles;
  uint32_t it = r0 + 1;
  constexpr uint32_t B = 16;
  for (; it + B <= r1; it += B) {
    float4 v[B], u[B];
#pragma unroll
    for (uint32_t q = 0; q < B; ++q) {
      v[q] = tiles[(size_t)(it + q) * TN + n0];
      u[q] = tiles[(size_t)(it + q) * TN + n1];
    }
#pragma unroll
    for (uint32_t q = 0; q < B; ++q) {
      a0.x += v[q].x; a0.y += v[q].y; a0.z += v[q].z; a0.w += v[q].w;
      a1.x += u[q].x; a1.y += u[q].y; a1.z += u[q].z; a1.w += u[q].w;
    }
  }
  constexpr uint32_t T = 4;
  for (; it < r1; it += T) {
    float4 v[T], u[T];
#pragma unroll
    for (uint32_t q = 0; q < T; ++q) {
      const size_t j = (size_t)min(it + q, r1 - 1u) * TN;
      v[q] = tiles[j + n0];
      u[q] = tiles[j + n1];
    }
#pragma unroll
    for (uint32_t q = 0; q < T; ++q) {
      const float m = it + q < r1 ? 1.0f : 0.0f;
      a0.x = __builtin_fmaf(v[q].x, m, a0.x); a0.y = __builtin_fmaf(v[q].y, m, a0.y);
      a0.z = __builtin_fmaf(v[q].z, m, a0.z); a0.w = __builtin_fmaf(v[q].w, m, a0.w);
      a1.x = __builtin_fmaf(u[q].x, m, a1.x); a1.y = __builtin_fmaf(u[q].y, m, a1.y);
      a1.z = __builtin_fmaf(u[q].z, m, a1.z); a1.w = __builtin_fmaf(u[q].w, m, a1.w);
    }
  }
  t0[n0] = a0;
  t0[n1] = a1;
}

// The crowded home blocks of one scene (the list its sort built): workgroup
// w combines the blocks whose index in the list is w modulo the grid.
__global__ void __launch_bounds__(kCombineThreads) k_tile_combine(mpmk::ReduceDesc d) {
  const uint32_t cq = d.crowded[0];
  for (uint32_t e = blockIdx.x; e < cq; e += gridDim.x) combine_block(d, (int64_t)d.crowded[1 + e]);
}

// The crowded home blocks (several items) of every scene, from the lists the
// sorts built: the scenes' lists are taken as one concatenation, and
// workgroup w combines the blocks whose index in it is w modulo the grid.
// Every wavefront reads the scenes' counts lane by lane (64 scenes per round
// trip) and their running sum, so all 4 take the same (uniform) loop.
// 512 workgroups: in a batch the combine runs beside the other chain's fused
// pass, and a larger grid takes it more of the chip than it gains (64 scenes:
// 1,024 workgroups cost ~1% per substep, 2,048 ~3%, 4,096 ~7%; 256 slows the
// single-scene combine of dense scenes).
constexpr int kCombineGrid = 512;
__global__ void __launch_bounds__(kCombineThreads)
    k_tile_combine_multi(const mpmk::ReduceDesc* __restrict__ descs, int nscenes) {
  const int lane = (int)(threadIdx.x & 63u);
  const uint32_t G = gridDim.x, w = blockIdx.x;
  uint32_t base = 0;  // blocks of the scenes before this chunk of 64
  for (int c = 0; c < nscenes; c += 64) {
    const int j = c + lane;
    const uint32_t cnt = j < nscenes ? descs[j].crowded[0] : 0u;
    uint32_t incl = cnt;  // inclusive sum over the lanes
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = (uint32_t)__shfl_up((int)incl, o);
      if (lane >= o) incl += u;
    }
    const int m = min(64, nscenes - c);
    for (int q = 0; q < m; ++q) {
      const uint32_t cq = (uint32_t)__shfl((int)cnt, q);
      if (cq == 0u) continue;
      const uint32_t first = base + (uint32_t)__shfl((int)incl, q) - cq;  // index of the scene's first block
      const mpmk::ReduceDesc& d = descs[c + q];
      for (uint32_t e = (w + G - first % G) % G; e < cq; e += G) combine_block(d, (int64_t)d.crowded[1 + e]);
    }
    base += (uint32_t)__shfl((int)incl, m - 1);
  }
}

__global__ void __launch_bounds__(256)
    k_tile_reduce_multi(const mpmk::ReduceDesc* __restrict__ descs, const uint32_t* __restrict__ prefix,
                        int nscenes) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t g = blockIdx.x * 4u + (uint32_t)wave;
  if (g >= prefix[nscenes]) return;  // uniform for the wavefront
  const int j = scene_of_wave(prefix, nscenes, g);
  reduce_block<false>(descs[j], (int64_t)(g - prefix[j]), lane);
}

namespace mpmk {

hipError_t g2p2g_desc(const ItemDesc& d, int nmats, int64_t launch, int kinds, hipStream_t s) {
  if (d.P.n <= 0 || launch <= 0) return hipSuccess;
  if (nmats < 1 || nmats > MPM_MAX_MATERIALS) return hipErrorInvalidValue;
  if ((d.flags & TILED_G2P) && !d.grid_v) return hipErrorInvalidValue;
  if (!d.tiles || !d.grid_atomic || !d.dirty) return hipErrorInvalidValue;
  if (launch > 0x7fffffff - kXcds) return hipErrorInvalidValue;
  if (d.gather && (d.Pdst.f == d.P.f || d.Pdst.n != d.P.n)) return hipErrorInvalidValue;
  if (!d.gather && d.Pdst.f != d.P.f) return hipErrorInvalidValue;
  if (d.flags & TILED_XCD) launch = (launch + kXcds - 1) / kXcds * kXcds;
  if (kinds == KINDS_LIQUID)
    hipLaunchKernelGGL((k_g2p2g_tiled<KINDS_LIQUID>), dim3((unsigned)launch), dim3(64 * kItemWaves), 0, s, d);
  else
    hipLaunchKernelGGL((k_g2p2g_tiled<KINDS_ALL>), dim3((unsigned)launch), dim3(64 * kItemWaves), 0, s, d);
  return hipGetLastError();
}

hipError_t g2p2g_multi(const ItemDesc* descs, const uint32_t* prefix, int nscenes, int64_t total,
                       int kinds, hipStream_t s) {
  if (nscenes <= 0 || total <= 0) return hipSuccess;
  if (!descs || !prefix || total > 0x7fffffff) return hipErrorInvalidValue;
  if (kinds == KINDS_LIQUID)
    hipLaunchKernelGGL((k_g2p2g_multi<KINDS_LIQUID>), dim3((unsigned)total), dim3(64 * kItemWaves), 0, s,
                       descs, prefix, nscenes);
  else
    hipLaunchKernelGGL((k_g2p2g_multi<KINDS_ALL>), dim3((unsigned)total), dim3(64 * kItemWaves), 0, s,
                       descs, prefix, nscenes);
  return hipGetLastError();
}

hipError_t tile_combine_multi(const ReduceDesc* descs, int nscenes, hipStream_t s) {
  if (nscenes <= 0) return hipSuccess;
  if (!descs) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tile_combine_multi, dim3(kCombineGrid), dim3(kCombineThreads), 0, s, descs, nscenes);
  return hipGetLastError();
}

hipError_t tile_reduce_multi(const ReduceDesc* descs, const uint32_t* prefix, int nscenes,
                             int64_t total_blocks, hipStream_t s) {
  if (nscenes <= 0 || total_blocks <= 0) return hipSuccess;
  if (!descs || !prefix || total_blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tile_reduce_multi, dim3((unsigned)((total_blocks + 3) / 4)), dim3(256), 0, s, descs,
                     prefix, nscenes);
  return hipGetLastError();
}

// F = cbrt(Jp) I for every particle: the F planes of a liquid state whose F
// the fused kernel derived from J
__global__ void __launch_bounds__(256) k_liquid_f(ParticlesDev P) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const float cj = cbrtf(P.f[PL_JP * P.stride + i]);
#pragma unroll
  for (int a = 0; a < 9; ++a) P.f[(PL_F + a) * P.stride + i] = (a % 4 == 0) ? cj : 0.0f;
}

hipError_t liquid_f(const ParticlesDev& P, hipStream_t s) {
  if (P.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_liquid_f, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, s, P);
  return hipGetLastError();
}

// node block of each node: zero where the last grid update skipped it
__global__ void __launch_bounds__(256) k_zero_dead(float4* __restrict__ grid, const uint32_t* __restrict__ live, int N) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nodes = (size_t)N * N * N;
  if (t >= nodes) return;
  const uint32_t u = (uint32_t)t, n = (uint32_t)N;  // (N <= 1024: a node index is below 2^30)
  const int z = (int)(u % n), y = (int)((u / n) % n), x = (int)(u / (n * n));
  const int nbx = (N + 3) / 4;
  if (!live[((size_t)(x >> 2) * nbx + (y >> 2)) * nbx + (z >> 2)]) grid[t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

hipError_t zero_dead_blocks(float4* grid, const uint32_t* live, int N, hipStream_t s) {
  const size_t nodes = (size_t)N * N * N;
  hipLaunchKernelGGL(k_zero_dead, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, s, grid, live, N);
  return hipGetLastError();
}

hipError_t tile_combine(const TileSums& in, int64_t nblk, hipStream_t s) {
  if (!in.tiles || !in.block_items || !in.crowded || nblk < 1) return hipErrorInvalidValue;
  ReduceDesc d{};
  d.tiles = in.tiles;
  d.block_items = in.block_items;
  d.crowded = in.crowded;
  d.nblk = nblk;
  const unsigned grid = (unsigned)std::min<int64_t>(kCombineGrid, nblk);
  hipLaunchKernelGGL(k_tile_combine, dim3(grid), dim3(kCombineThreads), 0, s, d);
  return hipGetLastError();
}

hipError_t tile_reduce(const TileSums& in, const StepConsts& k, float4* grid, int update, int mlp,
                       int combined, hipStream_t s) {
  if (!in.tiles || !in.atomic || !in.dirty || !in.block_items || !in.covered || !grid)
    return hipErrorInvalidValue;
  ReduceDesc d;
  d.tiles = in.tiles;
  d.block_items = in.block_items;
  d.covered = in.covered;
  d.grid_atomic = in.atomic;
  d.dirty = in.dirty;
  d.grid = grid;
  d.k = k;
  d.nbx = (k.N + 3) / 4;
  d.update = update;
  d.combined = combined;
  d.nblk = 0;
  d.crowded = nullptr;
  d.live = in.live;
  d.skip_dead = update && in.live;
  const int64_t nblocks = (int64_t)d.nbx * d.nbx * d.nbx;
  if (mlp)
    hipLaunchKernelGGL(k_tile_reduce<true>, dim3((unsigned)((nblocks + 3) / 4)), dim3(256), 0, s, d);
  else
    hipLaunchKernelGGL(k_tile_reduce<false>, dim3((unsigned)((nblocks + 3) / 4)), dim3(256), 0, s, d);
  return hipGetLastError();
}

}  // namespace mpmk
