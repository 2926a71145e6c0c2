// sort.hip -- cell-key sort of particles for gfx950.
//
// The reference processes particles in their storage order (mpm.cu:17,112).
// The engine keeps them ordered by grid cell so that a workgroup's particles
// cover a compact block of the grid (the LDS tiles of the fused kernel). The
// order is an implementation detail: downloads restore the reference's
// particle order.
//
// Key: blocked Morton order. Cells are grouped in BE^3 blocks (BE = 4 or 8);
// the key is morton(block) << 3 log2(BE) | (cell within block). Particles
// outside the grid are clamped to the nearest cell. After the sort the
// particles of each block are contiguous; build_items cuts them into work
// items of at most `chunk` particles of one block, the unit of work of the
// block-tiled substep kernel.
//
// Sort: least-significant-digit radix sort of (key, index) pairs, stable, so
// the permutation equals a CPU stable sort by key. Per digit pass:
//   k_radix_hist     per-tile digit histogram (LDS), digit-major in HBM
//   k_radix_scan     one workgroup per digit: exclusive scan over tiles
//   k_radix_scatter  stable in-tile ranking (wavefront ballots for equal
//                    digits, per-wave counts in LDS), then scatter
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "engine_types.hpp"
#include "kernels.hpp"

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kTile = kSortThreads * kSortItems;
constexpr int kWaves = kSortThreads / 64;

// Key of the cell containing x: blocks of BE^3 cells (BE = 1 << lb) in Morton
// order, cells inside a block in row-major order. Cells outside the grid are
// clamped to the nearest cell.
__device__ __forceinline__ uint32_t cell_key_of(const float x[3], const StepConsts& k, int lb) {
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int b = (int)(x[a] * k.dx_inv - 0.5f);
    b = b < 0 ? 0 : (b >= k.N ? k.N - 1 : b);
    c[a] = b;
  }
  const uint32_t m = mpmdev::morton3((uint32_t)c[0] >> lb, (uint32_t)c[1] >> lb, (uint32_t)c[2] >> lb);
  const uint32_t mask = (1u << lb) - 1u;
  const uint32_t local = ((((uint32_t)c[0] & mask) << lb | ((uint32_t)c[1] & mask)) << lb) |
                         ((uint32_t)c[2] & mask);
  return (m << (3 * lb)) | local;
}

__device__ __forceinline__ void cell_keys_body(const ParticlesDev& P, const StepConsts& k, int lb,
                                               uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                               uint32_t blk) {
  const int i = (int)(blk * 256u + threadIdx.x);
  if (i >= P.n) return;
  float x[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) x[a] = P.f[(PL_X + a) * P.stride + i];
  keys[i] = cell_key_of(x, k, lb);
  vals[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(256)
    k_cell_keys(ParticlesDev P, StepConsts k, int lb, uint32_t* __restrict__ keys,
                uint32_t* __restrict__ vals) {
  cell_keys_body(P, k, lb, keys, vals, blockIdx.x);
}

// ---- work items: runs of at most `chunk` particles of one home block -------
// block_range[2b], block_range[2b+1] = first and one-past-last sorted index of
// block b (written for occupied blocks only)
__device__ __forceinline__ void block_heads_body(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                 uint32_t* __restrict__ block_range, uint32_t blk) {
  const int64_t i = (int64_t)blk * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = keys[i] >> shift;
  if (i == 0 || (keys[i - 1] >> shift) != b) block_range[2 * (size_t)b] = (uint32_t)i;
  if (i == n - 1 || (keys[i + 1] >> shift) != b) block_range[2 * (size_t)b + 1] = (uint32_t)(i + 1);
}

__global__ void __launch_bounds__(256)
    k_block_heads(const uint32_t* __restrict__ keys, int64_t n, int shift,
                  uint32_t* __restrict__ block_range) {
  block_heads_body(keys, n, shift, block_range, blockIdx.x);
}

// flag[i] = 1 where a work item starts. A block of c particles is cut into
// ceil(c / chunk) items of near-equal size (at most chunk each).
__device__ __forceinline__ void item_flags_body(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                int chunk, const uint32_t* __restrict__ block_range,
                                                uint32_t* __restrict__ flag, uint32_t blk) {
  const int64_t i = (int64_t)blk * 256 + threadIdx.x;
  if (i >= n) return;
  const size_t b = keys[i] >> shift;
  const uint32_t s = block_range[2 * b], c = block_range[2 * b + 1] - s;
  const uint32_t items = (c + (uint32_t)chunk - 1u) / (uint32_t)chunk;
  const uint32_t size = (c + items - 1u) / items;
  flag[i] = (((uint32_t)i - s) % size == 0u) ? 1u : 0u;
}

__global__ void __launch_bounds__(256)
    k_item_flags(const uint32_t* __restrict__ keys, int64_t n, int shift, int chunk,
                 const uint32_t* __restrict__ block_range, uint32_t* __restrict__ flag) {
  item_flags_body(keys, n, shift, chunk, block_range, flag, blockIdx.x);
}

// Items in sorted order: item_start/item_block for every flagged particle, and
// for every occupied block its items [block_items[2b], block_items[2b+1]).
__device__ __forceinline__ void item_emit_body(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                               const uint32_t* __restrict__ flag,
                                               const uint32_t* __restrict__ pos,
                                               const uint32_t* __restrict__ total,
                                               uint32_t* __restrict__ item_start,
                                               uint32_t* __restrict__ item_block,
                                               uint32_t* __restrict__ block_items,
                                               uint32_t* __restrict__ covered, int nbx, uint32_t blk) {
  const int64_t i = (int64_t)blk * 256 + threadIdx.x;
  if (i < n) {
    const uint32_t b = keys[i] >> shift;
    if (flag[i]) {
      item_start[pos[i]] = (uint32_t)i;
      item_block[pos[i]] = b;
    }
    if (i == 0 || (keys[i - 1] >> shift) != b) {
      block_items[2 * (size_t)b] = pos[i];
      // the node blocks this home block's tiles can reach
      const int hx = (int)mpmdev::compact3(b >> 2), hy = (int)mpmdev::compact3(b >> 1),
                hz = (int)mpmdev::compact3(b);
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            const int x = hx + dx, y = hy + dy, z = hz + dz;
            if (x >= 0 && x < nbx && y >= 0 && y < nbx && z >= 0 && z < nbx)
              covered[((size_t)x * nbx + y) * nbx + z] = 1u;
          }
    }
    if (i == n - 1 || (keys[i + 1] >> shift) != b) block_items[2 * (size_t)b + 1] = pos[i] + flag[i];
  }
  if (i == 0) item_start[*total] = (uint32_t)n;  // sentinel: end of the last item
}

__global__ void __launch_bounds__(256)
    k_item_emit(const uint32_t* __restrict__ keys, int64_t n, int shift,
                const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                const uint32_t* __restrict__ total, uint32_t* __restrict__ item_start,
                uint32_t* __restrict__ item_block, uint32_t* __restrict__ block_items,
                uint32_t* __restrict__ covered, int nbx) {
  item_emit_body(keys, n, shift, flag, pos, total, item_start, item_block, block_items, covered, nbx,
                 blockIdx.x);
}

// the home blocks with several items: crowded[0] counts them (zero
// beforehand), crowded[1..] lists them (in no particular order)
__device__ __forceinline__ void crowded_body(const uint32_t* __restrict__ block_items, int64_t nblk,
                                             uint32_t* __restrict__ crowded, uint32_t blk) {
  const int64_t b = (int64_t)blk * 256 + threadIdx.x;
  if (b >= nblk) return;
  const uint2 r = reinterpret_cast<const uint2*>(block_items)[b];
  if (r.y > r.x + 1u) crowded[1 + atomicAdd(crowded, 1u)] = (uint32_t)b;
}

__global__ void __launch_bounds__(256)
    k_crowded(const uint32_t* __restrict__ block_items, int64_t nblk, uint32_t* __restrict__ crowded) {
  crowded_body(block_items, nblk, crowded, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Arrangement of one item's n particles over the first nw = ceil(n / 64) of
// the kItemWaves wavefronts of the substep kernel, which gives every wavefront a
// private P2G tile: the particle of rank r within its cell goes to wavefront
// (r + cell) mod nw, so particles sharing a cell land in different wavefronts
// as long as a cell holds at most nw of the item's particles. A wavefront
// takes at most 64 particles; the excess fills the free lanes of the other
// wavefronts in order (the first nw hold enough). gather[j] receives the
// original index perm[i] of the sorted particle i that takes slot j (a
// permutation inside the item); item_waves packs the per-wavefront counts
// (8 bits each). The layout is a function of the sorted
// keys only. One workgroup of 256 threads per item (n <= kMaxItemParticles).
__device__ __forceinline__ void item_arrange_body(const uint32_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ perm, int cell_bits,
                                                  const uint32_t* __restrict__ item_start,
                                                  const uint32_t* __restrict__ count,
                                                  uint32_t* __restrict__ gather,
                                                  uint32_t* __restrict__ item_waves, const uint32_t item) {
  if (item >= *count) return;  // uniform for the workgroup
  __shared__ uint32_t sk[256];
  __shared__ uint8_t orig[256];
  __shared__ uint32_t wcnt[4][64];
  __shared__ uint32_t cstart[64];
  __shared__ unsigned long long wsum[4];
  const int lane = threadIdx.x & 63;
  const uint32_t s = item_start[item];
  const int n = (int)(item_start[item + 1] - s);  // 1..kMaxItemParticles
  const uint32_t cmask = (1u << cell_bits) - 1u;  // cell within the home block (<= 64 cells)
  // -- stable counting sort of the item by cell (the keys arrive sorted by
  //    home block only): rank among equal cells from wavefront ballots, then
  //    per-wavefront counts and an exclusive scan over the cells
  {
    const int t0 = threadIdx.x, wv0 = t0 >> 6;
    for (int i = t0; i < 4 * 64; i += 256) (&wcnt[0][0])[i] = 0u;
    __syncthreads();
    const bool v0 = t0 < n;
    const uint32_t key0 = v0 ? keys[s + t0] : 0u;
    const uint32_t c0 = key0 & cmask;
    uint64_t peers = __ballot(v0);
    for (int b = 0; b < cell_bits; ++b) {
      const uint64_t m = __ballot((c0 >> b) & 1u);
      peers &= ((c0 >> b) & 1u) ? m : ~m;
    }
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t rank_w = (uint32_t)__popcll(peers & lt);
    if (v0 && (peers & lt) == 0ull) wcnt[wv0][c0] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (t0 < 64) {
      uint32_t acc = 0;
      for (int q = 0; q < 4; ++q) {
        const uint32_t v = wcnt[q][t0];
        wcnt[q][t0] = acc;
        acc += v;
      }
      // exclusive scan of the cell totals over the 64 lanes of wavefront 0
      uint32_t incl = acc;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
      }
      cstart[t0] = incl - acc;
    }
    __syncthreads();
    if (v0) {
      const uint32_t pos = cstart[c0] + wcnt[wv0][c0] + rank_w;
      sk[pos] = key0;
      orig[pos] = (uint8_t)t0;
    }
    __syncthreads();
  }
  // -- from here on thread t holds the t-th particle of the item in cell order
  const int t = threadIdx.x, wv = t >> 6;
  const bool valid = t < n;
  const uint32_t key = valid ? sk[t] : 0u;
  const int cell = (int)(key & cmask);
  const int rank = valid ? t - (int)cstart[cell] : 0;  // rank within its cell
  const int nw = (n + 63) >> 6;  // wavefronts the item needs
  const int w = (rank + cell) % nw;
  const unsigned long long one = valid ? (1ull << (16 * w)) : 0ull;
  unsigned long long incl = one;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long y = __shfl_up(incl, off);
    if (lane >= off) incl += y;
  }
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  unsigned long long before = 0ull, total = 0ull;
  for (int q = 0; q < 4; ++q) {  // the 4 wavefronts of this workgroup
    if (q < wv) before += wsum[q];
    total += wsum[q];
  }
  const int idx = (int)(((before + incl - one) >> (16 * w)) & 0xffffull);
  constexpr int NW = mpmk::kItemWaves;
  int cnt[NW], c[NW], fin[NW];
  int over = 0;
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    cnt[q] = (int)((total >> (16 * q)) & 0xffffull);
    c[q] = min(cnt[q], 64);
    over += cnt[q] - c[q];
  }
  // free lanes take the excess in wavefront order
  int left = over;
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    const int take = min(64 - c[q], left);
    fin[q] = c[q] + take;
    left -= take;
  }
  int fw = w, fl = idx;
  if (idx >= 64) {
    // the o-th excess particle takes the o-th free lane, wavefronts in order
    int o = idx - 64;
    for (int q = 0; q < w; ++q) o += cnt[q] - c[q];
    for (int q = 0; q < NW; ++q) {
      const int fr = 64 - c[q];
      if (o < fr) { fw = q; fl = c[q] + o; break; }
      o -= fr;
    }
  }
  int off = 0;
  for (int q = 0; q < fw; ++q) off += fin[q];
  if (valid) gather[s + (uint32_t)(off + fl)] = perm[s + orig[t]];
  if (t == 0) {
    uint32_t packed = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) packed |= (uint32_t)fin[q] << (8 * q);
    item_waves[item] = packed;
  }
}

__global__ void __launch_bounds__(256)
    k_item_arrange(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ perm,
                   int cell_bits, const uint32_t* __restrict__ item_start,
                   const uint32_t* __restrict__ count, uint32_t* __restrict__ gather,
                   uint32_t* __restrict__ item_waves) {
  item_arrange_body(keys, perm, cell_bits, item_start, count, gather, item_waves, blockIdx.x);
}

// ---- exclusive scan of u32 (tiles of kTile, then tile sums in one workgroup)
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  // 256 threads; returns the exclusive prefix of v over the workgroup
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t add = (tid >= off) ? sh[tid - off] : 0u;
    __syncthreads();
    sh[tid] += add;
    __syncthreads();
  }
  const uint32_t incl = sh[tid];
  if (total) *total = sh[255];
  __syncthreads();
  return incl - v;
}

__device__ __forceinline__ void scan_tiles_body(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                int64_t n, uint32_t* __restrict__ tile_sums, uint32_t blk) {
  __shared__ uint32_t sh[256];
  __shared__ uint32_t tot;
  const int64_t base = (int64_t)blk * kTile + (int64_t)threadIdx.x * kSortItems;
  uint32_t v[kSortItems], s = 0;
#pragma unroll
  for (int t = 0; t < kSortItems; ++t) {
    v[t] = (base + t < n) ? in[base + t] : 0u;
    s += v[t];
  }
  const uint32_t pre = block_exclusive_scan(s, sh, threadIdx.x == 255 ? &tot : nullptr);
  uint32_t run = pre;
#pragma unroll
  for (int t = 0; t < kSortItems; ++t) {
    if (base + t < n) out[base + t] = run;
    run += v[t];
  }
  __syncthreads();
  if (threadIdx.x == 0) tile_sums[blk] = tot;
}

__global__ void __launch_bounds__(256)
    k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int64_t n,
                 uint32_t* __restrict__ tile_sums) {
  scan_tiles_body(in, out, n, tile_sums, blockIdx.x);
}

__device__ __forceinline__ void scan_sums_body(uint32_t* __restrict__ sums, int ntiles,
                                               uint32_t* __restrict__ total) {
  __shared__ uint32_t sh[256];
  __shared__ uint32_t tot;
  uint32_t carry = 0;
  for (int c0 = 0; c0 < ntiles; c0 += 256) {
    const int i = c0 + threadIdx.x;
    const uint32_t v = (i < ntiles) ? sums[i] : 0u;
    const uint32_t pre = block_exclusive_scan(v, sh, threadIdx.x == 255 ? &tot : nullptr);
    __syncthreads();
    if (i < ntiles) sums[i] = carry + pre;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(256)
    k_scan_sums(uint32_t* __restrict__ sums, int ntiles, uint32_t* __restrict__ total) {
  scan_sums_body(sums, ntiles, total);
}

__device__ __forceinline__ void scan_add_body(uint32_t* __restrict__ out, int64_t n,
                                              const uint32_t* __restrict__ sums, uint32_t blk) {
  const int64_t base = (int64_t)blk * kTile + (int64_t)threadIdx.x * kSortItems;
  const uint32_t add = sums[blk];
#pragma unroll
  for (int t = 0; t < kSortItems; ++t)
    if (base + t < n) out[base + t] += add;
}

__global__ void __launch_bounds__(256)
    k_scan_add(uint32_t* __restrict__ out, int64_t n, const uint32_t* __restrict__ sums) {
  scan_add_body(out, n, sums, blockIdx.x);
}

template <int BITS>
__device__ __forceinline__ void radix_hist_body(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                uint32_t* __restrict__ hist, int ntiles, uint32_t blk) {
  constexpr int R = 1 << BITS;
  __shared__ uint32_t cnt[R];
  for (int d = threadIdx.x; d < R; d += kSortThreads) cnt[d] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blk * kTile;
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int64_t i = base + it * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (R - 1)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += kSortThreads) hist[(size_t)d * ntiles + blk] = cnt[d];
}

template <int BITS>
__global__ void __launch_bounds__(kSortThreads)
    k_radix_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                 uint32_t* __restrict__ hist, int ntiles) {
  radix_hist_body<BITS>(keys, n, shift, hist, ntiles, blockIdx.x);
}

// One workgroup per digit: exclusive scan of hist[d][0..ntiles) in place;
// total[d] receives the digit's count.
__device__ __forceinline__ void radix_scan_body(uint32_t* __restrict__ hist, int ntiles,
                                                uint32_t* __restrict__ total, uint32_t digit) {
  __shared__ uint32_t part[256];
  uint32_t* row = hist + (size_t)digit * ntiles;
  uint32_t carry = 0;
  for (int c0 = 0; c0 < ntiles; c0 += 256) {
    const int i = c0 + threadIdx.x;
    const uint32_t v = (i < ntiles) ? row[i] : 0u;
    part[threadIdx.x] = v;
    __syncthreads();
    // Hillis-Steele inclusive scan over 256 entries
    for (int off = 1; off < 256; off <<= 1) {
      const uint32_t add = (threadIdx.x >= (unsigned)off) ? part[threadIdx.x - off] : 0u;
      __syncthreads();
      part[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < ntiles) row[i] = carry + part[threadIdx.x] - v;
    carry += part[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) total[digit] = carry;
}

__global__ void __launch_bounds__(256)
    k_radix_scan(uint32_t* __restrict__ hist, int ntiles, uint32_t* __restrict__ total) {
  radix_scan_body(hist, ntiles, total, blockIdx.x);
}

template <int BITS>
__device__ __forceinline__ void radix_scatter_body(const uint32_t* __restrict__ keys_in,
                                                   const uint32_t* __restrict__ vals_in,
                                                   uint32_t* __restrict__ keys_out,
                                                   uint32_t* __restrict__ vals_out, int64_t n, int shift,
                                                   const uint32_t* __restrict__ hist, int ntiles,
                                                   const uint32_t* __restrict__ total, uint32_t blk) {
  constexpr int R = 1 << BITS;
  __shared__ uint32_t dbase[R];             // global start of each digit for this tile
  __shared__ uint32_t wcnt[kWaves][R];      // per-wave digit counts -> offsets
  __shared__ uint32_t run[R];               // digits already placed by earlier items
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  // digit bases: exclusive scan of the digit totals (R <= 256 entries), done
  // redundantly by every workgroup
  if (tid < R) run[tid] = total[tid];
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (int d = 0; d < R; ++d) {
      const uint32_t t = run[d];
      run[d] = acc;
      acc += t;
    }
  }
  __syncthreads();
  for (int d = tid; d < R; d += kSortThreads) {
    dbase[d] = run[d] + hist[(size_t)d * ntiles + blk];
  }
  __syncthreads();
  for (int d = tid; d < R; d += kSortThreads) run[d] = 0;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t base = (int64_t)blk * kTile;
  for (int it = 0; it < kSortItems; ++it) {
    for (int d = tid; d < kWaves * R; d += kSortThreads) (&wcnt[0][0])[d] = 0;
    __syncthreads();
    const int64_t i = base + it * kSortThreads + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? keys_in[i] : 0u;
    const uint32_t val = valid ? vals_in[i] : 0u;
    const uint32_t dig = (key >> shift) & (R - 1);
    // lanes of this wavefront holding the same digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
      const uint64_t m = __ballot((dig >> b) & 1u);
      peers &= ((dig >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    const bool leader = valid && (peers & lt_mask) == 0ull;
    if (leader) wcnt[wave][dig] = (uint32_t)__popcll(peers);
    __syncthreads();
    // per digit: exclusive offsets over waves, continuing the running count
    for (int d = tid; d < R; d += kSortThreads) {
      uint32_t r = run[d];
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = wcnt[w][d];
        wcnt[w][d] = r;
        r += c;
      }
      run[d] = r;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = dbase[dig] + wcnt[wave][dig] + rank;
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
  }
}

template <int BITS>
__global__ void __launch_bounds__(kSortThreads)
    k_radix_scatter(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
                    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, int64_t n,
                    int shift, const uint32_t* __restrict__ hist, int ntiles,
                    const uint32_t* __restrict__ total) {
  radix_scatter_body<BITS>(keys_in, vals_in, keys_out, vals_out, n, shift, hist, ntiles, total, blockIdx.x);
}

// dst[i] = src[perm[i]] for every plane, material and id
__global__ void __launch_bounds__(256)
    k_reorder(ParticlesDev src, ParticlesDev dst, const uint32_t* __restrict__ perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= src.n) return;
  const uint32_t j = perm[i];
#pragma unroll
  for (int p = 0; p < NUM_PLANES; ++p) dst.f[p * dst.stride + i] = src.f[p * src.stride + j];
  dst.mat[i] = src.mat[j];
  dst.id[i] = src.id[j];
}

// Removal of the particles with original index in [lo, hi): keep flags, then
// the kept ones packed in their order into dst (indices above the range move
// down by hi - lo) and the removed ones into `gone` (planes of hi - lo
// floats, slot id - lo) with their material indices.
__global__ void __launch_bounds__(256)
    k_keep_flags(ParticlesDev P, uint32_t lo, uint32_t hi, uint32_t* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const uint32_t id = P.id[i];
  flag[i] = (id >= lo && id < hi) ? 0u : 1u;
}

__global__ void __launch_bounds__(256)
    k_compact(ParticlesDev src, ParticlesDev dst, const uint32_t* __restrict__ pos, uint32_t lo, uint32_t hi,
              float* __restrict__ gone, uint8_t* __restrict__ gone_mat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= src.n) return;
  const uint32_t id = src.id[i];
  if (id >= lo && id < hi) {
    if (!gone) return;
    const uint32_t j = id - lo, m = hi - lo;
#pragma unroll
    for (int p = 0; p < NUM_PLANES; ++p) gone[(size_t)p * m + j] = src.f[p * src.stride + i];
    gone_mat[j] = src.mat[i];
    return;
  }
  const uint32_t j = pos[i];
#pragma unroll
  for (int p = 0; p < NUM_PLANES; ++p) dst.f[p * dst.stride + j] = src.f[p * src.stride + i];
  dst.mat[j] = src.mat[i];
  dst.id[j] = id >= hi ? id - (hi - lo) : id;
}

template <int BITS>
hipError_t radix_sort_t(uint32_t* keys[2], uint32_t* vals[2], int64_t n, int start_bit, int key_bits,
                        uint32_t* scratch, int* out_idx, hipStream_t s) {
  constexpr int R = 1 << BITS;
  const int ntiles = (int)((n + kTile - 1) / kTile);
  uint32_t* hist = scratch;
  uint32_t* total = scratch + (size_t)R * ntiles;
  int cur = 0;
  for (int shift = start_bit; shift < key_bits; shift += BITS) {
    hipLaunchKernelGGL(k_radix_hist<BITS>, dim3(ntiles), dim3(kSortThreads), 0, s, keys[cur], n,
                       shift, hist, ntiles);
    hipLaunchKernelGGL(k_radix_scan, dim3(R), dim3(256), 0, s, hist, ntiles, total);
    hipLaunchKernelGGL(k_radix_scatter<BITS>, dim3(ntiles), dim3(kSortThreads), 0, s, keys[cur],
                       vals[cur], keys[cur ^ 1], vals[cur ^ 1], n, shift, hist, ntiles, total);
    cur ^= 1;
  }
  *out_idx = cur;
  return hipGetLastError();
}

// ---- many scenes per launch: workgroup g of the launch is workgroup
// g - prefix[j] of scene j (prefix[j] <= g < prefix[j + 1]; scene_of_wave)
template <int PH>
__global__ void __launch_bounds__(256)
    k_sort_multi(const mpmk::SortDesc* __restrict__ descs, const uint32_t* __restrict__ prefix, int nscenes,
                 int pass) {
  const uint32_t g = blockIdx.x;
  const int j = mpmdev::scene_of_wave(prefix, nscenes, g);
  const mpmk::SortDesc& d = descs[j];
  const uint32_t blk = g - prefix[j];
  const int out = d.npasses & 1;
  const int shift = 3 * d.lb;
  if constexpr (PH == mpmk::SORT_FILL) {
    const int64_t i = (int64_t)blk * 256 + threadIdx.x;
    const int64_t nnb = (int64_t)d.nbx * d.nbx * d.nbx;
    if (i == 0) d.crowded[0] = 0u;
    if (i < 2 * d.nblk) d.block_items[i] = 0u;
    else if (i < 2 * d.nblk + nnb) d.covered[i - 2 * d.nblk] = 0u;
  } else if constexpr (PH == mpmk::SORT_KEYS) {
    cell_keys_body(d.P, d.k, d.lb, d.keys[0], d.vals[0], blk);
  } else if constexpr (PH == mpmk::SORT_HIST) {
    const int ntiles = (int)((d.n + kTile - 1) / kTile);
    radix_hist_body<8>(d.keys[pass & 1], d.n, d.start_bit + 8 * pass, d.scratch, ntiles, blk);
  } else if constexpr (PH == mpmk::SORT_RSCAN) {
    const int ntiles = (int)((d.n + kTile - 1) / kTile);
    radix_scan_body(d.scratch, ntiles, d.scratch + (size_t)256 * ntiles, blk);
  } else if constexpr (PH == mpmk::SORT_SCATTER) {
    const int ntiles = (int)((d.n + kTile - 1) / kTile);
    radix_scatter_body<8>(d.keys[pass & 1], d.vals[pass & 1], d.keys[(pass + 1) & 1], d.vals[(pass + 1) & 1],
                          d.n, d.start_bit + 8 * pass, d.scratch, ntiles, d.scratch + (size_t)256 * ntiles,
                          blk);
  } else if constexpr (PH == mpmk::SORT_HEADS) {
    block_heads_body(d.keys[out], d.n, shift, d.block_range, blk);
  } else if constexpr (PH == mpmk::SORT_FLAGS) {
    item_flags_body(d.keys[out], d.n, shift, d.chunk, d.block_range, d.flag, blk);
  } else if constexpr (PH == mpmk::SORT_SCAN_TILES) {
    scan_tiles_body(d.flag, d.pos, d.n, d.scratch, blk);
  } else if constexpr (PH == mpmk::SORT_SCAN_SUMS) {
    scan_sums_body(d.scratch, (int)((d.n + kTile - 1) / kTile), d.item_count);
  } else if constexpr (PH == mpmk::SORT_SCAN_ADD) {
    scan_add_body(d.pos, d.n, d.scratch, blk);
  } else if constexpr (PH == mpmk::SORT_EMIT) {
    item_emit_body(d.keys[out], d.n, shift, d.flag, d.pos, d.item_count, d.item_start, d.item_block,
                   d.block_items, d.covered, d.nbx, blk);
  } else if constexpr (PH == mpmk::SORT_ARRANGE) {
    item_arrange_body(d.keys[out], d.vals[out], shift, d.item_start, d.item_count, d.gather,
                      d.item_waves, blk);
  } else {
    crowded_body(d.block_items, d.nblk, d.crowded, blk);
  }
}

}  // namespace

namespace mpmk {

int64_t sort_blocks(const SortDesc& d, int phase, int pass) {
  const int64_t per256 = (d.n + 255) / 256, tiles = (d.n + kTile - 1) / kTile;
  switch (phase) {
    case SORT_FILL: return (2 * d.nblk + (int64_t)d.nbx * d.nbx * d.nbx + 255) / 256;
    case SORT_KEYS: case SORT_HEADS: case SORT_FLAGS: case SORT_EMIT: return per256;
    case SORT_HIST: case SORT_SCATTER: return pass < d.npasses ? tiles : 0;
    case SORT_RSCAN: return pass < d.npasses ? 256 : 0;
    case SORT_SCAN_TILES: case SORT_SCAN_ADD: return tiles;
    case SORT_SCAN_SUMS: return 1;
    case SORT_ARRANGE: return d.max_items;
    case SORT_CROWDED: return (d.nblk + 255) / 256;
    default: return 0;
  }
}

hipError_t sort_multi(int phase, int pass, const SortDesc* descs, const uint32_t* prefix, int nscenes,
                      int64_t total_blocks, hipStream_t s) {
  if (nscenes <= 0 || total_blocks <= 0) return hipSuccess;
  if (!descs || !prefix || total_blocks > 0x7fffffff) return hipErrorInvalidValue;
  const dim3 g((unsigned)total_blocks), b(256);
  switch (phase) {
#define MPM_SORT_CASE(PH) \
    case PH: hipLaunchKernelGGL(k_sort_multi<PH>, g, b, 0, s, descs, prefix, nscenes, pass); break;
    MPM_SORT_CASE(SORT_FILL)
    MPM_SORT_CASE(SORT_KEYS)
    MPM_SORT_CASE(SORT_HIST)
    MPM_SORT_CASE(SORT_RSCAN)
    MPM_SORT_CASE(SORT_SCATTER)
    MPM_SORT_CASE(SORT_HEADS)
    MPM_SORT_CASE(SORT_FLAGS)
    MPM_SORT_CASE(SORT_SCAN_TILES)
    MPM_SORT_CASE(SORT_SCAN_SUMS)
    MPM_SORT_CASE(SORT_SCAN_ADD)
    MPM_SORT_CASE(SORT_EMIT)
    MPM_SORT_CASE(SORT_ARRANGE)
    MPM_SORT_CASE(SORT_CROWDED)
#undef MPM_SORT_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int cell_key_bits(int N, int lb) {
  const int nb = (N + (1 << lb) - 1) >> lb;
  int b = 0;
  while ((1 << b) < nb) ++b;
  return 3 * b + 3 * lb;
}

size_t radix_scratch_words(int64_t n, int radix_bits) {
  const int64_t ntiles = (n + kTile - 1) / kTile;
  return (size_t)(1u << radix_bits) * (size_t)(ntiles > 0 ? ntiles : 1) + 256;
}

hipError_t cell_keys(const ParticlesDev& P, const StepConsts& k, int lb, uint32_t* keys,
                     uint32_t* vals, hipStream_t s) {
  if (P.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_keys, dim3((P.n + 255) / 256), dim3(256), 0, s, P, k, lb, keys, vals);
  return hipGetLastError();
}

hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* scratch,
                          uint32_t* total, hipStream_t s) {
  const int ntiles = (int)((n + kTile - 1) / kTile);
  if (n <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(256), 0, s, in, out, n, scratch);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, scratch, ntiles, total);
  hipLaunchKernelGGL(k_scan_add, dim3(ntiles), dim3(256), 0, s, out, n, scratch);
  return hipGetLastError();
}

hipError_t fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t build_items(const uint32_t* keys, const uint32_t* perm, int64_t n, int lb, int chunk,
                       int64_t nblk, uint32_t* block_range, uint32_t* block_items, uint32_t* covered,
                       int nbx, uint32_t* flag, uint32_t* pos, uint32_t* scratch, uint32_t* total,
                       uint32_t* item_start, uint32_t* item_block, uint32_t* item_waves,
                       uint32_t* gather, int64_t max_items, uint32_t* crowded, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (!crowded) return hipErrorInvalidValue;
  if (chunk < 1 || chunk > kMaxItemParticles || max_items < 1) return hipErrorInvalidValue;
  const int shift = 3 * lb;
  const unsigned g = (unsigned)((n + 255) / 256);
  const int64_t nnb = (int64_t)nbx * nbx * nbx;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((2 * nblk + 255) / 256)), dim3(256), 0, s,
                     block_items, 2 * nblk, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nnb + 255) / 256)), dim3(256), 0, s, covered, nnb, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, crowded, (int64_t)1, 0u);
  hipLaunchKernelGGL(k_block_heads, dim3(g), dim3(256), 0, s, keys, n, shift, block_range);
  hipLaunchKernelGGL(k_item_flags, dim3(g), dim3(256), 0, s, keys, n, shift, chunk, block_range, flag);
  hipError_t e = exclusive_scan(flag, pos, n, scratch, total, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_item_emit, dim3(g), dim3(256), 0, s, keys, n, shift, flag, pos, total,
                     item_start, item_block, block_items, covered, nbx);
  if (!perm || !gather || !item_waves) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_item_arrange, dim3((unsigned)max_items), dim3(256), 0, s, keys, perm, shift,
                     item_start, total, gather, item_waves);
  hipLaunchKernelGGL(k_crowded, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, s, block_items, nblk, crowded);
  return hipGetLastError();
}

hipError_t radix_sort(uint32_t* keys[2], uint32_t* vals[2], int64_t n, int start_bit, int key_bits,
                      int radix_bits, uint32_t* scratch, int* out_idx, hipStream_t s) {
  *out_idx = 0;
  if (n <= 1 || key_bits <= start_bit) return hipSuccess;
  switch (radix_bits) {
    case 4: return radix_sort_t<4>(keys, vals, n, start_bit, key_bits, scratch, out_idx, s);
    case 6: return radix_sort_t<6>(keys, vals, n, start_bit, key_bits, scratch, out_idx, s);
    case 8: return radix_sort_t<8>(keys, vals, n, start_bit, key_bits, scratch, out_idx, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t remove_range(const ParticlesDev& src, const ParticlesDev& dst, uint32_t lo, uint32_t hi,
                        uint32_t* flag, uint32_t* pos, uint32_t* scratch, uint32_t* total, float* gone,
                        uint8_t* gone_mat, hipStream_t s) {
  if (src.n <= 0) return hipSuccess;
  const unsigned g = (unsigned)((src.n + 255) / 256);
  hipLaunchKernelGGL(k_keep_flags, dim3(g), dim3(256), 0, s, src, lo, hi, flag);
  hipError_t e = exclusive_scan(flag, pos, src.n, scratch, total, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_compact, dim3(g), dim3(256), 0, s, src, dst, pos, lo, hi, gone, gone_mat);
  return hipGetLastError();
}

hipError_t reorder(const ParticlesDev& src, const ParticlesDev& dst, const uint32_t* perm,
                   hipStream_t s) {
  if (src.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_reorder, dim3((src.n + 255) / 256), dim3(256), 0, s, src, dst, perm);
  return hipGetLastError();
}

}  // namespace mpmk
