// kernels.hpp -- host-side launchers of the HIP kernels (defined in the .hip
// files of this directory). All launches are asynchronous on the given stream
// and return the launch status.
#pragma once
#include <hip/hip_runtime.h>

#include "engine_types.hpp"

namespace mpmk {

hipError_t fill_zero(float4* g, size_t n, hipStream_t s);
// The direct kernels over the TransferScheme policy of interpolation kernel
// `interp` (MPM_INTERP_QUADRATIC: the reference's MLS-APIC with the quadratic
// B-spline; MPM_INTERP_CUBIC: the cubic B-spline with the generic D^-1).
hipError_t p2g_direct(const ParticlesDev& P, const MaterialDev* mats, const StepConsts& k,
                      float4* grid, hipStream_t s, int interp = MPM_INTERP_QUADRATIC);
hipError_t grid_update(float4* g, float4* clear, const StepConsts& k, hipStream_t s);
hipError_t g2p(const ParticlesDev& P, const MaterialDev* mats, const StepConsts& k,
               const float4* grid, hipStream_t s, int interp = MPM_INTERP_QUADRATIC);
hipError_t g2p2g_direct(const ParticlesDev& P, const MaterialDev* mats, const StepConsts& k,
                        const float4* grid_in, float4* grid_out, hipStream_t s,
                        int interp = MPM_INTERP_QUADRATIC);
hipError_t svd_batch(const float* A, float* U, float* S, float* V, int64_t n, hipStream_t s);
hipError_t probe_kernel_image(int* nregs);

// sort.hip
int cell_key_bits(int N, int lb);
size_t radix_scratch_words(int64_t n, int radix_bits);
hipError_t cell_keys(const ParticlesDev& P, const StepConsts& k, int lb, uint32_t* keys,
                     uint32_t* vals, hipStream_t s);
// Stable LSD radix sort of (keys[0], vals[0]) on key bits [start_bit,
// key_bits); the result lands in keys[*out_idx], vals[*out_idx]. scratch
// holds radix_scratch_words() words.
hipError_t radix_sort(uint32_t* keys[2], uint32_t* vals[2], int64_t n, int start_bit, int key_bits,
                      int radix_bits, uint32_t* scratch, int* out_idx, hipStream_t s);
hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* scratch,
                          uint32_t* total, hipStream_t s);
hipError_t fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
// A work item runs on one workgroup of kItemWaves wavefronts, one particle
// per thread: at most 64 kItemWaves particles.
constexpr int kItemWaves = 3;
constexpr int kMaxItemParticles = 64 * kItemWaves;
// Home blocks of the tiled substep: 4^3 cells; a work item's tile spans the
// block, the 2-node reach of its stencil and a 1-cell drift margin on each
// side: kTileW = 8 nodes per axis.
constexpr int kBlockLog2 = 2;
constexpr int kTileW = 8;
constexpr int kTileNodes = kTileW * kTileW * kTileW;

// Work items over sorted keys: item_start[0..*total] (sentinel n at the end)
// and item_block[0..*total); block_items[2b], [2b+1]: the items of block b
// (cleared for all nblk blocks first). k_item_arrange then lays out each
// item over its wavefronts: item_waves (per-wavefront particle counts) and
// gather[j] = perm[i] for the sorted particle i that lands in slot j.
// Arrays of n+1 entries; block_range holds two entries per grid block
// (Morton space).
// covered (nbx^3 flags, one per 4^3 node block) marks the node blocks some
// item's tile can reach.
hipError_t build_items(const uint32_t* keys, const uint32_t* perm, int64_t n, int lb, int chunk,
                       int64_t nblk, uint32_t* block_range, uint32_t* block_items, uint32_t* covered,
                       int nbx, uint32_t* flag, uint32_t* pos, uint32_t* scratch, uint32_t* total,
                       uint32_t* item_start, uint32_t* item_block, uint32_t* item_waves,
                       uint32_t* gather, int64_t max_items, uint32_t* crowded, hipStream_t s);
// Removes the particles with original index in [lo, hi): the others packed
// into dst in their order, indices above the range lowered by hi - lo; the
// removed ones (when `gone` is non-null) into planes of hi - lo floats at
// slot index - lo, their materials into gone_mat. flag, pos: src.n entries;
// scratch: exclusive_scan's tile sums; total: one word.
hipError_t remove_range(const ParticlesDev& src, const ParticlesDev& dst, uint32_t lo, uint32_t hi,
                        uint32_t* flag, uint32_t* pos, uint32_t* scratch, uint32_t* total, float* gone,
                        uint8_t* gone_mat, hipStream_t s);
// dst[i] = src[perm[i]]
hipError_t reorder(const ParticlesDev& src, const ParticlesDev& dst, const uint32_t* perm,
                   hipStream_t s);

// p2g_block.hip: fused G2P + tiled P2G over work items, and the tile reduction
struct WorkItems {
  const uint32_t* start;   // n_items + 1 entries (last = P)
  const uint32_t* block;   // Morton index of the item's home block
  const uint32_t* waves;   // per-wavefront particle counts, 8 bits each
  const uint32_t* count;   // device scalar: number of items
  int64_t max_items;       // bound on the number of items
  int64_t launch;          // workgroups to launch (one per possible item)
};
// Tiled P2G output of one pass: one kTileNodes-node float4 tile per item,
// plus the grid `atomic` (kept zero between passes) that particles leaving
// their tile add into, with `dirty` flags (one per 4^3 node block) marking
// where they did.
struct TileSums {
  float4* tiles;
  float4* atomic;
  uint32_t* dirty;
  const uint32_t* block_items;  // 2 entries per home block (Morton space)
  const uint32_t* covered;      // per node block: some tile reaches it
  uint32_t* live;               // per node block: written by the last grid update
  const uint32_t* crowded;      // count, then the home blocks with several items
};
// With `gather`, P is the particle order before the last sort: thread slot i
// reads particle gather[i] of P and writes it to slot i of Pdst (the sort's
// reorder folded into this pass). Without it Pdst must be P.
// One scene's fused pass (the arguments of k_g2p2g_tiled) and one scene's
// tile reduction (those of k_tile_reduce): the multi-scene launches read one
// of these per scene from a device table.
struct ItemDesc {
  ParticlesDev P, Pdst;
  const uint32_t* gather;
  const MaterialDev* mats;
  StepConsts k;
  const uint32_t* item_start;
  const uint32_t* item_block;
  const uint32_t* item_waves;
  const uint32_t* item_count;
  const float4* grid_v;
  float4* tiles;
  float* grid_atomic;
  uint32_t* dirty;
  int flags;
};
struct ReduceDesc {
  const float4* tiles;
  const uint32_t* block_items;
  const uint32_t* covered;
  float4* grid_atomic;
  uint32_t* dirty;
  float4* grid;
  StepConsts k;
  int nbx;
  int update;
  int combined;   // the items of each home block were summed into its first tile
  int64_t nblk;   // home blocks (Morton space)
  // crowded[0]: the number of home blocks with several items, crowded[1..]:
  // those blocks, in no particular order (built at the sort)
  const uint32_t* crowded;
  // live[nb] = some tile or side-grid scatter reaches node block nb; with
  // skip_dead the other blocks keep their old values instead of zeros (no
  // particle stencil reads them: a particle's G2P reads the nodes its own
  // P2G reached)
  uint32_t* live;
  int skip_dead;
};
// flags of the fused pass: G2P of the staged grid first (else P2G only);
// F derived from J (liquid kinds only: F planes neither read nor written);
// the original-index guard may bind (some particle id exceeds k.guard);
// XCD-contiguous item order: the hardware hands workgroup b to XCD b % 8, so
// workgroup b runs item (b % 8) * ceil(n / 8) + b / 8 and each XCD's L2 serves
// a contiguous run of items (neighbouring home blocks share staged nodes)
enum TiledFlags : int { TILED_G2P = 1, TILED_DERIVE_F = 2, TILED_GUARD = 4, TILED_XCD = 8 };
// launches of the XCD-ordered kernels round their grid up to a multiple of this
constexpr int kXcds = 8;
// The same passes over several scenes in one launch: descs[j] is scene j,
// prefix[j] its first workgroup (prefix[n] = total; workgroup counts are any
// positive number for g2p2g, whose workgroup takes every count-th item of its
// scene from its index on, and the node-block counts for the reduction).
// kinds: KINDS of every scene of the launch.
hipError_t g2p2g_desc(const ItemDesc& d, int nmats, int64_t launch, int kinds, hipStream_t s);
hipError_t g2p2g_multi(const ItemDesc* descs, const uint32_t* prefix, int nscenes, int64_t total,
                       int kinds, hipStream_t s);
hipError_t tile_reduce_multi(const ReduceDesc* descs, const uint32_t* prefix, int nscenes,
                             int64_t total_blocks, hipStream_t s);
// zero the node blocks that the last grid update left out (before a download)
hipError_t zero_dead_blocks(float4* grid, const uint32_t* live, int N, hipStream_t s);
// sums the tiles of every home block holding several items into its first
// item's tile, in item order (prefix over the scenes' nblk); a reduction with
// `combined` set then reads one tile per home block
hipError_t tile_combine_multi(const ReduceDesc* descs, int nscenes, hipStream_t s);
// One scene's cell sort and work-item build (the arguments of the sort
// kernels), for sorts of many scenes in one launch per kernel.
struct SortDesc {
  ParticlesDev P;
  StepConsts k;
  uint32_t* keys[2];
  uint32_t* vals[2];
  uint32_t* scratch;  // radix histograms and totals; tile sums of the item scan
  int64_t n;
  int lb, start_bit, key_bits, npasses;
  uint32_t* block_range;
  uint32_t* block_items;
  uint32_t* covered;
  int64_t nblk;
  int nbx;
  uint32_t* flag;
  uint32_t* pos;
  uint32_t* item_count;
  uint32_t* item_start;
  uint32_t* item_block;
  uint32_t* item_waves;
  uint32_t* gather;
  int chunk;
  int64_t max_items;
  uint32_t* crowded;  // the home blocks with several items: count, then the list
};
// the sort as a sequence of phases; SORT_HIST/RSCAN/SCATTER run once per
// 8-bit pass (pass = 0, 1, ...)
enum SortPhase : int {
  SORT_FILL = 0, SORT_KEYS, SORT_HIST, SORT_RSCAN, SORT_SCATTER, SORT_HEADS, SORT_FLAGS,
  SORT_SCAN_TILES, SORT_SCAN_SUMS, SORT_SCAN_ADD, SORT_EMIT, SORT_ARRANGE, SORT_CROWDED, SORT_NUM_PHASES
};
// workgroups of one scene in a phase (0: the scene sits the phase out)
int64_t sort_blocks(const SortDesc& d, int phase, int pass);
hipError_t sort_multi(int phase, int pass, const SortDesc* descs, const uint32_t* prefix, int nscenes,
                      int64_t total_blocks, hipStream_t s);
// F = cbrt(Jp) I for every particle (writes the F planes a derived-F pass left)
hipError_t liquid_f(const ParticlesDev& P, hipStream_t s);
// Sums the tiles (and the dirty parts of the atomic grid, which it clears)
// into every node of `grid`; update=1 applies the grid update (velocities),
// update=0 leaves the raw P2G sums (momentum, mass). mlp=1 suits dense scenes
// (home blocks of several items: their further tiles are read 8 at a time,
// at the cost of registers); the result does not depend on it. combined: the
// tiles went through tile_combine (one tile per home block is read).
hipError_t tile_reduce(const TileSums& in, const StepConsts& k, float4* grid, int update, int mlp,
                       int combined, hipStream_t s);
// Sums the tiles of every crowded home block (in.crowded, the sort's list)
// into its first item's tile, in item order: tile_reduce with `combined` then
// gives the same bits as without the combine.
hipError_t tile_combine(const TileSums& in, int64_t nblk, hipStream_t s);

// mesh.hip: surface mesh of the particles. The distance field (side^3
// voxels of edge vdx, float bits) starts at `cutoff` and takes the distance
// to every particle whose box of voxels reaches it; mesh_count writes the
// triangle count of each of the (side-1)^3 cubes of S = rv - distance = 0,
// mesh_emit writes them (9 floats each, and the 3 grid edges their
// vertices lie on) at the scanned offsets.
hipError_t mesh_distance_field(const ParticlesDev& P, int side, float vdx, float cutoff, uint32_t* sdf,
                               hipStream_t s);
hipError_t mesh_count(const uint32_t* sdf, int side, double rv, uint32_t* count, hipStream_t s);
hipError_t mesh_emit(const uint32_t* sdf, int side, double rv, double vdx, const uint32_t* offset, float* xyz,
                     uint64_t* keys, hipStream_t s);

// adjoint.hip: the differentiable substep (24 state planes of n floats) and
// its adjoint; grid is 4 N^3 floats of scratch, tape 8 N^3 floats (N^3
// float4 velocity and mass after the update, then N^3 float4 velocity
// before the walls and wall flags).
hipError_t diff_forward(const StepConsts& k, float mass, float volume, int64_t n, const float* st,
                        float* out, const float* mu, const float* lam, float4* grid, float* tape,
                        hipStream_t s);
// diff_forward with the scatter grid in a static device array of the
// library (kNoAllocMaxN^3 nodes per device) instead of a caller's buffer: no
// allocation. Calls that overlap on one device race on the array.
constexpr int kNoAllocMaxN = 64;
hipError_t diff_forward_noalloc(const StepConsts& k, float mass, float volume, int64_t n,
                                const float* st, float* out, const float* mu, const float* lam,
                                float* tape, hipStream_t s);
// `steps` substeps from st0: the state after substep j+1 at states + j
// state_stride, its tape at tapes + j tape_stride (tape_stride 0: one tape)
hipError_t diff_forward_steps(const StepConsts& k, float mass, float volume, int64_t n, int32_t steps,
                              const float* st0, float* states, int64_t state_stride, const float* mu,
                              const float* lam, float4* grid, float* tapes, int64_t tape_stride,
                              hipStream_t s);
// the adjoint of `steps` substeps: states + j state_stride is the state
// before substep j, tapes + j tape_stride its tape; grid2 is 8 N^3 floats of
// scratch, work 24 n floats (adj_in may be work)
hipError_t diff_backward_steps(const StepConsts& k, float mass, float volume, int64_t n, int32_t steps,
                               const float* states, int64_t state_stride, const float* mu,
                               const float* lam, const float* tapes, int64_t tape_stride,
                               const float* adj_out, float* adj_in, float* adj_mu, float* adj_lam,
                               float4* grid2, float* work, hipStream_t s);
hipError_t diff_backward(const StepConsts& k, float mass, float volume, int64_t n, const float* st,
                         const float* mu, const float* lam, const float* tape, const float* adj_out,
                         float* adj_in, float* adj_mu, float* adj_lam, float4* grid_adj,
                         hipStream_t s);

// render.hip: the device particles as the reference's debug view draws them,
// into a W x H RGB8 image (rows top to bottom) on the device
hipError_t render_points(const ParticlesDev& P, int W, int H, uint8_t* img, hipStream_t s);

}  // namespace mpmk
