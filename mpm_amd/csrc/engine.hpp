// engine.hpp -- host-side state of one simulation handle (the object behind
// mpm_handle). Stands in for the reference's class Simulation
// (include/mpm.cuh:44-106): it owns the device particle state, the two grids,
// the material table and the sort scratch, and enqueues the substep.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpm_hip.h"
#include "engine_types.hpp"
#include "kernels.hpp"

namespace mpm_engine {

bool validate_params(const mpm_params& p, std::string* why);
StepConsts make_consts(const mpm_params& p);
MaterialDev make_material_dev(const mpm_material& m, const StepConsts& k);

// Tunables (mpm_set_option names in quotes).
struct Options {
  int64_t fused = 1;        // "fused": G2P of step n fused with P2G of step n+1
  int64_t p2g_mode = 1;     // "p2g_mode": 0 direct atomics, 1 sorted + LDS tile
  int64_t interp = MPM_INTERP_QUADRATIC;  // "interpolation": kernel of the TransferScheme policy
                                          // (the cubic one runs on the direct path)
  int64_t sort_every = 16;  // "sort_every": re-sort particles every k substeps (0: never)
  int64_t radix_bits = 8;   // "radix_bits": digit width of the radix sort (4, 6 or 8)
  int64_t chunk = 0;        // "chunk": max particles per work item (64..192, multiple of 64; 0 = 192)
  int64_t block = 0;        // "block": edge of a home block in cells (4; 0 = default)
  int64_t derive_f = 1;     // "derive_f": liquid-only scenes keep F = cbrt(J) I implicit
  int64_t xcd_map = 1;      // "xcd_map": XCD-contiguous order of the fused kernel's items
  int64_t reduce_mlp = -1;  // "reduce_mlp": tile reduction batching further items of a block
                            // (1), one at a time (0), or chosen from the uploaded density (-1)
  int64_t combine = -1;     // "combine": sum the tiles of each crowded home block into one before
                            // the tile reduction (1), not (0), or chosen from the uploaded
                            // density (-1)
  int64_t markers = 0;      // "markers": roctx ranges around each advance call and its phases
  int64_t timing = 0;       // "timing": k > 0 brackets the kernels of one substep in every k
                            // with HIP events (per-phase time); 0 off
};

// Phases whose kernels the timing option brackets with HIP events.
enum Phase : int { PH_GRID_UPDATE = 0, PH_G2P2G, PH_P2G, PH_G2P, PH_SORT, PH_CLEAR, NUM_PHASES };
const char* phase_name(int p);

class Sim {
 public:
  explicit Sim(const mpm_params& p);
  ~Sim();

  std::mutex& mutex() { return mu_; }
  const std::string& last_error() const { return err_; }
  int fail(int code, const char* fmt, ...);

  int init_device();
  int set_materials(const mpm_material* m, int32_t n);
  int add_particles_soa(int64_t n, const float* x, const float* v, const float* C, const float* F,
                        const float* Jp, const uint8_t* mat);
  int add_particles_aos104(const void* p, int64_t n);
  int clear_staging();
  int upload(hipStream_t s);
  int append(hipStream_t s);
  // removes the particles with original index in [first, first + count)
  int remove(int64_t first, int64_t count, void* removed, hipStream_t s);
  // surface mesh of the device particles (mesh.hip)
  int mesh_build(int side, int radius, int64_t* ntri, hipStream_t s);
  int mesh_triangles(float* xyz, uint64_t* keys);
  int mesh_distance(int side, float* out);
  int mesh_write_obj(const char* path);
  // the last mesh with one vertex per crossed grid edge (in edge-key order)
  // and 0-based faces; V, F may be null to ask for the counts
  int mesh_indexed(int64_t* nv, int64_t* nf, double* V, int32_t* F);
  // the debug view of the device particles into a host RGB8 image (render.hip)
  int render(int W, int H, uint8_t* rgb, hipStream_t s);
  int64_t count() const { return P_; }
  double time() const { return t_; }
  int advance(int32_t substeps, hipStream_t s);
  int download_soa(float* x, float* v, float* C, float* F, float* Jp, uint8_t* mat, hipStream_t s);
  int download_aos104(void* out, hipStream_t s);
  int download_grid(float* grid, hipStream_t s);
  int save_state(const char* path);
  int load_state(const char* path);
  int set_option(const char* name, int64_t v);
  int get_option(const char* name, int64_t* v);
  int debug_p2g(float* grid, hipStream_t s);
  int debug_sort(uint32_t* keys, uint32_t* perm, hipStream_t s);
  int timing_read(const char* phase, double* total_ms, int64_t* launches);
  int timing_reset();
  // measures the substep under candidate sort intervals and item sizes from
  // a snapshot of the current state, keeps the fastest, restores the state;
  // *report receives the trials as JSON
  int autotune(int32_t trial_substeps, std::string* report, hipStream_t s);
  // a device copy of the particle state to try settings from (the batch
  // tuner); restore puts it back and restarts the step bookkeeping
  int snapshot_take(hipStream_t s);
  hipError_t snapshot_restore(hipStream_t s);
  void snapshot_drop();
  const Options& options() const { return opt_; }
  void set_options(const Options& o) { opt_ = o; ++gen_; }

  // the substep runs the tiled path: p2g_mode 1 with the quadratic kernel
  // (the tiles are sized for its 3-node stencil; other kernels run direct)
  bool tiled() const { return opt_.p2g_mode == 1 && opt_.interp == MPM_INTERP_QUADRATIC; }

  // ---- replay of captured substeps (batch.cpp) ----
  // A capture cycle is 2 sort intervals: it starts when the next substep sorts
  // and ends in the same state with the same particle buffer current.
  int cycle_substeps() const { return 2 * (int)opt_.sort_every; }
  bool at_cycle_start() const {
    return uploaded_ && !mats_dirty_ && primed_ && items_valid_ && !pending_gather_ &&
           tiled() && opt_.fused == 1 && opt_.sort_every > 0 && !opt_.timing &&
           steps_since_sort_ == opt_.sort_every;
  }
  // identifies the buffers and settings a captured cycle was recorded with
  uint64_t capture_signature() const { return (gen_ << 1) | (uint64_t)cur_; }
  // host-side bookkeeping of `substeps` substeps that ran from a replay
  void replayed(int32_t substeps) {
    for (int32_t i = 0; i < substeps; ++i) t_ += (double)prm_.dt;
  }

  // ---- one tiled fused substep in parts, for batches that launch the parts
  // of many scenes together (batch.cpp); Sim::substep runs the same parts ----
  // the scene is in the steady state of the tiled fused mode (primed, items
  // built, materials uploaded, timing off)
  bool fused_ready() const {
    return uploaded_ && P_ > 0 && !mats_dirty_ && primed_ && items_valid_ && tiled() &&
           opt_.fused == 1 && opt_.sort_every > 0 && !opt_.timing;
  }
  // part 1: the tile reduction with the grid update into the velocity grid
  mpmk::ReduceDesc reduce_desc() const;
  int64_t node_blocks() const;
  // part 2: the cell sort, when due
  bool sort_due() const {
    return !items_valid_ || (opt_.sort_every > 0 && steps_since_sort_ >= opt_.sort_every);
  }
  int sort_now(hipStream_t s) { return sort_particles(s); }
  // the same sort as arguments of the multi-scene sort kernels; sort_done()
  // after they were launched
  bool multi_sort_ok() const { return opt_.radix_bits == 8 && tiles_cap_ >= max_items(); }
  int sort_desc(mpmk::SortDesc& d);
  void sort_done() {
    pending_gather_ = true;
    steps_since_sort_ = 0;
    items_valid_ = true;
  }
  // part 3: the fused pass (writes F planes out first on `s` if needed);
  // fused_pass_done() after it was launched
  int fused_pass(const float4* grid_v, int do_g2p, mpmk::ItemDesc& d, int& kinds, int64_t& launch,
                 hipStream_t s);
  void fused_pass_done(const mpmk::ItemDesc& d);
  int64_t batch_workgroups() const;
  // the end of a substep: velocity grid, sort counter, time
  void substep_done() {
    vel_grid_ = d_grid_[0];
    grid_dead_ = true;
    ++steps_since_sort_;
    t_ += (double)prm_.dt;
  }
  const float4* velocity_grid() const { return d_grid_[0]; }
  // every material is WC_LIQUID (the fused pass's liquid-only kernel)
  bool liquid_only() const {
    for (const auto& m : mats_)
      if (m.kind != MPM_MATERIAL_WC_LIQUID) return false;
    return !mats_.empty();
  }
  int material_count() const { return (int)mats_.size(); }

  // device views for other entry points of the library
  ParticlesDev particles() const;
  mpmk::WorkItems work_items() const;
  const StepConsts& consts() const { return k_; }
  const MaterialDev* device_materials() const { return d_mats_; }

 private:
  int ensure_capacity(int64_t n);
  int ensure_grids();
  int upload_materials();
  int sort_particles(hipStream_t s);
  int substep(hipStream_t s);
  int fetch_state(std::vector<float>& planes, std::vector<uint8_t>& mat,
                  std::vector<uint32_t>& id, hipStream_t s);
  int hip_fail(hipError_t e, const char* what);
  int check_materials_cover_particles(const char* who);
  // timing: events recorded around a phase when opt_.timing is set
  int phase_begin(int ph, hipStream_t s);
  int phase_end(int ph, hipStream_t s);
  bool timing_now_ = false;  // this substep is timed
  // F derived from J: f_canonical_ when every particle's F is exactly
  // cbrt(Jp) I (at upload: F = I, Jp = 1), f_stale_ when the fused liquid
  // kernel left the F planes of the current copy unwritten
  bool f_canonical_ = false;
  bool f_stale_ = false;
  bool derives_f() const;
  bool dense_blocks(int64_t* occupied) const;  // from the staged particles
  bool use_mlp() const;
  bool dense_ = false;
  int64_t occupied_ = 0;  // home blocks holding particles at the upload
  size_t staged_on_device_ = 0;  // staged particles the last upload / append copied
  // mesh buffers: distance field, per-cube counts and offsets, scan scratch,
  // triangles; the last build's grid side and triangle count
  uint32_t* d_sdf_ = nullptr;
  uint32_t* d_mcount_ = nullptr;
  uint32_t* d_moff_ = nullptr;
  uint32_t* d_mscratch_ = nullptr;
  float* d_mxyz_ = nullptr;
  uint64_t* d_mkeys_ = nullptr;
  size_t sdf_cap_ = 0, mcube_cap_ = 0, mtri_cap_ = 0;
  uint8_t* d_img_ = nullptr;  // the last rendered image
  size_t img_cap_ = 0;
  int mesh_side_ = 0;
  int64_t mesh_ntri_ = -1;
  int ensure_f(hipStream_t s);
  int64_t timed_subs_ = 0;   // substeps since timing_reset

  struct Snapshot {
    float* f = nullptr;
    uint8_t* m = nullptr;
    uint32_t* id = nullptr;
    double t = 0.0;
    bool canonical = false;
    Options opt;
  } snap_;

  std::mutex mu_;
  std::string err_;
  int device_ = 0;
  int cus_ = 1;  // compute units of the device
  mpm_params prm_;
  StepConsts k_;
  Options opt_;
  std::vector<mpm_material> mats_;
  MaterialDev* d_mats_ = nullptr;
  bool mats_dirty_ = true;

  // host staging (SoA, upload order)
  std::vector<float> hx_, hv_, hC_, hF_, hJp_;
  std::vector<uint8_t> hmat_;

  // device particles: two buffers (reorder ping-pong)
  int64_t P_ = 0, cap_ = 0;
  size_t stride_ = 0;
  float* d_f_[2] = {nullptr, nullptr};
  uint8_t* d_mat_[2] = {nullptr, nullptr};
  uint32_t* d_id_[2] = {nullptr, nullptr};
  int cur_ = 0;
  bool uploaded_ = false;
  uint8_t max_mat_ = 0;  // largest material index among uploaded particles

  // grids d_grid_[0], d_grid_[1]. Direct P2G path: d_grid_[gcur_] holds the
  // primed P2G of the next substep. Tiled path: d_grid_[0] receives the
  // updated grid, d_grid_[1] is the side grid of particles that leave their
  // tiles (all zero, with d_dirty_, between passes while side_clean_).
  float4* d_grid_[2] = {nullptr, nullptr};
  size_t grid_nodes_ = 0;
  int gcur_ = 0;
  bool primed_ = false;              // the P2G of the current particle state is pending
  const float4* vel_grid_ = nullptr;  // grid holding the velocities of the last update
  bool side_clean_ = false;
  uint32_t* d_dirty_ = nullptr;      // one flag per 4^3 node block
  uint32_t* d_covered_ = nullptr;    // one flag per 4^3 node block (set at the sort)
  uint32_t* d_live_ = nullptr;       // one flag per 4^3 node block (set by the grid update)
  bool grid_dead_ = false;           // the velocity grid holds stale values in dead blocks
  size_t dirty_cap_ = 0;

  // work items of the tiled kernel (valid after a sort)
  int64_t nblk_ = 0;                   // Morton space of 4^3 home blocks
  uint32_t* d_block_start_ = nullptr;  // [first, end) particles of each home block
  uint32_t* d_block_items_ = nullptr;  // [first, end) items of each home block
  uint32_t* d_crowded_ = nullptr;      // count, then the home blocks with several items
  size_t block_start_cap_ = 0;
  uint32_t* d_flag_ = nullptr;         // cap entries
  uint32_t* d_pos_ = nullptr;          // cap entries
  uint32_t* d_gather_ = nullptr;       // cap entries: source slot of each new slot
  bool pending_gather_ = false;        // a sort's reorder waits for the next tiled pass
  uint32_t* d_item_start_ = nullptr;   // cap + 1 entries
  uint32_t* d_item_block_ = nullptr;   // cap entries
  uint32_t* d_item_waves_ = nullptr;   // cap entries
  uint32_t* d_item_count_ = nullptr;   // device scalar
  float4* d_tiles_ = nullptr;          // tiles_cap_ tiles of kTileNodes nodes
  int64_t tiles_cap_ = 0;
  const int lb_ = mpmk::kBlockLog2;    // log2 of the home block edge
  int chunk_ = 256;                    // particles per work item in use
  int kinds_ = 0;                      // 1: every material is WC_LIQUID
  bool items_valid_ = false;
  int64_t max_items() const;
  mpmk::TileSums tile_sums() const;
  // the tiles of the last pass (and the side grid) into `grid`: the grid
  // update (update = 1) or the raw sums; crowded blocks combined first
  int reduce_tiles(float4* grid, int update, hipStream_t s);
  bool use_combine() const;
  int tiled_p2g(hipStream_t s);        // P2G of the current state into the tiles
  // the fused kernel on the current particles; performs a pending reorder
  int run_tiled(const float4* grid_v, int do_g2p, hipStream_t s);

  // sort scratch
  uint32_t* d_keys_[2] = {nullptr, nullptr};
  uint32_t* d_vals_[2] = {nullptr, nullptr};
  uint32_t* d_hist_ = nullptr;
  size_t hist_cap_ = 0;
  int64_t steps_since_sort_ = 1 << 30;

  double t_ = 0.0;
  uint64_t gen_ = 0;  // bumped by every change of settings, materials or particles

  struct EventPair { int phase; hipEvent_t a, b; };
  std::vector<EventPair> ev_;      // recorded, not yet read
  std::vector<hipEvent_t> ev_pool_;
  double phase_ms_[NUM_PHASES] = {0};
  int64_t phase_n_[NUM_PHASES] = {0};
  hipEvent_t open_[NUM_PHASES] = {nullptr};
};

}  // namespace mpm_engine
