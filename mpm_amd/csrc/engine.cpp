// engine.cpp -- host orchestration of one simulation handle.
//
// Reference counterparts (kekeblom/mpm):
//   Sim::upload          Simulation::initCuda / particlesToDevice, src/mpm.cu:197-207,278-286
//   Sim::advance         Simulation::advance, src/mpm.cu:323-329 (t += dt in double, :328)
//   Sim::download_*      Simulation::syncDevice / particlesToHost, src/mpm.cu:209-211,288-306
//   make_consts          SimulationParameters, include/TransferScheme.h:22-28
//   make_material_dev    MaterialModel fields + p2g_prepare_particle's stress factor,
//                        include/TransferScheme.h:83
#include "engine.hpp"
#include "mesh_post.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "kernels.hpp"
#include <rocprofiler-sdk-roctx/roctx.h>

#define HIPCHK(expr, what)                 \
  do {                                     \
    hipError_t e_ = (expr);                \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

namespace mpm_engine {

bool validate_params(const mpm_params& p, std::string* why) {
  if (!(p.dt > 0.0f) || !std::isfinite(p.dt)) { *why = "dt must be positive and finite"; return false; }
  if (p.N < 4 || p.N > 1024) { *why = "N must be in [4, 1024]"; return false; }
  if (!std::isfinite(p.gravity)) { *why = "gravity must be finite"; return false; }
  if (p.boundary_mode != MPM_BOUNDARY_FRACTION && p.boundary_mode != MPM_BOUNDARY_CELLS) {
    *why = "unknown boundary mode";
    return false;
  }
  if (!std::isfinite(p.boundary) || p.boundary < 0.0f) { *why = "boundary must be >= 0"; return false; }
  if (p.boundary_mode == MPM_BOUNDARY_CELLS && p.boundary != std::floor(p.boundary)) {
    *why = "boundary must be a whole number of node layers in MPM_BOUNDARY_CELLS mode";
    return false;
  }
  if (p.compat != MPM_COMPAT_REFERENCE && p.compat != MPM_COMPAT_FIXED) { *why = "unknown compat mode"; return false; }
  return true;
}

StepConsts make_consts(const mpm_params& p) {
  StepConsts k{};
  k.N = (int32_t)p.N;
  k.dt = p.dt;
  k.dx = (float)(1.0 / (double)p.N);          // TransferScheme.h:26
  k.dx_inv = (float)(1.0 / (double)k.dx);     // TransferScheme.h:27
  k.dinv = (4.0f * k.dx_inv) * k.dx_inv;      // InterpolationKernel.cuh:72
  k.gravity = p.gravity;
  k.boundary_mode = p.boundary_mode;
  k.bfrac = p.boundary;
  k.bcells = (int32_t)p.boundary;
  k.compat = p.compat;
  const uint64_t n3 = (uint64_t)p.N * p.N * p.N;
  k.guard = (p.compat == MPM_COMPAT_REFERENCE) ? (uint32_t)n3 : 0xffffffffu;  // mpm.cu:114
  return k;
}

MaterialDev make_material_dev(const mpm_material& m, const StepConsts& k) {
  MaterialDev d{};
  d.kind = m.kind;
  d.volume = m.particle_volume;
  d.mass = m.particle_mass;
  d.mu0 = m.mu0;
  d.lambda0 = m.lambda0;
  d.hardening = m.hardening;
  d.clamp_lo = m.clamp_lo;
  d.clamp_hi = m.clamp_hi;
  d.bulk = m.bulk_modulus;
  d.stress_scale = ((-k.dinv) * k.dt) * m.particle_volume;  // TransferScheme.h:83
  return d;
}

Sim::Sim(const mpm_params& p) : prm_(p), k_(make_consts(p)) {}

Sim::~Sim() {
  for (int b = 0; b < 2; ++b) {
    if (d_f_[b]) (void)hipFree(d_f_[b]);
    if (d_mat_[b]) (void)hipFree(d_mat_[b]);
    if (d_id_[b]) (void)hipFree(d_id_[b]);
    if (d_grid_[b]) (void)hipFree(d_grid_[b]);
    if (d_keys_[b]) (void)hipFree(d_keys_[b]);
    if (d_vals_[b]) (void)hipFree(d_vals_[b]);
  }
  if (d_hist_) (void)hipFree(d_hist_);
  if (d_mats_) (void)hipFree(d_mats_);
  for (uint32_t* p : {d_block_start_, d_block_items_, d_crowded_, d_flag_, d_pos_, d_gather_, d_item_start_,
                      d_item_block_, d_item_waves_, d_item_count_, d_dirty_, d_covered_, d_live_,
                      d_sdf_, d_mcount_, d_moff_, d_mscratch_})
    if (p) (void)hipFree(p);
  if (d_tiles_) (void)hipFree(d_tiles_);
  if (d_img_) (void)hipFree(d_img_);
  if (d_mxyz_) (void)hipFree(d_mxyz_);
  if (d_mkeys_) (void)hipFree(d_mkeys_);
  for (auto& p : ev_) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto ev : ev_pool_) (void)hipEventDestroy(ev);
  for (auto ev : open_) if (ev) (void)hipEventDestroy(ev);
}

const char* phase_name(int p) {
  static const char* names[NUM_PHASES] = {"grid_update", "g2p2g", "p2g", "g2p", "sort", "clear"};
  return (p >= 0 && p < NUM_PHASES) ? names[p] : "?";
}

// Timing events skip the system-scope fence: with it, every record writes
// back and invalidates the caches, which slows the kernels it brackets
// (measured: 14 us per substep at the dam break). Only the elapsed time
// between the pair is read, after the stream has been synchronised.
static hipError_t timing_event(hipEvent_t* ev) {
  return hipEventCreateWithFlags(ev, hipEventDisableSystemFence);
}

int Sim::phase_begin(int ph, hipStream_t s) {
  if (opt_.markers) roctxRangePushA(phase_name(ph));
  if (!timing_now_) return MPM_OK;
  hipEvent_t ev;
  if (!ev_pool_.empty()) { ev = ev_pool_.back(); ev_pool_.pop_back(); }
  else HIPCHK(timing_event(&ev), "hipEventCreate");
  HIPCHK(hipEventRecord(ev, s), "hipEventRecord");
  open_[ph] = ev;
  return MPM_OK;
}

int Sim::phase_end(int ph, hipStream_t s) {
  if (opt_.markers) roctxRangePop();
  if (!timing_now_ || !open_[ph]) return MPM_OK;
  hipEvent_t ev;
  if (!ev_pool_.empty()) { ev = ev_pool_.back(); ev_pool_.pop_back(); }
  else HIPCHK(timing_event(&ev), "hipEventCreate");
  HIPCHK(hipEventRecord(ev, s), "hipEventRecord");
  ev_.push_back({ph, open_[ph], ev});
  open_[ph] = nullptr;
  return MPM_OK;
}

int Sim::timing_read(const char* phase, double* total_ms, int64_t* launches) {
  if (!phase || !total_ms || !launches) return fail(MPM_ERR_INVALID, "mpm_timing_read: null argument");
  for (auto& p : ev_) {
    HIPCHK(hipEventSynchronize(p.b), "hipEventSynchronize");
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
    phase_ms_[p.phase] += ms;
    phase_n_[p.phase] += 1;
    ev_pool_.push_back(p.a);
    ev_pool_.push_back(p.b);
  }
  ev_.clear();
  for (int ph = 0; ph < NUM_PHASES; ++ph)
    if (std::string(phase) == phase_name(ph)) {
      *total_ms = phase_ms_[ph];
      *launches = phase_n_[ph];
      return MPM_OK;
    }
  return fail(MPM_ERR_INVALID, "mpm_timing_read: unknown phase '%s'", phase);
}

int Sim::timing_reset() {
  double d;
  int64_t n;
  int rc = timing_read("g2p2g", &d, &n);  // drain pending events
  if (rc) return rc;
  for (int ph = 0; ph < NUM_PHASES; ++ph) { phase_ms_[ph] = 0.0; phase_n_[ph] = 0; }
  timed_subs_ = 0;
  return MPM_OK;
}

int Sim::fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  err_ = buf;
  return code;
}

int Sim::hip_fail(hipError_t e, const char* what) {
  return fail(MPM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}


int Sim::init_device() {
  HIPCHK(hipGetDevice(&device_), "hipGetDevice");
  HIPCHK(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, device_),
         "hipDeviceGetAttribute(CUs)");
  if (cus_ < 1) cus_ = 1;
  HIPCHK(hipMalloc(&d_mats_, sizeof(MaterialDev) * MPM_MAX_MATERIALS), "hipMalloc(materials)");
  return ensure_grids();
}

int Sim::ensure_grids() {
  const size_t nodes = (size_t)prm_.N * prm_.N * prm_.N;
  if (d_grid_[0] && grid_nodes_ == nodes) return MPM_OK;
  for (int b = 0; b < 2; ++b) {
    if (d_grid_[b]) (void)hipFree(d_grid_[b]);
    d_grid_[b] = nullptr;
    HIPCHK(hipMalloc(&d_grid_[b], nodes * sizeof(float4)), "hipMalloc(grid)");
    HIPCHK(hipMemset(d_grid_[b], 0, nodes * sizeof(float4)), "hipMemset(grid)");
  }
  grid_nodes_ = nodes;
  const size_t nbx = ((size_t)prm_.N + 3) / 4;
  for (uint32_t** p : {&d_dirty_, &d_covered_, &d_live_}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    HIPCHK(hipMalloc(p, nbx * nbx * nbx * sizeof(uint32_t)), "hipMalloc(node block flags)");
    HIPCHK(hipMemset(*p, 0, nbx * nbx * nbx * sizeof(uint32_t)), "hipMemset(node block flags)");
  }
  dirty_cap_ = nbx * nbx * nbx;
  side_clean_ = true;
  vel_grid_ = nullptr;
  primed_ = false;
  // Morton space of the 4^3 home blocks
  int bits = 0;
  while ((1 << bits) < (int)nbx) ++bits;
  nblk_ = (int64_t)1 << (3 * bits);
  if ((size_t)nblk_ > block_start_cap_) {
    for (uint32_t** p : {&d_block_start_, &d_block_items_, &d_crowded_}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    HIPCHK(hipMalloc(&d_block_start_, 2 * sizeof(uint32_t) * nblk_), "hipMalloc(block ranges)");
    HIPCHK(hipMalloc(&d_block_items_, 2 * sizeof(uint32_t) * nblk_), "hipMalloc(block items)");
    HIPCHK(hipMemset(d_block_items_, 0, 2 * sizeof(uint32_t) * nblk_), "hipMemset(block items)");
    HIPCHK(hipMalloc(&d_crowded_, sizeof(uint32_t) * (nblk_ + 1)), "hipMalloc(crowded blocks)");
    HIPCHK(hipMemset(d_crowded_, 0, sizeof(uint32_t) * (nblk_ + 1)), "hipMemset(crowded blocks)");
    block_start_cap_ = (size_t)nblk_;
  }
  items_valid_ = false;
  return MPM_OK;
}

int Sim::set_materials(const mpm_material* m, int32_t n) {
  ++gen_;
  if (n < 0 || n > MPM_MAX_MATERIALS || (n > 0 && !m))
    return fail(MPM_ERR_INVALID, "mpm_set_materials: need 0..%d materials", MPM_MAX_MATERIALS);
  for (int32_t i = 0; i < n; ++i) {
    if (m[i].kind < MPM_MATERIAL_SNOW || m[i].kind > MPM_MATERIAL_WC_LIQUID)
      return fail(MPM_ERR_INVALID, "mpm_set_materials: material %d has unknown kind %d", i, m[i].kind);
    if (!(m[i].particle_mass > 0.0f) || !(m[i].particle_volume > 0.0f))
      return fail(MPM_ERR_INVALID, "mpm_set_materials: material %d needs positive volume and mass", i);
  }
  mats_.assign(m, m + n);
  mats_dirty_ = true;
  primed_ = false;
  return MPM_OK;
}

int Sim::upload_materials() {
  if (!mats_dirty_) return MPM_OK;
  std::vector<MaterialDev> d(MPM_MAX_MATERIALS);
  for (size_t i = 0; i < mats_.size(); ++i) d[i] = make_material_dev(mats_[i], k_);
  HIPCHK(hipMemcpy(d_mats_, d.data(), sizeof(MaterialDev) * d.size(), hipMemcpyHostToDevice),
         "upload materials");
  mats_dirty_ = false;
  return MPM_OK;
}

int Sim::add_particles_soa(int64_t n, const float* x, const float* v, const float* C,
                           const float* F, const float* Jp, const uint8_t* mat) {
  if (n < 0) return fail(MPM_ERR_INVALID, "mpm_add_particles_soa: negative count");
  if (n == 0) return MPM_OK;
  if (!x || !mat) return fail(MPM_ERR_INVALID, "mpm_add_particles_soa: x and material are required");
  if ((int64_t)hJp_.size() + n > (int64_t)0x7fffffff - 64)
    return fail(MPM_ERR_CAPACITY, "mpm_add_particles_soa: more than 2^31 particles");
  const size_t o = hJp_.size();
  hx_.resize(3 * (o + n));
  hv_.resize(3 * (o + n));
  hC_.resize(9 * (o + n));
  hF_.resize(9 * (o + n));
  hJp_.resize(o + n);
  hmat_.resize(o + n);
  std::memcpy(&hx_[3 * o], x, sizeof(float) * 3 * n);
  if (v) std::memcpy(&hv_[3 * o], v, sizeof(float) * 3 * n);
  else std::fill(hv_.begin() + 3 * o, hv_.end(), 0.0f);
  if (C) std::memcpy(&hC_[9 * o], C, sizeof(float) * 9 * n);
  else std::fill(hC_.begin() + 9 * o, hC_.end(), 0.0f);
  if (F) {
    std::memcpy(&hF_[9 * o], F, sizeof(float) * 9 * n);
  } else {
    for (int64_t i = 0; i < n; ++i)
      for (int a = 0; a < 9; ++a) hF_[9 * (o + i) + a] = (a % 4 == 0) ? 1.0f : 0.0f;
  }
  if (Jp) std::memcpy(&hJp_[o], Jp, sizeof(float) * n);
  else std::fill(hJp_.begin() + o, hJp_.end(), 1.0f);
  std::memcpy(&hmat_[o], mat, n);
  return MPM_OK;
}

// Reference particle bytes: include/types.h:24-35, include/TransferScheme.h:46-54
// (u8 material@0, x@4, v@16, F@28 column-major, C@64 column-major, Jp@100).
int Sim::add_particles_aos104(const void* p, int64_t n) {
  if (n < 0 || (n > 0 && !p)) return fail(MPM_ERR_INVALID, "mpm_add_particles_aos104: bad arguments");
  const uint8_t* b = (const uint8_t*)p;
  std::vector<float> x(3 * n), v(3 * n), C(9 * n), F(9 * n), Jp(n);
  std::vector<uint8_t> mat(n);
  for (int64_t i = 0; i < n; ++i) {
    const uint8_t* r = b + 104 * i;
    mat[i] = r[0];
    std::memcpy(&x[3 * i], r + 4, 12);
    std::memcpy(&v[3 * i], r + 16, 12);
    float fc[9], cc[9];
    std::memcpy(fc, r + 28, 36);
    std::memcpy(cc, r + 64, 36);
    for (int row = 0; row < 3; ++row)
      for (int col = 0; col < 3; ++col) {
        F[9 * i + 3 * row + col] = fc[3 * col + row];
        C[9 * i + 3 * row + col] = cc[3 * col + row];
      }
    std::memcpy(&Jp[i], r + 100, 4);
  }
  return add_particles_soa(n, x.data(), v.data(), C.data(), F.data(), Jp.data(), mat.data());
}

int Sim::clear_staging() {
  hx_.clear(); hv_.clear(); hC_.clear(); hF_.clear(); hJp_.clear(); hmat_.clear();
  staged_on_device_ = 0;
  return MPM_OK;
}

int Sim::ensure_capacity(int64_t n) {
  if (n <= cap_ && d_f_[0]) return MPM_OK;
  const int64_t cap = std::max<int64_t>(64, ((n + 63) / 64) * 64);
  // the tiled kernel addresses the particle planes through one buffer
  // resource, whose byte range is 32-bit
  if ((uint64_t)cap * NUM_PLANES * sizeof(float) > 0xffffffffull)
    return fail(MPM_ERR_CAPACITY, "%lld particles exceed one handle's 4 GiB particle planes",
                (long long)n);
  for (int b = 0; b < 2; ++b) {
    if (d_f_[b]) (void)hipFree(d_f_[b]);
    if (d_mat_[b]) (void)hipFree(d_mat_[b]);
    if (d_id_[b]) (void)hipFree(d_id_[b]);
    if (d_keys_[b]) (void)hipFree(d_keys_[b]);
    if (d_vals_[b]) (void)hipFree(d_vals_[b]);
    d_f_[b] = nullptr; d_mat_[b] = nullptr; d_id_[b] = nullptr; d_keys_[b] = nullptr; d_vals_[b] = nullptr;
  }
  cap_ = 0;
  for (int b = 0; b < 2; ++b) {
    HIPCHK(hipMalloc(&d_f_[b], sizeof(float) * NUM_PLANES * cap), "hipMalloc(particles)");
    HIPCHK(hipMalloc(&d_mat_[b], cap), "hipMalloc(material ids)");
    HIPCHK(hipMalloc(&d_id_[b], sizeof(uint32_t) * cap), "hipMalloc(particle ids)");
    HIPCHK(hipMalloc(&d_keys_[b], sizeof(uint32_t) * cap), "hipMalloc(sort keys)");
    HIPCHK(hipMalloc(&d_vals_[b], sizeof(uint32_t) * cap), "hipMalloc(sort values)");
  }
  for (uint32_t** p : {&d_flag_, &d_pos_, &d_gather_, &d_item_start_, &d_item_block_, &d_item_waves_}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  HIPCHK(hipMalloc(&d_flag_, sizeof(uint32_t) * cap), "hipMalloc(item flags)");
  HIPCHK(hipMalloc(&d_pos_, sizeof(uint32_t) * cap), "hipMalloc(item positions)");
  HIPCHK(hipMalloc(&d_gather_, sizeof(uint32_t) * cap), "hipMalloc(item arrangement)");
  HIPCHK(hipMalloc(&d_item_start_, sizeof(uint32_t) * (cap + 1)), "hipMalloc(item starts)");
  HIPCHK(hipMalloc(&d_item_block_, sizeof(uint32_t) * cap), "hipMalloc(item blocks)");
  HIPCHK(hipMalloc(&d_item_waves_, sizeof(uint32_t) * cap), "hipMalloc(item waves)");
  if (!d_item_count_) HIPCHK(hipMalloc(&d_item_count_, sizeof(uint32_t)), "hipMalloc(item count)");
  const size_t words = mpmk::radix_scratch_words(cap, 8);
  if (words > hist_cap_) {
    if (d_hist_) (void)hipFree(d_hist_);
    d_hist_ = nullptr;
    HIPCHK(hipMalloc(&d_hist_, words * sizeof(uint32_t)), "hipMalloc(sort scratch)");
    hist_cap_ = words;
  }
  cap_ = cap;
  stride_ = (size_t)cap;
  return MPM_OK;
}

ParticlesDev Sim::particles() const {
  ParticlesDev P;
  P.f = d_f_[cur_];
  P.mat = d_mat_[cur_];
  P.id = d_id_[cur_];
  P.stride = stride_;
  P.n = (int32_t)P_;
  return P;
}

int Sim::upload(hipStream_t s) {
  ++gen_;
  const int64_t n = (int64_t)hJp_.size();
  for (int64_t i = 0; i < n; ++i)
    if (hmat_[i] >= mats_.size())
      return fail(MPM_ERR_INVALID, "mpm_upload: particle %lld uses material %d but only %zu are set",
                  (long long)i, (int)hmat_[i], mats_.size());
  int rc = ensure_capacity(n);
  if (rc) return rc;
  rc = upload_materials();
  if (rc) return rc;
  // transpose the staging set into planes
  std::vector<float> planes((size_t)NUM_PLANES * stride_, 0.0f);
  for (int64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      planes[(PL_X + a) * stride_ + i] = hx_[3 * i + a];
      planes[(PL_V + a) * stride_ + i] = hv_[3 * i + a];
    }
    for (int a = 0; a < 9; ++a) {
      planes[(PL_C + a) * stride_ + i] = hC_[9 * i + a];
      planes[(PL_F + a) * stride_ + i] = hF_[9 * i + a];
    }
    planes[PL_JP * stride_ + i] = hJp_[i];
  }
  std::vector<uint32_t> ids(n);
  std::iota(ids.begin(), ids.end(), 0u);
  f_canonical_ = true;
  for (int64_t i = 0; i < n && f_canonical_; ++i) {
    f_canonical_ = hJp_[i] == 1.0f;
    for (int a = 0; a < 9; ++a) f_canonical_ = f_canonical_ && hF_[9 * i + a] == ((a % 4 == 0) ? 1.0f : 0.0f);
  }
  f_stale_ = false;
  dense_ = dense_blocks(&occupied_);
  max_mat_ = 0;
  for (int64_t i = 0; i < n; ++i) max_mat_ = std::max<uint8_t>(max_mat_, hmat_[i]);
  cur_ = 0;
  HIPCHK(hipMemcpyAsync(d_f_[0], planes.data(), sizeof(float) * planes.size(), hipMemcpyHostToDevice, s), "upload particles");
  if (n > 0) {
    HIPCHK(hipMemcpyAsync(d_mat_[0], hmat_.data(), n, hipMemcpyHostToDevice, s), "upload materials ids");
    HIPCHK(hipMemcpyAsync(d_id_[0], ids.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s), "upload ids");
  }
  HIPCHK(hipStreamSynchronize(s), "upload sync");
  P_ = n;
  staged_on_device_ = (size_t)n;
  uploaded_ = true;
  items_valid_ = false;
  pending_gather_ = false;
  primed_ = false;
  vel_grid_ = nullptr;
  steps_since_sort_ = 1 << 30;
  t_ = 0.0;
  return MPM_OK;
}

// Adds the particles staged since the last upload / append (or clear) after
// the ones on the device, which stay as they are (no copy back to the
// host); the new ones take the next original indices. Simulated time goes
// on. The items are rebuilt at the next substep (a sort, then the P2G of
// the whole set).
int Sim::append(hipStream_t s) {
  ++gen_;
  if (!uploaded_) return fail(MPM_ERR_STATE, "mpm_append_particles: nothing uploaded");
  const size_t total = hJp_.size();
  if (total < staged_on_device_) staged_on_device_ = 0;
  const int64_t n = (int64_t)(total - staged_on_device_), o = (int64_t)staged_on_device_;
  if (n == 0) return MPM_OK;
  for (int64_t i = o; i < (int64_t)total; ++i)
    if (hmat_[i] >= mats_.size())
      return fail(MPM_ERR_INVALID, "mpm_append_particles: particle %lld uses material %d but only %zu are set",
                  (long long)(i - o), (int)hmat_[i], mats_.size());
  if (P_ + n > (int64_t)0x7fffffff - 64)
    return fail(MPM_ERR_CAPACITY, "mpm_append_particles: more than 2^31 particles");
  bool canon = true;
  for (int64_t i = o; i < (int64_t)total && canon; ++i) {
    canon = hJp_[i] == 1.0f;
    for (int a = 0; a < 9; ++a) canon = canon && hF_[9 * i + a] == ((a % 4 == 0) ? 1.0f : 0.0f);
  }
  if (!canon) {  // F stops being implicit: write out the device particles' F first
    int rc = ensure_f(s);
    if (rc) return rc;
  }
  int rc = upload_materials();
  if (rc) return rc;
  if (P_ + n > cap_) {
    // grow: keep the device particles in a compact copy across the reallocation
    float* tf = nullptr;
    uint8_t* tm = nullptr;
    uint32_t* ti = nullptr;
    const size_t P = (size_t)P_;
    hipError_t e = hipSuccess;
    if (P > 0) {
      e = hipMalloc(&tf, sizeof(float) * NUM_PLANES * P);
      if (e == hipSuccess) e = hipMalloc(&tm, P);
      if (e == hipSuccess) e = hipMalloc(&ti, sizeof(uint32_t) * P);
      if (e == hipSuccess)
        e = hipMemcpy2DAsync(tf, sizeof(float) * P, d_f_[cur_], sizeof(float) * stride_, sizeof(float) * P,
                             NUM_PLANES, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess) e = hipMemcpyAsync(tm, d_mat_[cur_], P, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess) e = hipMemcpyAsync(ti, d_id_[cur_], sizeof(uint32_t) * P, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e == hipSuccess) {
      rc = ensure_capacity(P_ + n);
      cur_ = 0;
      if (!rc && P > 0) {
        e = hipMemcpy2DAsync(d_f_[0], sizeof(float) * stride_, tf, sizeof(float) * P, sizeof(float) * P,
                             NUM_PLANES, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(d_mat_[0], tm, P, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(d_id_[0], ti, sizeof(uint32_t) * P, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
      }
    }
    if (tf) (void)hipFree(tf);
    if (tm) (void)hipFree(tm);
    if (ti) (void)hipFree(ti);
    if (rc) return rc;
    HIPCHK(e, "mpm_append_particles: keeping the device particles");
  }
  // the new particles as planes of n floats, copied behind the P_ on the device
  std::vector<float> planes((size_t)NUM_PLANES * n, 0.0f);
  for (int64_t j = 0; j < n; ++j) {
    const int64_t i = o + j;
    for (int a = 0; a < 3; ++a) {
      planes[(PL_X + a) * n + j] = hx_[3 * i + a];
      planes[(PL_V + a) * n + j] = hv_[3 * i + a];
    }
    for (int a = 0; a < 9; ++a) {
      planes[(PL_C + a) * n + j] = hC_[9 * i + a];
      planes[(PL_F + a) * n + j] = hF_[9 * i + a];
    }
    planes[PL_JP * n + j] = hJp_[i];
  }
  std::vector<uint32_t> ids((size_t)n);
  std::iota(ids.begin(), ids.end(), (uint32_t)P_);
  HIPCHK(hipMemcpy2DAsync(d_f_[cur_] + P_, sizeof(float) * stride_, planes.data(), sizeof(float) * n,
                          sizeof(float) * n, NUM_PLANES, hipMemcpyHostToDevice, s),
         "append particles");
  HIPCHK(hipMemcpyAsync(d_mat_[cur_] + P_, &hmat_[o], n, hipMemcpyHostToDevice, s), "append material ids");
  HIPCHK(hipMemcpyAsync(d_id_[cur_] + P_, ids.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s),
         "append ids");
  HIPCHK(hipStreamSynchronize(s), "append sync");
  for (int64_t i = o; i < (int64_t)total; ++i) max_mat_ = std::max<uint8_t>(max_mat_, hmat_[i]);
  f_canonical_ = f_canonical_ && canon;
  P_ += n;
  staged_on_device_ = total;
  items_valid_ = false;
  pending_gather_ = false;  // the next sort starts from the current order
  primed_ = false;
  side_clean_ = false;      // the last pass's side-grid scatters are dropped with its tiles
  steps_since_sort_ = 1 << 30;
  return MPM_OK;
}

// Removes the particles with original index in [first, first + count) on the
// device: the others are packed into the other particle copy in their order,
// with the indices above the range lowered by count, so downloads list them
// as before without the removed ones. `removed` (when non-null) receives the
// removed particles' state in the reference's 104-byte layout, in index
// order. The items are rebuilt at the next substep (a sort, then the P2G of
// the remaining set), as after an append.
int Sim::remove(int64_t first, int64_t count, void* removed, hipStream_t s) {
  ++gen_;
  if (!uploaded_) return fail(MPM_ERR_STATE, "mpm_remove_particles: nothing uploaded");
  if (first < 0 || count < 0 || first + count > P_)
    return fail(MPM_ERR_INVALID, "mpm_remove_particles: range [%lld, %lld) outside the %lld particles",
                (long long)first, (long long)(first + count), (long long)P_);
  if (count == 0) return MPM_OK;
  if (removed) {  // the F planes, when the liquid's pass left them implicit
    const int rc = ensure_f(s);
    if (rc) return rc;
  }
  float* gone = nullptr;
  uint8_t* gone_mat = nullptr;
  if (removed) {
    HIPCHK(hipMalloc(&gone, sizeof(float) * NUM_PLANES * (size_t)count), "hipMalloc(removed particles)");
    hipError_t e = hipMalloc(&gone_mat, (size_t)count);
    if (e != hipSuccess) {
      (void)hipFree(gone);
      return hip_fail(e, "hipMalloc(removed particles)");
    }
  }
  ParticlesDev src = particles();
  ParticlesDev dst = src;
  dst.f = d_f_[cur_ ^ 1];
  dst.mat = d_mat_[cur_ ^ 1];
  dst.id = d_id_[cur_ ^ 1];
  hipError_t e = mpmk::remove_range(src, dst, (uint32_t)first, (uint32_t)(first + count), d_flag_, d_pos_,
                                    d_hist_, d_item_count_, gone, gone_mat, s);
  std::vector<float> planes;
  std::vector<uint8_t> mats;
  if (e == hipSuccess && removed) {
    planes.resize((size_t)NUM_PLANES * (size_t)count);
    mats.resize((size_t)count);
    e = hipMemcpyAsync(planes.data(), gone, sizeof(float) * planes.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(mats.data(), gone_mat, (size_t)count, hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (gone) (void)hipFree(gone);
  if (gone_mat) (void)hipFree(gone_mat);
  HIPCHK(e, "mpm_remove_particles");
  if (removed) {
    // planes -> the reference's particle bytes (types.h:24-35, TransferScheme.h:46-54)
    uint8_t* out = static_cast<uint8_t*>(removed);
    const size_t m = (size_t)count;
    for (size_t j = 0; j < m; ++j) {
      uint8_t* rec = out + 104 * j;
      std::memset(rec, 0, 104);
      rec[0] = mats[j];
      float x[3], v[3], F[9], C[9];
      for (int a = 0; a < 3; ++a) {
        x[a] = planes[(PL_X + a) * m + j];
        v[a] = planes[(PL_V + a) * m + j];
      }
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {  // row-major planes -> column-major bytes
          F[3 * c + r] = planes[(PL_F + 3 * r + c) * m + j];
          C[3 * c + r] = planes[(PL_C + 3 * r + c) * m + j];
        }
      const float jp = planes[PL_JP * m + j];
      std::memcpy(rec + 4, x, 12);
      std::memcpy(rec + 16, v, 12);
      std::memcpy(rec + 28, F, 36);
      std::memcpy(rec + 64, C, 36);
      std::memcpy(rec + 100, &jp, 4);
    }
  }
  cur_ ^= 1;
  P_ -= count;
  items_valid_ = false;
  pending_gather_ = false;  // the next sort starts from the packed order
  primed_ = false;
  side_clean_ = false;      // the last pass's side-grid scatters are dropped with its tiles
  steps_since_sort_ = 1 << 30;
  return MPM_OK;
}

// Surface mesh (mesh.hip) of the device particles, as the reference's
// MeshBuilder::computeMesh does it on the host (mesh_builder.h:181-211):
// voxel_dx = N dx / side (float N dx, then double), distance cut-off
// (radius + 1) voxel_dx, iso-surface at radius * voxel_dx.
int Sim::mesh_build(int side, int radius, int64_t* ntri, hipStream_t s) {
  ++gen_;
  if (!uploaded_) return fail(MPM_ERR_STATE, "mpm_mesh_build: nothing uploaded");
  if (side < 2 || side > 1024) return fail(MPM_ERR_INVALID, "mpm_mesh_build: grid side must be 2..1024");
  if (radius < 0 || radius > side) return fail(MPM_ERR_INVALID, "mpm_mesh_build: radius must be 0..side");
  const double vdx = (double)((float)k_.N * k_.dx) / (double)side;
  const float vdx_f = (float)vdx, cutoff = (float)((double)(radius + 1) * vdx);
  const double rv = (double)radius * vdx;
  const size_t nv = (size_t)side * side * side, nc = (size_t)(side - 1) * (side - 1) * (side - 1);
  if (nv > sdf_cap_) {
    if (d_sdf_) (void)hipFree(d_sdf_);
    d_sdf_ = nullptr;
    sdf_cap_ = 0;
    HIPCHK(hipMalloc(&d_sdf_, sizeof(uint32_t) * nv), "hipMalloc(distance field)");
    sdf_cap_ = nv;
  }
  if (nc > mcube_cap_) {
    for (uint32_t** q : {&d_mcount_, &d_moff_, &d_mscratch_}) {
      if (*q) (void)hipFree(*q);
      *q = nullptr;
    }
    mcube_cap_ = 0;
    HIPCHK(hipMalloc(&d_mcount_, sizeof(uint32_t) * nc), "hipMalloc(mesh counts)");
    HIPCHK(hipMalloc(&d_moff_, sizeof(uint32_t) * (nc + 1)), "hipMalloc(mesh offsets)");
    // scan scratch: one sum per tile of 2048 counts, and the total at the end
    HIPCHK(hipMalloc(&d_mscratch_, sizeof(uint32_t) * (nc / 2048 + 2)), "hipMalloc(mesh scan)");
    mcube_cap_ = nc;
  }
  HIPCHK(mpmk::mesh_distance_field(particles(), side, vdx_f, cutoff, d_sdf_, s), "mesh distance field");
  HIPCHK(mpmk::mesh_count(d_sdf_, side, rv, d_mcount_, s), "mesh count");
  uint32_t* total = d_moff_ + nc;
  HIPCHK(mpmk::exclusive_scan(d_mcount_, d_moff_, (int64_t)nc, d_mscratch_, total, s), "mesh scan");
  uint32_t n = 0;
  HIPCHK(hipMemcpyAsync(&n, total, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "mesh size");
  HIPCHK(hipStreamSynchronize(s), "mesh size");
  if ((size_t)n > mtri_cap_) {
    if (d_mxyz_) (void)hipFree(d_mxyz_);
    if (d_mkeys_) (void)hipFree(d_mkeys_);
    d_mxyz_ = nullptr;
    d_mkeys_ = nullptr;
    mtri_cap_ = 0;
    HIPCHK(hipMalloc(&d_mxyz_, sizeof(float) * 9 * (size_t)n), "hipMalloc(mesh triangles)");
    HIPCHK(hipMalloc(&d_mkeys_, sizeof(uint64_t) * 3 * (size_t)n), "hipMalloc(mesh edges)");
    mtri_cap_ = n;
  }
  if (n > 0) HIPCHK(mpmk::mesh_emit(d_sdf_, side, rv, vdx, d_moff_, d_mxyz_, d_mkeys_, s), "mesh emit");
  HIPCHK(hipStreamSynchronize(s), "mesh emit");
  mesh_side_ = side;
  mesh_ntri_ = n;
  if (ntri) *ntri = n;
  return MPM_OK;
}

int Sim::mesh_triangles(float* xyz, uint64_t* keys) {
  if (mesh_ntri_ < 0) return fail(MPM_ERR_STATE, "mpm_mesh_triangles: no mesh built");
  const size_t n = (size_t)mesh_ntri_;
  if (n == 0) return MPM_OK;
  if (xyz) HIPCHK(hipMemcpy(xyz, d_mxyz_, sizeof(float) * 9 * n, hipMemcpyDeviceToHost), "mesh triangles");
  if (keys) HIPCHK(hipMemcpy(keys, d_mkeys_, sizeof(uint64_t) * 3 * n, hipMemcpyDeviceToHost), "mesh edges");
  return MPM_OK;
}

int Sim::mesh_distance(int side, float* out) {
  if (mesh_ntri_ < 0) return fail(MPM_ERR_STATE, "mpm_mesh_distance_field: no mesh built");
  if (!out || side != mesh_side_)
    return fail(MPM_ERR_INVALID, "mpm_mesh_distance_field: the last mesh has grid side %d", mesh_side_);
  const size_t nv = (size_t)mesh_side_ * mesh_side_ * mesh_side_;
  HIPCHK(hipMemcpy(out, d_sdf_, sizeof(float) * nv, hipMemcpyDeviceToHost), "mesh distance field");
  return MPM_OK;
}

// The last mesh indexed: one vertex per grid edge the surface crosses (the
// triangles that share an edge share its vertex bits), in the order of the
// edge keys; faces 0-based.
int Sim::mesh_indexed(int64_t* nv, int64_t* nf, double* V, int32_t* F) {
  if (mesh_ntri_ < 0) return fail(MPM_ERR_STATE, "mpm_mesh_indexed: no mesh built");
  const size_t n = (size_t)mesh_ntri_;
  std::vector<float> xyz(9 * n);
  std::vector<uint64_t> keys(3 * n);
  int rc = mesh_triangles(xyz.data(), keys.data());
  if (rc) return rc;
  std::vector<uint64_t> order(keys);
  std::sort(order.begin(), order.end());
  order.erase(std::unique(order.begin(), order.end()), order.end());
  if (order.size() > (size_t)INT32_MAX) return fail(MPM_ERR_CAPACITY, "mpm_mesh_indexed: too many vertices");
  if (nv) *nv = (int64_t)order.size();
  if (nf) *nf = (int64_t)n;
  if (!V && !F) return MPM_OK;
  for (size_t t = 0; t < 3 * n; ++t) {
    const size_t v = (size_t)(std::lower_bound(order.begin(), order.end(), keys[t]) - order.begin());
    if (F) F[t] = (int32_t)v;
    if (V)
      for (int a = 0; a < 3; ++a) V[3 * v + a] = (double)xyz[3 * t + a];
  }
  return MPM_OK;
}

// The reference's debug view (Renderer::render, renderer.h:48-59) of the
// device particles, without a copy of them to the host. Blocking.
int Sim::render(int W, int H, uint8_t* rgb, hipStream_t s) {
  if (!uploaded_) return fail(MPM_ERR_STATE, "mpm_render_frame: nothing uploaded");
  if (W < 1 || H < 1 || W > 16384 || H > 16384 || !rgb)
    return fail(MPM_ERR_INVALID, "mpm_render_frame: need 1 <= width, height <= 16384 and an image buffer");
  const size_t bytes = (size_t)3 * (size_t)W * (size_t)H;
  if (bytes > img_cap_) {
    if (d_img_) (void)hipFree(d_img_);
    d_img_ = nullptr;
    img_cap_ = 0;
    HIPCHK(hipMalloc(&d_img_, bytes), "hipMalloc(image)");
    img_cap_ = bytes;
  }
  HIPCHK(mpmk::render_points(particles(), W, H, d_img_, s), "render");
  HIPCHK(hipMemcpyAsync(rgb, d_img_, bytes, hipMemcpyDeviceToHost, s), "image copy");
  HIPCHK(hipStreamSynchronize(s), "image copy");
  return MPM_OK;
}

// OBJ of the last mesh as built (no smoothing or decimation)
int Sim::mesh_write_obj(const char* path) {
  if (mesh_ntri_ < 0) return fail(MPM_ERR_STATE, "mpm_mesh_write_obj: no mesh built");
  if (!path) return fail(MPM_ERR_INVALID, "mpm_mesh_write_obj: null path");
  int64_t nv = 0, nf = 0;
  int rc = mesh_indexed(&nv, &nf, nullptr, nullptr);
  if (rc) return rc;
  std::vector<double> V(3 * (size_t)nv);
  std::vector<int32_t> F(3 * (size_t)nf);
  rc = mesh_indexed(&nv, &nf, V.data(), F.data());
  if (rc) return rc;
  if (!mpm_mesh::write_obj(path, nv, V.data(), nf, F.data()))
    return fail(MPM_ERR_IO, "mpm_mesh_write_obj: cannot write %s", path);
  return MPM_OK;
}

#define PHASE(ph, expr)                 \
  do {                                  \
    int rc_ = phase_begin(ph, s);       \
    if (rc_) return rc_;                \
    expr;                               \
    rc_ = phase_end(ph, s);             \
    if (rc_) return rc_;                \
  } while (0)

int Sim::sort_particles(hipStream_t s) {
  chunk_ = opt_.chunk > 0 ? (int)opt_.chunk : mpmk::kMaxItemParticles;
  const int64_t items = std::max<int64_t>(1, max_items());
  if (items > tiles_cap_) {
    if (d_tiles_) (void)hipFree(d_tiles_);
    d_tiles_ = nullptr;
    tiles_cap_ = 0;
    HIPCHK(hipMalloc(&d_tiles_, sizeof(float4) * mpmk::kTileNodes * (size_t)items), "hipMalloc(tiles)");
    tiles_cap_ = items;
  }
  if (P_ <= 0) {
    // no items: every home block has an empty item range, no node block is reached
    HIPCHK(mpmk::fill_u32(d_block_items_, 2 * nblk_, 0u, s), "clear block items");
    HIPCHK(mpmk::fill_u32(d_crowded_, 1, 0u, s), "clear crowded blocks");
    HIPCHK(mpmk::fill_u32(d_covered_, (int64_t)dirty_cap_, 0u, s), "clear node block flags");
    steps_since_sort_ = 0;
    items_valid_ = true;
    return MPM_OK;
  }
  const ParticlesDev src = particles();
  int out = 0;
  PHASE(PH_SORT, {
    HIPCHK(mpmk::cell_keys(src, k_, lb_, d_keys_[0], d_vals_[0], s), "cell keys");
    // by home block only: k_item_arrange orders each item by cell
    HIPCHK(mpmk::radix_sort(d_keys_, d_vals_, P_, 3 * lb_, mpmk::cell_key_bits(k_.N, lb_),
                            (int)opt_.radix_bits, d_hist_, &out, s),
           "radix sort");
    HIPCHK(mpmk::build_items(d_keys_[out], d_vals_[out], P_, lb_, chunk_, nblk_, d_block_start_,
                             d_block_items_, d_covered_, (k_.N + 3) / 4, d_flag_, d_pos_, d_hist_,
                             d_item_count_, d_item_start_, d_item_block_, d_item_waves_, d_gather_,
                             items, d_crowded_, s),
           "work items");
  });
  // the particles move to their new slots in the next tiled pass
  pending_gather_ = true;
  steps_since_sort_ = 0;
  items_valid_ = true;
  return MPM_OK;
}

int Sim::sort_desc(mpmk::SortDesc& d) {
  chunk_ = opt_.chunk > 0 ? (int)opt_.chunk : mpmk::kMaxItemParticles;
  const int64_t items = std::max<int64_t>(1, max_items());
  // (the multi-scene sort always uses 8-bit digits: a stable sort gives the
  // same order as the scene's own sort with any "radix_bits")
  if (P_ <= 0 || items > tiles_cap_)
    return fail(MPM_ERR_STATE, "sort_desc: needs particles and allocated tiles");
  d.P = particles();
  d.k = k_;
  d.keys[0] = d_keys_[0];
  d.keys[1] = d_keys_[1];
  d.vals[0] = d_vals_[0];
  d.vals[1] = d_vals_[1];
  d.scratch = d_hist_;
  d.n = P_;
  d.lb = lb_;
  d.start_bit = 3 * lb_;
  d.key_bits = mpmk::cell_key_bits(k_.N, lb_);
  d.npasses = (P_ > 1 && d.key_bits > d.start_bit) ? (d.key_bits - d.start_bit + 7) / 8 : 0;
  d.block_range = d_block_start_;
  d.block_items = d_block_items_;
  d.covered = d_covered_;
  d.crowded = d_crowded_;
  d.nblk = nblk_;
  d.nbx = (k_.N + 3) / 4;
  d.flag = d_flag_;
  d.pos = d_pos_;
  d.item_count = d_item_count_;
  d.item_start = d_item_start_;
  d.item_block = d_item_block_;
  d.item_waves = d_item_waves_;
  d.gather = d_gather_;
  d.chunk = chunk_;
  d.max_items = items;
  return MPM_OK;
}

// Whether the staged particles crowd their home blocks: more than 1 in 20
// occupied blocks needs a further work item (more particles than one item
// takes). The grid update then reads several tiles per node and block,
// which the mlp variant of the tile reduction batches.
bool Sim::dense_blocks(int64_t* occupied_out) const {
  const int64_t n = (int64_t)hJp_.size();
  *occupied_out = 0;
  if (n == 0) return false;
  const int N = (int)k_.N, nb = (N + 3) / 4;
  const int64_t per = opt_.chunk > 0 ? opt_.chunk : mpmk::kMaxItemParticles;
  std::vector<uint64_t> keys((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    uint64_t key = 0;
    for (int a = 0; a < 3; ++a) {
      const int c = std::min(N - 1, std::max(0, (int)std::floor(hx_[3 * i + a] * (float)N)));
      key = key * (uint64_t)nb + (uint64_t)(c >> 2);
    }
    keys[(size_t)i] = key;
  }
  std::sort(keys.begin(), keys.end());
  int64_t occupied = 0, extra = 0;
  for (size_t i = 0; i < keys.size();) {
    size_t j = i;
    while (j < keys.size() && keys[j] == keys[i]) ++j;
    ++occupied;
    extra += ((int64_t)(j - i) + per - 1) / per - 1;
    i = j;
  }
  *occupied_out = occupied;
  return extra * 20 > occupied;
}

// Workgroups of this scene's share of a batch's fused launch: its items (at
// most one per `chunk` particles plus one per occupied home block) with the
// upload's occupancy and room for it to grow by half, plus 64; a workgroup
// takes further items of the scene in turn when there are more. At most
// max_items(), the bound that needs no loop.
int64_t Sim::batch_workgroups() const {
  const int64_t est = (P_ + chunk_ - 1) / chunk_ + occupied_ + occupied_ / 2 + 64;
  return std::max<int64_t>(1, std::min(max_items(), est));
}

bool Sim::use_mlp() const { return opt_.reduce_mlp < 0 ? dense_ : opt_.reduce_mlp != 0; }
bool Sim::use_combine() const { return opt_.combine < 0 ? dense_ : opt_.combine != 0; }

int Sim::reduce_tiles(float4* grid, int update, hipStream_t s) {
  const mpmk::TileSums t = tile_sums();
  const bool comb = use_combine();
  if (comb) HIPCHK(mpmk::tile_combine(t, nblk_, s), "tile combine");
  // (combined: one tile per block, which the plain variant reads best)
  HIPCHK(mpmk::tile_reduce(t, k_, grid, update, comb ? 0 : use_mlp(), comb ? 1 : 0, s), "grid update");
  return MPM_OK;
}

bool Sim::derives_f() const {
  // only the fused tiled path: the unfused one's G2P kernel reads the F planes
  if (!opt_.derive_f || !f_canonical_ || !tiled() || !opt_.fused || mats_.empty()) return false;
  for (const auto& m : mats_)
    if (m.kind != MPM_MATERIAL_WC_LIQUID) return false;
  return true;
}

int Sim::ensure_f(hipStream_t s) {
  if (!f_stale_) return MPM_OK;
  HIPCHK(mpmk::liquid_f(particles(), s), "liquid F");
  f_stale_ = false;
  return MPM_OK;
}

int Sim::fused_pass(const float4* grid_v, int do_g2p, mpmk::ItemDesc& d, int& kinds, int64_t& launch,
                    hipStream_t s) {
  // the liquid-only instantiation when every material is WC_LIQUID
  kinds_ = 1;
  for (const auto& m : mats_)
    if (m.kind != MPM_MATERIAL_WC_LIQUID) kinds_ = 0;
  const bool derive = kinds_ == 1 && derives_f();
  if (!derive) {
    int rc = ensure_f(s);
    if (rc) return rc;
  }
  int flags = do_g2p ? mpmk::TILED_G2P : 0;
  if (derive) flags |= mpmk::TILED_DERIVE_F;
  if (P_ > 0 && (uint64_t)(P_ - 1) > (uint64_t)k_.guard) flags |= mpmk::TILED_GUARD;
  if (opt_.xcd_map) flags |= mpmk::TILED_XCD;
  const ParticlesDev P = particles();
  d.P = P;
  d.Pdst = P;
  d.gather = nullptr;
  if (pending_gather_) {
    d.Pdst.f = d_f_[cur_ ^ 1];
    d.Pdst.mat = d_mat_[cur_ ^ 1];
    d.Pdst.id = d_id_[cur_ ^ 1];
    d.gather = d_gather_;
  }
  const mpmk::WorkItems w = work_items();
  const mpmk::TileSums t = tile_sums();
  d.mats = d_mats_;
  d.k = k_;
  d.item_start = w.start;
  d.item_block = w.block;
  d.item_waves = w.waves;
  d.item_count = w.count;
  d.grid_v = grid_v;
  d.tiles = t.tiles;
  d.grid_atomic = (float*)t.atomic;
  d.dirty = t.dirty;
  d.flags = flags;
  kinds = kinds_;
  launch = w.launch;
  return MPM_OK;
}

void Sim::fused_pass_done(const mpmk::ItemDesc& d) {
  if (pending_gather_) {
    cur_ ^= 1;
    pending_gather_ = false;
  }
  if (d.flags & mpmk::TILED_DERIVE_F) f_stale_ = true;
}

mpmk::ReduceDesc Sim::reduce_desc() const {
  const mpmk::TileSums t = tile_sums();
  mpmk::ReduceDesc r;
  r.tiles = t.tiles;
  r.block_items = t.block_items;
  r.covered = t.covered;
  r.grid_atomic = t.atomic;
  r.dirty = t.dirty;
  r.grid = d_grid_[0];
  r.k = k_;
  r.nbx = (k_.N + 3) / 4;
  r.update = 1;
  r.combined = 0;
  r.nblk = nblk_;
  r.crowded = d_crowded_;
  r.live = d_live_;
  r.skip_dead = 1;
  return r;
}

int64_t Sim::node_blocks() const {
  const int64_t nbx = (k_.N + 3) / 4;
  return nbx * nbx * nbx;
}

int Sim::run_tiled(const float4* grid_v, int do_g2p, hipStream_t s) {
  mpmk::ItemDesc d;
  int kinds = 0;
  int64_t launch = 0;
  int rc = fused_pass(grid_v, do_g2p, d, kinds, launch, s);
  if (rc) return rc;
  HIPCHK(mpmk::g2p2g_desc(d, (int)mats_.size(), launch, kinds, s), do_g2p ? "g2p2g" : "p2g");
  fused_pass_done(d);
  return MPM_OK;
}

int64_t Sim::max_items() const {
  // one item per started `chunk` particles, plus one per occupied block
  return (P_ + chunk_ - 1) / chunk_ + std::min<int64_t>(P_, nblk_);
}

mpmk::WorkItems Sim::work_items() const {
  mpmk::WorkItems w;
  w.start = d_item_start_;
  w.block = d_item_block_;
  w.waves = d_item_waves_;
  w.count = d_item_count_;
  w.max_items = max_items();
  // one workgroup per possible item; those beyond the item count exit at once
  w.launch = std::max<int64_t>(1, w.max_items);
  return w;
}

mpmk::TileSums Sim::tile_sums() const {
  mpmk::TileSums t;
  t.tiles = d_tiles_;
  t.atomic = d_grid_[1];
  t.dirty = d_dirty_;
  t.block_items = d_block_items_;
  t.covered = d_covered_;
  t.live = d_live_;
  t.crowded = d_crowded_;
  return t;
}

// P2G of the current particle state into the tiles (and the side grid).
// Sorts first when the items are stale.
int Sim::tiled_p2g(hipStream_t s) {
  if (!items_valid_ || (opt_.sort_every > 0 && steps_since_sort_ >= opt_.sort_every)) {
    int rc = sort_particles(s);
    if (rc) return rc;
  }
  if (!side_clean_) {
    PHASE(PH_CLEAR, {
      HIPCHK(mpmk::fill_zero(d_grid_[1], grid_nodes_, s), "clear side grid");
      HIPCHK(mpmk::fill_u32(d_dirty_, (int64_t)dirty_cap_, 0u, s), "clear dirty flags");
    });
    side_clean_ = true;
  }
  int rc = MPM_OK;
  PHASE(PH_P2G, rc = run_tiled(nullptr, 0, s));
  return rc;
}

int Sim::substep(hipStream_t s) {
  // timing k: substep i is timed when i % k == (i / k) % k, so the timed
  // substeps visit every position of a k-substep cycle (the sort interval)
  const int64_t k = opt_.timing;
  timing_now_ = k > 0 && timed_subs_ % k == (timed_subs_ / k) % k;
  if (k > 0) ++timed_subs_;
  // every path but the fused tiled one reads the F planes
  if (!(tiled() && opt_.fused)) {
    int rc = ensure_f(s);
    if (rc) return rc;
  }
  const ParticlesDev P0 = particles();
  if (tiled()) {
    if (!opt_.fused) {
      // reference order: P2G, grid update, G2P
      int rc = tiled_p2g(s);
      if (rc) return rc;
      PHASE(PH_GRID_UPDATE, rc = reduce_tiles(d_grid_[0], 1, s));
      if (rc) return rc;
      PHASE(PH_G2P, HIPCHK(mpmk::g2p(particles(), d_mats_, k_, d_grid_[0], s), "g2p"));
      vel_grid_ = d_grid_[0];
      grid_dead_ = true;
      primed_ = false;
      ++steps_since_sort_;
      return MPM_OK;
    }
    if (!primed_) {
      int rc = tiled_p2g(s);
      if (rc) return rc;
      primed_ = true;
    }
    // the tiles hold the P2G of this substep (items of the current order)
    {
      int rc = MPM_OK;
      PHASE(PH_GRID_UPDATE, rc = reduce_tiles(d_grid_[0], 1, s));
      if (rc) return rc;
    }
    if (!items_valid_ || (opt_.sort_every > 0 && steps_since_sort_ >= opt_.sort_every)) {
      int rc = sort_particles(s);
      if (rc) return rc;
    }
    int rc = MPM_OK;
    PHASE(PH_G2P2G, rc = run_tiled(d_grid_[0], 1, s));
    if (rc) return rc;
    vel_grid_ = d_grid_[0];
    grid_dead_ = true;
    ++steps_since_sort_;
    return MPM_OK;
  }
  // direct P2G (global atomics), two grids
  side_clean_ = false;
  grid_dead_ = false;
  float4* gA = d_grid_[gcur_];
  float4* gB = d_grid_[gcur_ ^ 1];
  if (!opt_.fused) {
    // reference order: clear, P2G, grid update, G2P
    PHASE(PH_CLEAR, HIPCHK(mpmk::fill_zero(gA, grid_nodes_, s), "clear grid"));
    PHASE(PH_P2G, HIPCHK(mpmk::p2g_direct(P0, d_mats_, k_, gA, s, (int)opt_.interp), "p2g"));
    PHASE(PH_GRID_UPDATE, HIPCHK(mpmk::grid_update(gA, nullptr, k_, s), "grid update"));
    PHASE(PH_G2P, HIPCHK(mpmk::g2p(P0, d_mats_, k_, gA, s, (int)opt_.interp), "g2p"));
    vel_grid_ = gA;
    gcur_ ^= 1;
    primed_ = false;
    return MPM_OK;
  }
  if (!primed_) {
    PHASE(PH_CLEAR, HIPCHK(mpmk::fill_zero(gA, grid_nodes_, s), "clear grid"));
    PHASE(PH_P2G, HIPCHK(mpmk::p2g_direct(P0, d_mats_, k_, gA, s, (int)opt_.interp), "p2g"));
    primed_ = true;
  }
  // gA: P2G of this substep -> velocities; gB cleared in the same pass
  PHASE(PH_GRID_UPDATE, HIPCHK(mpmk::grid_update(gA, gB, k_, s), "grid update"));
  PHASE(PH_G2P2G, HIPCHK(mpmk::g2p2g_direct(P0, d_mats_, k_, gA, gB, s, (int)opt_.interp), "g2p2g"));
  vel_grid_ = gA;
  gcur_ ^= 1;  // gB (primed P2G of the next substep) becomes current
  return MPM_OK;
}

int Sim::check_materials_cover_particles(const char* who) {
  if (P_ > 0 && (int)max_mat_ >= (int)mats_.size())
    return fail(MPM_ERR_STATE, "%s: particles use material %d but only %zu materials are set", who,
                (int)max_mat_, mats_.size());
  return MPM_OK;
}

int Sim::advance(int32_t substeps, hipStream_t s) {
  if (!uploaded_) return fail(MPM_ERR_STATE, "mpm_advance: call mpm_upload first");
  if (substeps < 0) return fail(MPM_ERR_INVALID, "mpm_advance: negative substep count");
  int rc0 = check_materials_cover_particles("mpm_advance");
  if (rc0) return rc0;
  int rc = upload_materials();
  if (rc) return rc;
  // a host-side range over the enqueue of the whole call (the phases inside
  // it open their own, see phase_begin)
  struct Range {
    bool on;
    explicit Range(bool o) : on(o) { if (on) roctxRangePushA("mpm_advance"); }
    ~Range() { if (on) roctxRangePop(); }
  } range(opt_.markers != 0);
  for (int32_t i = 0; i < substeps; ++i) {
    rc = substep(s);
    if (rc) return rc;
    t_ += (double)prm_.dt;  // mpm.cu:328
  }
  return MPM_OK;
}

int Sim::fetch_state(std::vector<float>& planes, std::vector<uint8_t>& mat,
                     std::vector<uint32_t>& id, hipStream_t s) {
  planes.resize((size_t)NUM_PLANES * stride_);
  mat.resize(P_);
  id.resize(P_);
  if (P_ == 0) return MPM_OK;
  int rc = ensure_f(s);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(planes.data(), d_f_[cur_], sizeof(float) * planes.size(), hipMemcpyDeviceToHost, s), "download particles");
  HIPCHK(hipMemcpyAsync(mat.data(), d_mat_[cur_], P_, hipMemcpyDeviceToHost, s), "download material ids");
  HIPCHK(hipMemcpyAsync(id.data(), d_id_[cur_], sizeof(uint32_t) * P_, hipMemcpyDeviceToHost, s), "download ids");
  HIPCHK(hipStreamSynchronize(s), "download sync");
  return MPM_OK;
}

int Sim::download_soa(float* x, float* v, float* C, float* F, float* Jp, uint8_t* mat,
                      hipStream_t s) {
  if (!uploaded_) return fail(MPM_ERR_STATE, "mpm_download_soa: nothing uploaded");
  std::vector<float> pl;
  std::vector<uint8_t> m;
  std::vector<uint32_t> id;
  int rc = fetch_state(pl, m, id, s);
  if (rc) return rc;
  for (int64_t i = 0; i < P_; ++i) {
    const size_t o = id[i];
    if ((int64_t)o >= P_) return fail(MPM_ERR_STATE, "mpm_download_soa: corrupt particle id");
    for (int a = 0; a < 3; ++a) {
      if (x) x[3 * o + a] = pl[(PL_X + a) * stride_ + i];
      if (v) v[3 * o + a] = pl[(PL_V + a) * stride_ + i];
    }
    for (int a = 0; a < 9; ++a) {
      if (C) C[9 * o + a] = pl[(PL_C + a) * stride_ + i];
      if (F) F[9 * o + a] = pl[(PL_F + a) * stride_ + i];
    }
    if (Jp) Jp[o] = pl[PL_JP * stride_ + i];
    if (mat) mat[o] = m[i];
  }
  return MPM_OK;
}

int Sim::download_aos104(void* out, hipStream_t s) {
  if (!out && P_ > 0) return fail(MPM_ERR_INVALID, "mpm_download_aos104: null output");
  std::vector<float> x(3 * P_), v(3 * P_), C(9 * P_), F(9 * P_), Jp(P_);
  std::vector<uint8_t> mat(P_);
  int rc = download_soa(x.data(), v.data(), C.data(), F.data(), Jp.data(), mat.data(), s);
  if (rc) return rc;
  uint8_t* b = (uint8_t*)out;
  for (int64_t i = 0; i < P_; ++i) {
    uint8_t* r = b + 104 * i;
    std::memset(r, 0, 104);
    r[0] = mat[i];
    std::memcpy(r + 4, &x[3 * i], 12);
    std::memcpy(r + 16, &v[3 * i], 12);
    float fc[9], cc[9];
    for (int row = 0; row < 3; ++row)
      for (int col = 0; col < 3; ++col) {
        fc[3 * col + row] = F[9 * i + 3 * row + col];
        cc[3 * col + row] = C[9 * i + 3 * row + col];
      }
    std::memcpy(r + 28, fc, 36);
    std::memcpy(r + 64, cc, 36);
    std::memcpy(r + 100, &Jp[i], 4);
  }
  return MPM_OK;
}

int Sim::download_grid(float* grid, hipStream_t s) {
  if (!grid) return fail(MPM_ERR_INVALID, "mpm_download_grid: null output");
  if (!vel_grid_) return fail(MPM_ERR_STATE, "mpm_download_grid: no grid update has run yet");
  if (grid_dead_ && vel_grid_ == d_grid_[0]) {
    // node blocks the tiled grid update skipped hold old values: zero them
    HIPCHK(mpmk::zero_dead_blocks(d_grid_[0], d_live_, k_.N, s), "zero dead node blocks");
    grid_dead_ = false;
  }
  HIPCHK(hipMemcpyAsync(grid, vel_grid_, grid_nodes_ * sizeof(float4), hipMemcpyDeviceToHost, s), "download grid");
  HIPCHK(hipStreamSynchronize(s), "download grid sync");
  return MPM_OK;
}

int Sim::debug_p2g(float* grid, hipStream_t s) {
  ++gen_;
  if (!uploaded_) return fail(MPM_ERR_STATE, "mpm_debug_p2g: nothing uploaded");
  if (!grid) return fail(MPM_ERR_INVALID, "mpm_debug_p2g: null output");
  int rc = check_materials_cover_particles("mpm_debug_p2g");
  if (rc) return rc;
  rc = upload_materials();
  if (rc) return rc;
  rc = ensure_f(s);
  if (rc) return rc;
  float4* g = nullptr;
  if (tiled()) {
    // fresh sort and P2G into the tiles; the raw sums land in d_grid_[0]
    items_valid_ = false;
    rc = tiled_p2g(s);
    if (rc) return rc;
    g = d_grid_[0];
    rc = reduce_tiles(g, 0, s);
    if (rc) return rc;
  } else {
    // the non-current grid, so a primed direct-path grid survives
    g = d_grid_[gcur_ ^ 1];
    HIPCHK(mpmk::fill_zero(g, grid_nodes_, s), "clear grid");
    HIPCHK(mpmk::p2g_direct(particles(), d_mats_, k_, g, s, (int)opt_.interp), "p2g");
  }
  HIPCHK(hipMemcpyAsync(grid, g, grid_nodes_ * sizeof(float4), hipMemcpyDeviceToHost, s), "download grid");
  HIPCHK(hipStreamSynchronize(s), "p2g sync");
  if (tiled()) primed_ = false;  // the reduction consumed the side grid
  vel_grid_ = nullptr;
  return MPM_OK;
}

int Sim::debug_sort(uint32_t* keys, uint32_t* perm, hipStream_t s) {
  if (!uploaded_) return fail(MPM_ERR_STATE, "mpm_debug_sort: nothing uploaded");
  if (P_ == 0) return MPM_OK;
  if (!keys || !perm) return fail(MPM_ERR_INVALID, "mpm_debug_sort: null output");
  const ParticlesDev P = particles();
  HIPCHK(mpmk::cell_keys(P, k_, lb_, d_keys_[0], d_vals_[0], s), "cell keys");
  int out = 0;
  HIPCHK(mpmk::radix_sort(d_keys_, d_vals_, P_, 0, mpmk::cell_key_bits(k_.N, lb_), (int)opt_.radix_bits, d_hist_, &out, s), "radix sort");
  std::vector<uint32_t> vals(P_), ids(P_);
  HIPCHK(hipMemcpyAsync(keys, d_keys_[out], sizeof(uint32_t) * P_, hipMemcpyDeviceToHost, s), "download keys");
  HIPCHK(hipMemcpyAsync(vals.data(), d_vals_[out], sizeof(uint32_t) * P_, hipMemcpyDeviceToHost, s), "download perm");
  HIPCHK(hipMemcpyAsync(ids.data(), d_id_[cur_], sizeof(uint32_t) * P_, hipMemcpyDeviceToHost, s), "download ids");
  HIPCHK(hipStreamSynchronize(s), "sort sync");
  for (int64_t i = 0; i < P_; ++i) perm[i] = ids[vals[i]];
  return MPM_OK;
}

// ---- checkpoint: magic, version, params, materials, count, time, SoA arrays
namespace {
const char kMagic[8] = {'M', 'P', 'M', 'S', 'T', 'A', 'T', '1'};
}

int Sim::save_state(const char* path) {
  if (!path) return fail(MPM_ERR_INVALID, "mpm_save_state: null path");
  if (!uploaded_) return fail(MPM_ERR_STATE, "mpm_save_state: nothing uploaded");
  std::vector<float> x(3 * P_), v(3 * P_), C(9 * P_), F(9 * P_), Jp(P_);
  std::vector<uint8_t> mat(P_);
  int rc = download_soa(x.data(), v.data(), C.data(), F.data(), Jp.data(), mat.data(), nullptr);
  if (rc) return rc;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(MPM_ERR_IO, "mpm_save_state: cannot open %s", path);
  const int32_t ver = 1, nm = (int32_t)mats_.size();
  const int64_t n = P_;
  bool ok = std::fwrite(kMagic, 8, 1, f) == 1 && std::fwrite(&ver, 4, 1, f) == 1 &&
            std::fwrite(&prm_, sizeof(prm_), 1, f) == 1 && std::fwrite(&nm, 4, 1, f) == 1 &&
            (nm == 0 || std::fwrite(mats_.data(), sizeof(mpm_material), nm, f) == (size_t)nm) &&
            std::fwrite(&n, 8, 1, f) == 1 && std::fwrite(&t_, 8, 1, f) == 1;
  auto wr = [&](const void* p, size_t bytes) { if (ok && bytes) ok = std::fwrite(p, bytes, 1, f) == 1; };
  wr(x.data(), 4 * x.size()); wr(v.data(), 4 * v.size()); wr(C.data(), 4 * C.size());
  wr(F.data(), 4 * F.size()); wr(Jp.data(), 4 * Jp.size()); wr(mat.data(), mat.size());
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(MPM_ERR_IO, "mpm_save_state: write to %s failed", path);
  return MPM_OK;
}

int Sim::load_state(const char* path) {
  ++gen_;
  if (!path) return fail(MPM_ERR_INVALID, "mpm_load_state: null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(MPM_ERR_IO, "mpm_load_state: cannot open %s", path);
  char magic[8];
  int32_t ver = 0, nm = 0;
  mpm_params p{};
  int64_t n = 0;
  double t = 0;
  bool ok = std::fread(magic, 8, 1, f) == 1 && std::memcmp(magic, kMagic, 8) == 0 &&
            std::fread(&ver, 4, 1, f) == 1 && ver == 1 && std::fread(&p, sizeof(p), 1, f) == 1 &&
            std::fread(&nm, 4, 1, f) == 1 && nm >= 0 && nm <= MPM_MAX_MATERIALS;
  std::vector<mpm_material> mats(ok ? nm : 0);
  if (ok && nm) ok = std::fread(mats.data(), sizeof(mpm_material), nm, f) == (size_t)nm;
  ok = ok && std::fread(&n, 8, 1, f) == 1 && n >= 0 && n < (int64_t)0x7fffff00 && std::fread(&t, 8, 1, f) == 1;
  std::vector<float> x, v, C, F, Jp;
  std::vector<uint8_t> mat;
  if (ok) {
    x.resize(3 * n); v.resize(3 * n); C.resize(9 * n); F.resize(9 * n); Jp.resize(n); mat.resize(n);
  }
  auto rd = [&](void* q, size_t bytes) { if (ok && bytes) ok = std::fread(q, bytes, 1, f) == 1; };
  rd(x.data(), 4 * x.size()); rd(v.data(), 4 * v.size()); rd(C.data(), 4 * C.size());
  rd(F.data(), 4 * F.size()); rd(Jp.data(), 4 * Jp.size()); rd(mat.data(), mat.size());
  std::fclose(f);
  if (!ok) return fail(MPM_ERR_IO, "mpm_load_state: %s is not a valid state file", path);
  std::string why;
  if (!validate_params(p, &why)) return fail(MPM_ERR_IO, "mpm_load_state: bad parameters: %s", why.c_str());
  prm_ = p;
  k_ = make_consts(p);
  int rc = ensure_grids();
  if (rc) return rc;
  rc = set_materials(mats.data(), nm);
  if (rc) return rc;
  clear_staging();
  rc = add_particles_soa(n, x.data(), v.data(), C.data(), F.data(), Jp.data(), mat.data());
  if (rc) return rc;
  rc = upload(nullptr);
  if (rc) return rc;
  t_ = t;
  return MPM_OK;
}

// A copy of the current particle state (planes, materials, ids), the
// simulated time and the F bookkeeping, to try settings from and come back to.
int Sim::snapshot_take(hipStream_t s) {
  snapshot_drop();
  const size_t plane_bytes = sizeof(float) * NUM_PLANES * stride_;
  HIPCHK(hipMalloc(&snap_.f, plane_bytes), "hipMalloc(snapshot)");
  HIPCHK(hipMalloc(&snap_.m, (size_t)std::max<int64_t>(P_, 1)), "hipMalloc(snapshot)");
  HIPCHK(hipMalloc(&snap_.id, sizeof(uint32_t) * (size_t)std::max<int64_t>(P_, 1)), "hipMalloc(snapshot)");
  HIPCHK(hipMemcpyAsync(snap_.f, d_f_[cur_], plane_bytes, hipMemcpyDeviceToDevice, s), "snapshot copy");
  HIPCHK(hipMemcpyAsync(snap_.m, d_mat_[cur_], (size_t)P_, hipMemcpyDeviceToDevice, s), "snapshot copy");
  HIPCHK(hipMemcpyAsync(snap_.id, d_id_[cur_], sizeof(uint32_t) * (size_t)P_, hipMemcpyDeviceToDevice, s),
         "snapshot copy");
  snap_.t = t_;
  snap_.canonical = f_canonical_;
  snap_.opt = opt_;
  return MPM_OK;
}

// Puts the snapshot's particles back (into copy 0) and restarts the step
// bookkeeping from them: the next substep sorts and primes. Options are left
// as they are.
hipError_t Sim::snapshot_restore(hipStream_t s) {
  if (!snap_.f) return hipErrorInvalidValue;
  const size_t plane_bytes = sizeof(float) * NUM_PLANES * stride_;
  hipError_t r = hipMemcpyAsync(d_f_[0], snap_.f, plane_bytes, hipMemcpyDeviceToDevice, s);
  if (r == hipSuccess) r = hipMemcpyAsync(d_mat_[0], snap_.m, (size_t)P_, hipMemcpyDeviceToDevice, s);
  if (r == hipSuccess) r = hipMemcpyAsync(d_id_[0], snap_.id, sizeof(uint32_t) * (size_t)P_, hipMemcpyDeviceToDevice, s);
  cur_ = 0;
  items_valid_ = false;
  primed_ = false;
  pending_gather_ = false;
  steps_since_sort_ = 1 << 30;
  side_clean_ = false;
  vel_grid_ = nullptr;
  f_stale_ = false;
  f_canonical_ = snap_.canonical;
  t_ = snap_.t;
  ++gen_;
  return r;
}

void Sim::snapshot_drop() {
  if (snap_.f) (void)hipFree(snap_.f);
  if (snap_.m) (void)hipFree(snap_.m);
  if (snap_.id) (void)hipFree(snap_.id);
  snap_.f = nullptr;
  snap_.m = nullptr;
  snap_.id = nullptr;
}

int Sim::autotune(int32_t trial_substeps, std::string* report, hipStream_t s) {
  ++gen_;
  if (!uploaded_ || P_ <= 0) return fail(MPM_ERR_STATE, "mpm_autotune: nothing uploaded");
  if (opt_.interp != MPM_INTERP_QUADRATIC)
    return fail(MPM_ERR_STATE, "mpm_autotune: tunes the tiled path, which implements the quadratic kernel only");
  if (trial_substeps < 1 || trial_substeps > 100000)
    return fail(MPM_ERR_INVALID, "mpm_autotune: trial_substeps must be 1..100000");
  int rc = check_materials_cover_particles("mpm_autotune");
  if (rc) return rc;
  rc = upload_materials();
  if (rc) return rc;
  rc = ensure_f(s);
  if (rc) return rc;
  // snapshot of the current particle copy
  rc = snapshot_take(s);
  if (rc) return rc;
  hipError_t e = hipSuccess;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const Options saved = opt_;
  auto restore = [&]() -> hipError_t { return snapshot_restore(s); };
  struct Trial { int64_t sort_every, chunk, radix_bits; double us; };
  std::vector<Trial> trials;
  auto run_trial = [&](int64_t se, int64_t ch, int64_t rb) -> bool {
    if (e != hipSuccess) return false;
    e = restore();
    if (e != hipSuccess) return false;
    opt_ = saved;
    opt_.p2g_mode = 1;
    opt_.fused = 1;
    opt_.timing = 0;
    opt_.sort_every = se;
    opt_.chunk = ch;
    opt_.radix_bits = rb;
    const int32_t n = std::max<int32_t>(trial_substeps, (int32_t)(2 * se));
    rc = advance((int32_t)se, s);  // one sort interval to the steady state
    if (!rc) e = hipEventRecord(e0, s);
    if (!rc && e == hipSuccess) rc = advance(n, s);
    if (!rc && e == hipSuccess) e = hipEventRecord(e1, s);
    if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (rc || e != hipSuccess) return false;
    trials.push_back({se, ch, rb, 1e3 * (double)ms / n});
    return true;
  };
  const int64_t sorts[] = {8, 12, 16, 24, 32, 48, 64, 96};
  const int64_t chunks[] = {mpmk::kMaxItemParticles, 128};
  bool ok = true;
  for (int64_t ch : chunks)
    for (int64_t se : sorts)
      if (ok) ok = run_trial(se, ch, 8);
  // then the radix digit width (sort bin width) at the best interval and
  // item size: a stable sort gives the same order with any width
  if (ok) {
    size_t b = 0;
    for (size_t i = 1; i < trials.size(); ++i)
      if (trials[i].us < trials[b].us) b = i;
    const Trial t = trials[b];
    for (int64_t rb : {6, 4})
      if (ok) ok = run_trial(t.sort_every, t.chunk, rb);
  }
  const hipError_t er = restore();
  if (e == hipSuccess) e = er;
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  snapshot_drop();
  opt_ = saved;
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "mpm_autotune");
  size_t best = 0;
  for (size_t i = 1; i < trials.size(); ++i)
    if (trials[i].us < trials[best].us) best = i;
  opt_.p2g_mode = 1;
  opt_.fused = 1;
  opt_.sort_every = trials[best].sort_every;
  opt_.chunk = trials[best].chunk == mpmk::kMaxItemParticles ? 0 : trials[best].chunk;
  opt_.radix_bits = trials[best].radix_bits;
  if (report) {
    char buf[200];
    *report = "{\"sort_every\": " + std::to_string(opt_.sort_every) + ", \"chunk\": " +
              std::to_string(trials[best].chunk) + ", \"radix_bits\": " + std::to_string(opt_.radix_bits) +
              ", \"trials\": [";
    for (size_t i = 0; i < trials.size(); ++i) {
      std::snprintf(buf, sizeof(buf),
                    "%s{\"sort_every\": %lld, \"chunk\": %lld, \"radix_bits\": %lld, \"us_per_substep\": %.3f}",
                    i ? ", " : "", (long long)trials[i].sort_every, (long long)trials[i].chunk,
                    (long long)trials[i].radix_bits, trials[i].us);
      *report += buf;
    }
    *report += "]}";
  }
  return MPM_OK;
}

int Sim::set_option(const char* name, int64_t v) {
  ++gen_;
  if (!name) return fail(MPM_ERR_INVALID, "mpm_set_option: null name");
  const std::string n = name;
  if (n == "fused") { if (v != 0 && v != 1) return fail(MPM_ERR_INVALID, "fused must be 0 or 1"); opt_.fused = v; primed_ = false; return MPM_OK; }
  if (n == "p2g_mode") {
    if (v != 0 && v != 1) return fail(MPM_ERR_INVALID, "p2g_mode must be 0 or 1");
    if (v != opt_.p2g_mode) primed_ = false;
    opt_.p2g_mode = v;
    return MPM_OK;
  }
  if (n == "interpolation") {
    if (v != MPM_INTERP_QUADRATIC && v != MPM_INTERP_CUBIC)
      return fail(MPM_ERR_INVALID, "interpolation must be %d (quadratic) or %d (cubic)", MPM_INTERP_QUADRATIC,
                  MPM_INTERP_CUBIC);
    if (v != opt_.interp) primed_ = false;
    opt_.interp = v;
    return MPM_OK;
  }
  if (n == "sort_every") { if (v < 0) return fail(MPM_ERR_INVALID, "sort_every must be >= 0"); opt_.sort_every = v; return MPM_OK; }
  if (n == "derive_f") { if (v != 0 && v != 1) return fail(MPM_ERR_INVALID, "derive_f must be 0 or 1"); opt_.derive_f = v; return MPM_OK; }
  if (n == "reduce_mlp") { if (v < -1 || v > 1) return fail(MPM_ERR_INVALID, "reduce_mlp must be -1, 0 or 1"); opt_.reduce_mlp = v; return MPM_OK; }
  if (n == "combine") { if (v < -1 || v > 1) return fail(MPM_ERR_INVALID, "combine must be -1, 0 or 1"); opt_.combine = v; return MPM_OK; }
  if (n == "xcd_map") { if (v != 0 && v != 1) return fail(MPM_ERR_INVALID, "xcd_map must be 0 or 1"); opt_.xcd_map = v; return MPM_OK; }
  if (n == "markers") { if (v != 0 && v != 1) return fail(MPM_ERR_INVALID, "markers must be 0 or 1"); opt_.markers = v; return MPM_OK; }
  if (n == "radix_bits") { if (v != 4 && v != 6 && v != 8) return fail(MPM_ERR_INVALID, "radix_bits must be 4, 6 or 8"); opt_.radix_bits = v; return MPM_OK; }
  if (n == "chunk") {
    if (v < 0 || v > mpmk::kMaxItemParticles || v % 64 != 0)
      return fail(MPM_ERR_INVALID, "chunk must be 0 (auto = %d) or a multiple of 64 up to %d",
                  mpmk::kMaxItemParticles, mpmk::kMaxItemParticles);
    opt_.chunk = v;
    items_valid_ = false;
    return MPM_OK;
  }
  if (n == "block") {
    if (v != 0 && v != 4) return fail(MPM_ERR_INVALID, "block must be 0 (default) or 4");
    opt_.block = v;
    return MPM_OK;
  }
  if (n == "timing") {
    if (v < 0 || v > 4096) return fail(MPM_ERR_INVALID, "timing must be 0 (off) or 1..4096 (one substep in k)");
    opt_.timing = v;
    timed_subs_ = 0;
    return MPM_OK;
  }
  return fail(MPM_ERR_INVALID, "mpm_set_option: unknown option '%s'", name);
}

int Sim::get_option(const char* name, int64_t* v) {
  if (!name || !v) return fail(MPM_ERR_INVALID, "mpm_get_option: null argument");
  const std::string n = name;
  if (n == "fused") *v = opt_.fused;
  else if (n == "p2g_mode") *v = opt_.p2g_mode;
  else if (n == "interpolation") *v = opt_.interp;
  else if (n == "sort_every") *v = opt_.sort_every;
  else if (n == "radix_bits") *v = opt_.radix_bits;
  else if (n == "chunk") *v = opt_.chunk > 0 ? opt_.chunk : mpmk::kMaxItemParticles;
  else if (n == "block") *v = 1 << lb_;
  else if (n == "timing") *v = opt_.timing;
  else if (n == "derive_f") *v = opt_.derive_f;
  else if (n == "xcd_map") *v = opt_.xcd_map;
  else if (n == "markers") *v = opt_.markers;
  else if (n == "reduce_mlp") *v = opt_.reduce_mlp;
  else if (n == "combine") *v = opt_.combine;
  else return fail(MPM_ERR_INVALID, "mpm_get_option: unknown option '%s'", name);
  return MPM_OK;
}

}  // namespace mpm_engine
