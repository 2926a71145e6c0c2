// capi.cpp -- the C-ABI of libmpm_hip.so (declared in include/mpm_hip.h).
//
// Host side of the engine: owns the device buffers of one simulation, stages
// particles on the host, and enqueues the substep kernels. It stands in for
// the reference's class Simulation (include/mpm.cuh:44-106,
// src/mpm.cu:180-394); each entry point names the method it replaces in the
// header.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpm_hip.h"
#include "engine.hpp"
#include "mc_table.hpp"
#include "mesh_post.hpp"
#include "host_error.hpp"
#include "kernels.hpp"

using mpm_host::set_global_error;
using mpm_engine::Sim;

extern "C" int mpm_abi_version(void) { return MPM_ABI_VERSION; }

extern "C" const char* mpm_last_error(mpm_handle h) {
  if (!h) return mpm_host::global_error();
  return reinterpret_cast<Sim*>(h)->last_error().c_str();
}

extern "C" int mpm_create(const mpm_params* params, mpm_handle* out) {
  if (!params || !out) {
    set_global_error("mpm_create: null argument");
    return MPM_ERR_INVALID;
  }
  *out = nullptr;
  std::string why;
  if (!mpm_engine::validate_params(*params, &why)) {
    set_global_error("mpm_create: %s", why.c_str());
    return MPM_ERR_INVALID;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    set_global_error("mpm_create: no HIP device visible");
    return MPM_ERR_NOGPU;
  }
  Sim* s = new Sim(*params);
  const int rc = s->init_device();
  if (rc != MPM_OK) {
    set_global_error("mpm_create: %s", s->last_error().c_str());
    delete s;
    return rc;
  }
  *out = reinterpret_cast<mpm_handle>(s);
  return MPM_OK;
}

extern "C" int mpm_destroy(mpm_handle h) {
  delete reinterpret_cast<Sim*>(h);
  return MPM_OK;
}

#define MPM_GET(h)                                             \
  if (!h) {                                                    \
    set_global_error("%s: null handle", __func__);             \
    return MPM_ERR_INVALID;                                    \
  }                                                            \
  Sim* s = reinterpret_cast<Sim*>(h);                          \
  std::lock_guard<std::mutex> lock_(s->mutex());

extern "C" int mpm_set_materials(mpm_handle h, const mpm_material* mats, int32_t n) {
  MPM_GET(h);
  return s->set_materials(mats, n);
}

extern "C" int mpm_add_particles_soa(mpm_handle h, int64_t n, const float* x, const float* v,
                                     const float* C, const float* F, const float* Jp,
                                     const uint8_t* material) {
  MPM_GET(h);
  return s->add_particles_soa(n, x, v, C, F, Jp, material);
}

extern "C" int mpm_add_particles_aos104(mpm_handle h, const void* particles, int64_t n) {
  MPM_GET(h);
  return s->add_particles_aos104(particles, n);
}

extern "C" int mpm_clear_particles(mpm_handle h) {
  MPM_GET(h);
  return s->clear_staging();
}

extern "C" int mpm_upload(mpm_handle h, void* stream) {
  MPM_GET(h);
  return s->upload((hipStream_t)stream);
}

extern "C" int mpm_append_particles(mpm_handle h, void* stream) {
  MPM_GET(h);
  return s->append((hipStream_t)stream);
}

extern "C" int mpm_mesh_build(mpm_handle h, int32_t grid_side, int32_t particle_radius, int64_t* triangles,
                              void* stream) {
  MPM_GET(h);
  return s->mesh_build(grid_side, particle_radius, triangles, (hipStream_t)stream);
}

extern "C" int mpm_mesh_triangles(mpm_handle h, float* xyz, uint64_t* edges) {
  MPM_GET(h);
  return s->mesh_triangles(xyz, edges);
}

extern "C" int mpm_mesh_distance_field(mpm_handle h, int32_t grid_side, float* out) {
  MPM_GET(h);
  return s->mesh_distance(grid_side, out);
}

extern "C" int mpm_mesh_indexed(mpm_handle h, int64_t* nv, int64_t* nf, double* V, int32_t* F) {
  MPM_GET(h);
  if (!nv || !nf) return s->fail(MPM_ERR_INVALID, "mpm_mesh_indexed: null count");
  return s->mesh_indexed(nv, nf, V, F);
}

extern "C" int mpm_render_frame(mpm_handle h, int32_t width, int32_t height, uint8_t* rgb, void* stream) {
  MPM_GET(h);
  return s->render(width, height, rgb, (hipStream_t)stream);
}

extern "C" int mpm_mesh_case_table(uint8_t* ntri, int8_t* tri) {
  if (!ntri || !tri) {
    set_global_error("mpm_mesh_case_table: null argument");
    return MPM_ERR_INVALID;
  }
  static constexpr mc::Table t = mc::make_table();
  std::memcpy(ntri, t.ntri, sizeof(t.ntri));
  std::memcpy(tri, t.tri, sizeof(t.tri));
  return MPM_OK;
}

namespace {
int mesh_args(const char* fn, int64_t nv, const void* V, int64_t nf, const int32_t* F) {
  std::string why;
  if (nv < 0 || nf < 0 || nv > INT32_MAX || (nv > 0 && !V) || (nf > 0 && !F)) {
    set_global_error("%s: bad mesh arguments", fn);
    return MPM_ERR_INVALID;
  }
  if (!mpm_mesh::check_faces(nv, nf, F, &why)) {
    set_global_error("%s: %s", fn, why.c_str());
    return MPM_ERR_INVALID;
  }
  return MPM_OK;
}
}  // namespace

extern "C" int mpm_mesh_smooth(int64_t nv, double* V, int64_t nf, const int32_t* F) {
  int rc = mesh_args(__func__, nv, V, nf, F);
  if (rc) return rc;
  std::string why;
  if (mpm_mesh::smooth(nv, V, nf, F, &why) != 0) {
    set_global_error("%s", why.c_str());
    return MPM_ERR_INVALID;
  }
  return MPM_OK;
}

extern "C" int mpm_mesh_decimate(int64_t nv, const double* V, int64_t nf, const int32_t* F, int64_t max_faces,
                                 double* U, int64_t* nu, int32_t* G, int64_t* ng) {
  int rc = mesh_args(__func__, nv, V, nf, F);
  if (rc) return rc;
  if (max_faces < 0 || !nu || !ng || (nv > 0 && !U) || (nf > 0 && !G)) {
    set_global_error("mpm_mesh_decimate: need max_faces >= 0 and output buffers");
    return MPM_ERR_INVALID;
  }
  std::string why;
  if (mpm_mesh::decimate(nv, V, nf, F, max_faces, U, nu, G, ng, &why) != 0) {
    set_global_error("mpm_mesh_decimate: %s", why.c_str());
    return MPM_ERR_INVALID;
  }
  return MPM_OK;
}

extern "C" int mpm_obj_write(const char* path, int64_t nv, const double* V, int64_t nf, const int32_t* F) {
  int rc = mesh_args(__func__, nv, V, nf, F);
  if (rc) return rc;
  if (!path) {
    set_global_error("mpm_obj_write: null path");
    return MPM_ERR_INVALID;
  }
  if (!mpm_mesh::write_obj(path, nv, V, nf, F)) {
    set_global_error("mpm_obj_write: cannot write %s", path);
    return MPM_ERR_IO;
  }
  return MPM_OK;
}

extern "C" int mpm_bgeo_write(const char* path, int64_t n, const float* x, const float* v) {
  if (!path || n < 0 || n > INT32_MAX || (n > 0 && (!x || !v))) {
    set_global_error("mpm_bgeo_write: bad arguments");
    return MPM_ERR_INVALID;
  }
  if (!mpm_mesh::write_bgeo(path, n, x, v)) {
    set_global_error("mpm_bgeo_write: cannot write %s", path);
    return MPM_ERR_IO;
  }
  return MPM_OK;
}

extern "C" int mpm_mesh_write_obj(mpm_handle h, const char* path) {
  MPM_GET(h);
  return s->mesh_write_obj(path);
}

extern "C" int mpm_remove_particles(mpm_handle h, int64_t first, int64_t count, void* removed, void* stream) {
  MPM_GET(h);
  return s->remove(first, count, removed, (hipStream_t)stream);
}

extern "C" int mpm_count(mpm_handle h, int64_t* out) {
  MPM_GET(h);
  if (!out) return s->fail(MPM_ERR_INVALID, "mpm_count: null output");
  *out = s->count();
  return MPM_OK;
}

extern "C" int mpm_advance(mpm_handle h, int32_t substeps, void* stream) {
  MPM_GET(h);
  return s->advance(substeps, (hipStream_t)stream);
}

extern "C" int mpm_time(mpm_handle h, double* out) {
  MPM_GET(h);
  if (!out) return s->fail(MPM_ERR_INVALID, "mpm_time: null output");
  *out = s->time();
  return MPM_OK;
}

extern "C" int mpm_download_soa(mpm_handle h, float* x, float* v, float* C, float* F, float* Jp,
                                uint8_t* material, void* stream) {
  MPM_GET(h);
  return s->download_soa(x, v, C, F, Jp, material, (hipStream_t)stream);
}

extern "C" int mpm_download_aos104(mpm_handle h, void* particles, void* stream) {
  MPM_GET(h);
  return s->download_aos104(particles, (hipStream_t)stream);
}

extern "C" int mpm_download_grid(mpm_handle h, float* grid, void* stream) {
  MPM_GET(h);
  return s->download_grid(grid, (hipStream_t)stream);
}

extern "C" int mpm_save_state(mpm_handle h, const char* path) {
  MPM_GET(h);
  return s->save_state(path);
}

extern "C" int mpm_load_state(mpm_handle h, const char* path) {
  MPM_GET(h);
  return s->load_state(path);
}

extern "C" int mpm_set_option(mpm_handle h, const char* name, int64_t value) {
  MPM_GET(h);
  return s->set_option(name, value);
}

extern "C" int mpm_get_option(mpm_handle h, const char* name, int64_t* value) {
  MPM_GET(h);
  return s->get_option(name, value);
}

extern "C" int mpm_timing_read(mpm_handle h, const char* phase, double* total_ms, int64_t* launches) {
  MPM_GET(h);
  return s->timing_read(phase, total_ms, launches);
}

extern "C" int mpm_autotune(mpm_handle h, int32_t trial_substeps, char* report, int32_t report_len,
                            void* stream) {
  MPM_GET(h);
  std::string r;
  const int rc = s->autotune(trial_substeps, &r, (hipStream_t)stream);
  if (rc) return rc;
  if (report && report_len > 0) std::snprintf(report, (size_t)report_len, "%s", r.c_str());
  return MPM_OK;
}

extern "C" int mpm_timing_reset(mpm_handle h) {
  MPM_GET(h);
  return s->timing_reset();
}

extern "C" int mpm_debug_p2g(mpm_handle h, float* grid, void* stream) {
  MPM_GET(h);
  return s->debug_p2g(grid, (hipStream_t)stream);
}

extern "C" int mpm_debug_sort(mpm_handle h, uint32_t* keys, uint32_t* perm, void* stream) {
  MPM_GET(h);
  return s->debug_sort(keys, perm, (hipStream_t)stream);
}

extern "C" int mpm_debug_svd(int64_t n, const float* A, float* U, float* sigma, float* V) {
  if (n < 0 || (n > 0 && (!A || !U || !sigma || !V))) {
    set_global_error("mpm_debug_svd: bad arguments");
    return MPM_ERR_INVALID;
  }
  if (n == 0) return MPM_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    set_global_error("mpm_debug_svd: no HIP device visible");
    return MPM_ERR_NOGPU;
  }
  // device layout: A (9n) | U (9n) | sigma (3n) | V (9n)
  float* d = nullptr;
  const size_t bytes = (size_t)n * 30 * sizeof(float);
  hipError_t e = hipMalloc(&d, bytes);
  if (e != hipSuccess) {
    set_global_error("mpm_debug_svd: hipMalloc: %s", hipGetErrorString(e));
    return MPM_ERR_HIP;
  }
  float *dA = d, *dU = d + 9 * n, *dS = d + 18 * n, *dV = d + 21 * n;
  const char* step = "hipMemcpy(host to device)";
  e = hipMemcpy(dA, A, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) { step = "kernel launch"; e = mpmk::svd_batch(dA, dU, dS, dV, n, nullptr); }
  if (e == hipSuccess) { step = "hipMemcpy(device to host)"; e = hipMemcpy(U, dU, (size_t)n * 9 * sizeof(float), hipMemcpyDeviceToHost); }
  if (e == hipSuccess) e = hipMemcpy(sigma, dS, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(V, dV, (size_t)n * 9 * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) {
    set_global_error("mpm_debug_svd: %s: %s", step, hipGetErrorString(e));
    return MPM_ERR_HIP;
  }
  return MPM_OK;
}

extern "C" int mpm_device_info(char* buf, int32_t len) {
  if (!buf || len <= 0) {
    set_global_error("mpm_device_info: null buffer");
    return MPM_ERR_INVALID;
  }
  int rt = 0, drv = 0, ndev = 0, dev = 0;
  (void)hipRuntimeGetVersion(&rt);
  (void)hipDriverGetVersion(&drv);
  const hipError_t ce = hipGetDeviceCount(&ndev);
  std::string out = "hip_runtime=" + std::to_string(rt) + " hip_driver=" + std::to_string(drv) +
                    " devices=" + std::to_string(ndev) + " count_status=" + hipGetErrorString(ce);
  if (ce == hipSuccess && ndev > 0 && hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
      out += std::string(" name=") + p.name + " arch=" + p.gcnArchName +
             " cus=" + std::to_string(p.multiProcessorCount) +
             " lds_per_block=" + std::to_string(p.sharedMemPerBlock);
    }
    int nregs = -1;
    const hipError_t fe = mpmk::probe_kernel_image(&nregs);
    out += std::string(" kernel_image=") + hipGetErrorString(fe) + " svd_kernel_vgprs=" + std::to_string(nregs);
  }
  (void)hipGetLastError();
  std::snprintf(buf, (size_t)len, "%s", out.c_str());
  return (ce == hipSuccess && ndev > 0) ? MPM_OK : MPM_ERR_NOGPU;
}

extern "C" int mpm_debug_selftest(char* buf, int32_t len) {
  if (!buf || len <= 0) {
    set_global_error("mpm_debug_selftest: null buffer");
    return MPM_ERR_INVALID;
  }
  std::string out;
  auto rec = [&](const char* what, hipError_t e) {
    out += std::string(what) + "=" + hipGetErrorName(e) + "; ";
    return e == hipSuccess;
  };
  int ndev = 0;
  if (!rec("count", hipGetDeviceCount(&ndev)) || ndev <= 0) {
    std::snprintf(buf, (size_t)len, "%s", out.c_str());
    return MPM_ERR_NOGPU;
  }
  float host[256], back[256];
  for (int i = 0; i < 256; ++i) host[i] = (float)i + 0.5f;
  float* d = nullptr;
  bool ok = rec("malloc", hipMalloc(&d, 4096));
  if (ok) ok = rec("h2d", hipMemcpy(d, host, sizeof(host), hipMemcpyHostToDevice));
  if (ok) {
    ok = rec("d2h", hipMemcpy(back, d, sizeof(back), hipMemcpyDeviceToHost));
    out += std::string("d2h_match=") + (std::memcmp(host, back, sizeof(host)) == 0 ? "yes" : "no") + "; ";
  }
  if (ok) {
    rec("fill_launch", mpmk::fill_zero((float4*)d, 64, nullptr));
    rec("fill_sync", hipDeviceSynchronize());
    ok = rec("fill_d2h", hipMemcpy(back, d, sizeof(back), hipMemcpyDeviceToHost));
    bool zero = true;
    for (int i = 0; i < 256; ++i) zero = zero && back[i] == 0.0f;
    out += std::string("fill_zero=") + (zero ? "yes" : "no") + "; ";
  }
  if (ok) {
    float A[27] = {1, 2, 1, 1, 3, 1, 1, 8, 1, 0, 1, 0, -1, 2, -1, -1, 0.001f, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
    ok = rec("svd_h2d", hipMemcpy(d, A, sizeof(A), hipMemcpyHostToDevice));
    if (ok) rec("svd_launch", mpmk::svd_batch(d, d + 27, d + 54, d + 63, 3, nullptr));
    rec("svd_sync", hipDeviceSynchronize());
    rec("svd_lasterror", hipGetLastError());
    ok = rec("svd_d2h", hipMemcpy(back, d, 90 * sizeof(float), hipMemcpyDeviceToHost));
    if (ok) {
      char tmp[128];
      std::snprintf(tmp, sizeof(tmp), "sigma0=%g,%g,%g; ", back[54], back[55], back[56]);
      out += tmp;
    }
  }
  float* pinned = nullptr;
  if (rec("hostmalloc", hipHostMalloc((void**)&pinned, 1024, 0))) {
    rec("d2h_pinned", hipMemcpy(pinned, d, 1024, hipMemcpyDeviceToHost));
    (void)hipHostFree(pinned);
  }
  if (d) rec("free", hipFree(d));
  (void)hipGetLastError();
  std::snprintf(buf, (size_t)len, "%s", out.c_str());
  return MPM_OK;
}

// ---------------------------------------------------------------------------
// differentiable substep (adjoint.hip)
// ---------------------------------------------------------------------------
namespace {

// grids and tapes are read as float4
bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

int diff_check(const char* fn, const mpm_params* p, float mass, float volume, int64_t n,
               StepConsts* k) {
  if (!p) {
    set_global_error("%s: null params", fn);
    return MPM_ERR_INVALID;
  }
  std::string why;
  if (!mpm_engine::validate_params(*p, &why)) {
    set_global_error("%s: %s", fn, why.c_str());
    return MPM_ERR_INVALID;
  }
  if (p->compat != MPM_COMPAT_FIXED) {
    set_global_error("%s: the differentiable substep implements compat FIXED only", fn);
    return MPM_ERR_INVALID;
  }
  if (n < 0 || !(mass > 0.0f) || !(volume > 0.0f) || !std::isfinite(mass) || !std::isfinite(volume)) {
    set_global_error("%s: need n >= 0 and finite positive particle mass and volume", fn);
    return MPM_ERR_INVALID;
  }
  *k = mpm_engine::make_consts(*p);
  return MPM_OK;
}

}  // namespace

extern "C" int mpm_diff_sizes(const mpm_params* params, int64_t* grid_floats, int64_t* tape_floats) {
  std::string why;
  if (!params || !grid_floats || !tape_floats || !mpm_engine::validate_params(*params, &why)) {
    set_global_error("mpm_diff_sizes: bad arguments %s", why.c_str());
    return MPM_ERR_INVALID;
  }
  const int64_t nodes = (int64_t)params->N * params->N * params->N;
  *grid_floats = 4 * nodes;
  *tape_floats = 8 * nodes;
  return MPM_OK;
}

extern "C" int mpm_diff_substep(const mpm_params* params, float particle_mass, float particle_volume,
                                int64_t n, const float* state, float* state_out, const float* mu,
                                const float* lambda, float* grid, float* tape, void* stream) {
  StepConsts k;
  int rc = diff_check(__func__, params, particle_mass, particle_volume, n, &k);
  if (rc) return rc;
  if (!grid || !tape || (n > 0 && (!state || !state_out || !mu || !lambda))) {
    set_global_error("mpm_diff_substep: null buffer");
    return MPM_ERR_INVALID;
  }
  if (n > 0 && state == state_out) {
    set_global_error("mpm_diff_substep: state_out must not alias state");
    return MPM_ERR_INVALID;
  }
  if (!aligned16(grid) || !aligned16(tape)) {
    set_global_error("mpm_diff_substep: grid and tape must be 16-byte aligned");
    return MPM_ERR_INVALID;
  }
  const hipError_t e = mpmk::diff_forward(k, particle_mass, particle_volume, n, state, state_out, mu,
                                          lambda, (float4*)grid, tape, (hipStream_t)stream);
  if (e != hipSuccess) {
    set_global_error("mpm_diff_substep: %s", hipGetErrorString(e));
    return MPM_ERR_HIP;
  }
  return MPM_OK;
}

extern "C" int mpm_diff_substep_noalloc(const mpm_params* params, float particle_mass,
                                        float particle_volume, int64_t n, const float* state,
                                        float* state_out, const float* mu, const float* lambda,
                                        float* tape, void* stream) {
  StepConsts k;
  int rc = diff_check(__func__, params, particle_mass, particle_volume, n, &k);
  if (rc) return rc;
  if (params->N > (uint32_t)mpmk::kNoAllocMaxN) {
    set_global_error("mpm_diff_substep_noalloc: N = %u exceeds the static scratch grid (N <= %d)",
                     params->N, mpmk::kNoAllocMaxN);
    return MPM_ERR_CAPACITY;
  }
  if (!tape || (n > 0 && (!state || !state_out || !mu || !lambda))) {
    set_global_error("mpm_diff_substep_noalloc: null buffer");
    return MPM_ERR_INVALID;
  }
  if (n > 0 && state == state_out) {
    set_global_error("mpm_diff_substep_noalloc: state_out must not alias state");
    return MPM_ERR_INVALID;
  }
  if (!aligned16(tape)) {
    set_global_error("mpm_diff_substep_noalloc: tape must be 16-byte aligned");
    return MPM_ERR_INVALID;
  }
  const hipError_t e = mpmk::diff_forward_noalloc(k, particle_mass, particle_volume, n, state,
                                                  state_out, mu, lambda, tape, (hipStream_t)stream);
  if (e != hipSuccess) {
    set_global_error("mpm_diff_substep_noalloc: %s", hipGetErrorString(e));
    return MPM_ERR_HIP;
  }
  return MPM_OK;
}

extern "C" int mpm_diff_forward_steps(const mpm_params* params, float particle_mass,
                                      float particle_volume, int64_t n, int32_t steps, const float* state0,
                                      float* states, int64_t state_stride, const float* mu,
                                      const float* lambda, float* grid, float* tapes, int64_t tape_stride,
                                      void* stream) {
  StepConsts k;
  int rc = diff_check(__func__, params, particle_mass, particle_volume, n, &k);
  if (rc) return rc;
  const int64_t nodes = (int64_t)params->N * params->N * params->N;
  if (steps < 1 || state_stride < 24 * n || (tape_stride != 0 && (tape_stride < 8 * nodes || tape_stride % 4))) {
    set_global_error("mpm_diff_forward_steps: need steps >= 1, state_stride >= 24 n and tape_stride 0 or a "
                     "multiple of 4 >= 8 N^3");
    return MPM_ERR_INVALID;
  }
  if (!grid || !tapes || (n > 0 && (!state0 || !states || !mu || !lambda))) {
    set_global_error("mpm_diff_forward_steps: null buffer");
    return MPM_ERR_INVALID;
  }
  if (n > 0 && state0 >= states && state0 < states + (size_t)steps * (size_t)state_stride) {
    set_global_error("mpm_diff_forward_steps: state0 must not lie in the output states");
    return MPM_ERR_INVALID;
  }
  if (!aligned16(grid) || !aligned16(tapes)) {
    set_global_error("mpm_diff_forward_steps: grid and tapes must be 16-byte aligned");
    return MPM_ERR_INVALID;
  }
  const hipError_t e = mpmk::diff_forward_steps(k, particle_mass, particle_volume, n, steps, state0, states,
                                                state_stride, mu, lambda, (float4*)grid, tapes, tape_stride,
                                                (hipStream_t)stream);
  if (e != hipSuccess) {
    set_global_error("mpm_diff_forward_steps: %s", hipGetErrorString(e));
    return MPM_ERR_HIP;
  }
  return MPM_OK;
}

extern "C" int mpm_diff_backward_steps(const mpm_params* params, float particle_mass,
                                       float particle_volume, int64_t n, int32_t steps, const float* states,
                                       int64_t state_stride, const float* mu, const float* lambda,
                                       const float* tapes, int64_t tape_stride, const float* adj_out,
                                       float* adj_in, float* adj_mu, float* adj_lambda, float* grid,
                                       float* work, void* stream) {
  StepConsts k;
  int rc = diff_check(__func__, params, particle_mass, particle_volume, n, &k);
  if (rc) return rc;
  const int64_t nodes = (int64_t)params->N * params->N * params->N;
  if (steps < 1 || state_stride < 24 * n || (steps > 1 && (tape_stride < 8 * nodes || tape_stride % 4))) {
    set_global_error("mpm_diff_backward_steps: need steps >= 1, state_stride >= 24 n and tape_stride a "
                     "multiple of 4 >= 8 N^3");
    return MPM_ERR_INVALID;
  }
  if (!grid || !tapes ||
      (n > 0 && (!states || !mu || !lambda || !adj_out || !adj_in || !adj_mu || !adj_lambda || !work))) {
    set_global_error("mpm_diff_backward_steps: null buffer");
    return MPM_ERR_INVALID;
  }
  if (n > 0 && (adj_in == adj_out || work == adj_out)) {
    set_global_error("mpm_diff_backward_steps: adj_in and work must not alias adj_out");
    return MPM_ERR_INVALID;
  }
  if (!aligned16(grid) || !aligned16(tapes)) {
    set_global_error("mpm_diff_backward_steps: grid and tapes must be 16-byte aligned");
    return MPM_ERR_INVALID;
  }
  const hipError_t e = mpmk::diff_backward_steps(k, particle_mass, particle_volume, n, steps, states,
                                                 state_stride, mu, lambda, tapes, tape_stride, adj_out,
                                                 adj_in, adj_mu, adj_lambda, (float4*)grid, work,
                                                 (hipStream_t)stream);
  if (e != hipSuccess) {
    set_global_error("mpm_diff_backward_steps: %s", hipGetErrorString(e));
    return MPM_ERR_HIP;
  }
  return MPM_OK;
}

extern "C" int mpm_diff_substep_backward(const mpm_params* params, float particle_mass,
                                         float particle_volume, int64_t n, const float* state,
                                         const float* mu, const float* lambda, const float* tape,
                                         const float* adj_out, float* adj_in, float* adj_mu,
                                         float* adj_lambda, float* grid, void* stream) {
  StepConsts k;
  int rc = diff_check(__func__, params, particle_mass, particle_volume, n, &k);
  if (rc) return rc;
  if (!grid || !tape ||
      (n > 0 && (!state || !mu || !lambda || !adj_out || !adj_in || !adj_mu || !adj_lambda))) {
    set_global_error("mpm_diff_substep_backward: null buffer");
    return MPM_ERR_INVALID;
  }
  if (n > 0 && adj_in == adj_out) {
    set_global_error("mpm_diff_substep_backward: adj_in must not alias adj_out");
    return MPM_ERR_INVALID;
  }
  if (!aligned16(grid) || !aligned16(tape)) {
    set_global_error("mpm_diff_substep_backward: grid and tape must be 16-byte aligned");
    return MPM_ERR_INVALID;
  }
  const hipError_t e = mpmk::diff_backward(k, particle_mass, particle_volume, n, state, mu, lambda,
                                           tape, adj_out, adj_in, adj_mu, adj_lambda, (float4*)grid,
                                           (hipStream_t)stream);
  if (e != hipSuccess) {
    set_global_error("mpm_diff_substep_backward: %s", hipGetErrorString(e));
    return MPM_ERR_HIP;
  }
  return MPM_OK;
}
