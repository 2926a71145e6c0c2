// material.cpp -- host-only parts of the C-ABI that need no GPU: the error
// message of calls without a handle, and material construction. Kept apart
// from capi.cpp so that host-only builds (the sanitizer test of the scene
// input path) link the same code.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/mpm_hip.h"
#include "host_error.hpp"

namespace mpm_host {

namespace {
thread_local std::string g_last_error;
}  // namespace

void set_global_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

const char* global_error() { return g_last_error.c_str(); }

}  // namespace mpm_host

using mpm_host::set_global_error;

// ---------------------------------------------------------------------------
// material construction (MaterialModel.cuh:27-29,50-54,75-82; the TOML values
// are doubles converted to the constructors' float parameters, main.cu:36-42)
// ---------------------------------------------------------------------------
extern "C" int mpm_material_init(mpm_material* m, int32_t kind, double volume, double density,
                                 double E, double nu, double hardening, double clamp_lo,
                                 double clamp_hi, double bulk_modulus) {
  if (!m) {
    set_global_error("mpm_material_init: null material");
    return MPM_ERR_INVALID;
  }
  if (kind < MPM_MATERIAL_SNOW || kind > MPM_MATERIAL_WC_LIQUID) {
    set_global_error("mpm_material_init: unknown material kind %d", kind);
    return MPM_ERR_INVALID;
  }
  const float fv = (float)volume, fd = (float)density, fE = (float)E, fnu = (float)nu;
  std::memset(m, 0, sizeof(*m));
  m->kind = kind;
  m->particle_volume = fv;
  m->particle_mass = fd * fv;
  m->mu0 = fE / (2.0f * (1.0f + fnu));
  m->lambda0 = fE * fnu / ((1.0f + fnu) * (1.0f - 2.0f * fnu));
  m->hardening = (float)hardening;
  m->clamp_lo = (float)clamp_lo;
  m->clamp_hi = (float)clamp_hi;
  m->bulk_modulus = (float)bulk_modulus;
  return MPM_OK;
}
