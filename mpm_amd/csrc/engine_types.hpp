// engine_types.hpp -- types shared by the host side (capi.cpp) and the HIP
// kernels: the device particle layout, per-material constants and per-step
// constants.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../include/mpm_hip.h"

// ---------------------------------------------------------------------------
// Particle state in HBM: structure of arrays, one float plane per scalar
// component, each plane `stride` floats long (stride is a multiple of 64 so
// every plane starts 256-byte aligned and a wavefront's 64 lanes read one
// full, aligned 256-byte segment per plane).
// ---------------------------------------------------------------------------
enum Plane : int {
  PL_X = 0,    // x0 x1 x2
  PL_V = 3,    // v0 v1 v2
  PL_C = 6,    // C row-major, 9 planes
  PL_F = 15,   // F row-major, 9 planes
  PL_JP = 24,  // Jp
  NUM_PLANES = 25
};

struct ParticlesDev {
  float* f;          // NUM_PLANES * stride floats
  uint8_t* mat;      // material index
  uint32_t* id;      // index in upload order (the reference's particle index)
  size_t stride;
  int32_t n;
};

struct MaterialDev {
  int32_t kind;
  float volume, mass, mu0, lambda0, hardening, clamp_lo, clamp_hi, bulk;
  float stress_scale;  // -Dinv * dt * volume: the stress-to-affine factor
};

#define MPM_MAX_MATERIALS 64

struct StepConsts {
  float dt, dx, dx_inv, dinv, gravity, bfrac;
  int32_t N;
  int32_t boundary_mode;
  int32_t bcells;
  int32_t compat;
  uint32_t guard;  // particles with id > guard skip G2P (compat REFERENCE: N^3)
};
