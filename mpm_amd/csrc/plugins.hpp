// plugins.hpp -- the reference's compile-time plugin surface as device policy
// templates for gfx950.
//
// The reference binds its physics at compile time (include/mpm.cuh:24-27):
//   InterpolationKernel = QuadraticInterpolationKernel
//   TransferScheme      = MLS_APIC_Scheme<InterpolationKernel>
//   MaterialModel       = MMSnow<Particle>
// The material model is a runtime kind here (mpm_material.kind, dispatched in
// device_math.hpp). The other two are the policies below: the direct kernels
// of substep.hip are templates over a TransferScheme policy, which is a
// template over an InterpolationKernel policy.
//
// InterpolationKernel policy (InterpolationKernelBase<N, D_is_const>,
// include/InterpolationKernel.cuh:12-53):
//   static constexpr int size()          nodes per direction
//   static constexpr bool d_is_const()   D^-1 is the same for every particle
//   weights_per_direction(x, dx_inv, base, w)
//                                        first node per axis and the
//                                        weights w[axis][node]
//   D_inv_const(dx_inv)                  the constant D^-1 = c I (returns c),
//                                        used when d_is_const()
// and the generic D^-1 of InterpolationKernel.cuh:21-50 is d_inv_generic<K>
// below, used when !d_is_const().
//
// TransferScheme policy (TransferSchemeBase + MLS_APIC_Scheme,
// include/TransferScheme.h:33-143): per particle, after prepare():
//   base[3], w[3][S]                     the stencil
//   affine(tau, C, material)             p2g_prepare_particle's affine
//                                        momentum matrix (TransferScheme.h:83-85)
//   node_dinv_row(d, out)                d^T D^-1 of a node offset d, the G2P
//                                        affine factor (TransferScheme.h:131)
// The node loops themselves (p2g_node_contribution, g2p_node_contribution,
// g2p_finish_particle) are written once in substep.hip over any policy.
//
// Bound instances: MlsApicScheme<QuadraticKernel> (the reference's) for every
// path; MlsApicScheme<CubicKernel> (no reference counterpart, the example of
// a second kernel through the generic D^-1) for the direct path. The tiled
// fused path (p2g_block.hip) sizes its tiles for the quadratic stencil and
// implements that instance only.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "engine_types.hpp"

namespace mpmplug {

// QuadraticInterpolationKernel (InterpolationKernel.cuh:55-74)
struct QuadraticKernel {
  static constexpr int size() { return 3; }
  static constexpr bool d_is_const() { return true; }
  __device__ static __forceinline__ void weights_per_direction(const float x[3], float dx_inv, int base[3],
                                                               float w[3][3]) {
    mpmdev::quad_weights(x, dx_inv, base, w);
  }
  // Identity * 4 * dx_inv * dx_inv (InterpolationKernel.cuh:71-73)
  __device__ static __forceinline__ float D_inv_const(float dx_inv) { return (4.0f * dx_inv) * dx_inv; }
};

// Cubic B-spline N(t), t = |node - particle| / dx: (1/2) t^3 - t^2 + 2/3 on
// [0, 1), (2 - t)^3 / 6 on [1, 2), 0 beyond.
__device__ __forceinline__ float cubic_bspline(float t) {
  const float t2 = t * t;
  const float inner = ((0.5f * t2) * t - t2) + (2.0f / 3.0f);
  const float u = 2.0f - t;
  const float outer = ((u * u) * u) * (1.0f / 6.0f);
  return t < 1.0f ? inner : (t < 2.0f ? outer : 0.0f);
}

// A cubic B-spline kernel: 4 nodes per direction from base = int(x/dx - 1)
// (truncated toward zero like the quadratic kernel's .cast<int>()); its D^-1
// goes through the generic inverse (d_is_const() false).
struct CubicKernel {
  static constexpr int size() { return 4; }
  static constexpr bool d_is_const() { return false; }
  __device__ static __forceinline__ void weights_per_direction(const float x[3], float dx_inv, int base[3],
                                                               float w[3][4]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float xg = x[a] * dx_inv;
      base[a] = (int)(xg - 1.0f);
      const float fx = xg - (float)base[a];
#pragma unroll
      for (int o = 0; o < 4; ++o) w[a][o] = cubic_bspline(fabsf(fx - (float)o));
    }
  }
  __device__ static __forceinline__ float D_inv_const(float) { return 0.0f; }  // not constant
};

// InterpolationKernelBase::D_inv (InterpolationKernel.cuh:21-50): D = sum
// over the stencil of w_i w_j w_k (x_node - x_p)(x_node - x_p)^T, inverted by
// its adjugate. Row-major Dinv[9].
template <class K>
__device__ __forceinline__ void d_inv_generic(const float x[3], const int base[3],
                                              const float w[3][K::size()], float dx, float Dinv[9]) {
  constexpr int S = K::size();
  float D[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) D[q] = 0.0f;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const float d0 = (float)(base[0] + i) * dx - x[0];
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const float d1 = (float)(base[1] + j) * dx - x[1];
      const float wij = w[0][i] * w[1][j];
#pragma unroll
      for (int kk = 0; kk < S; ++kk) {
        const float d2 = (float)(base[2] + kk) * dx - x[2];
        const float wt = wij * w[2][kk];
        const float d[3] = {d0, d1, d2};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const float wd = wt * d[r];
#pragma unroll
          for (int c = 0; c < 3; ++c) D[3 * r + c] += wd * d[c];
        }
      }
    }
  }
  float cof[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      cof[3 * r + c] = D[3 * r1 + c1] * D[3 * r2 + c2] - D[3 * r1 + c2] * D[3 * r2 + c1];
    }
  const float det = cof[0] * D[0] + (cof[3] * D[3] + cof[6] * D[6]);
  const float inv = 1.0f / det;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Dinv[3 * r + c] = cof[3 * c + r] * inv;
}

// MLS_APIC_Scheme<InterpolationKernel> (TransferScheme.h:57-143).
template <class K>
struct MlsApicScheme {
  using Kernel = K;
  static constexpr int S = K::size();
  int base[3];
  float w[3][S];
  float Dinv[K::d_is_const() ? 1 : 9];  // the particle's D^-1 (generic kernels only)

  // weights and base (weights_per_direction), then D^-1 when it is not
  // constant; false when no node of the stencil is in the grid (the
  // reference skips such particles, mpm.cu:31-35,129-133)
  __device__ __forceinline__ bool prepare(const float x[3], const StepConsts& k) {
    K::weights_per_direction(x, k.dx_inv, base, w);
    const int N = k.N;
    if (base[0] + S < 0 || base[0] >= N || base[1] + S < 0 || base[1] >= N || base[2] + S < 0 ||
        base[2] >= N)
      return false;
    if constexpr (!K::d_is_const()) d_inv_generic<K>(x, base, w, k.dx, Dinv);
    return true;
  }

  // affine = -Dinv dt V tau + m C (p2g_prepare_particle, TransferScheme.h:83-85).
  // Constant D^-1: m.stress_scale holds ((-Dinv) dt) V for the quadratic
  // kernel's constant (engine.cpp make_material_dev).
  __device__ __forceinline__ void affine(const float tau[9], const float C[9], const MaterialDev& m,
                                         const StepConsts& k, float aff[9]) const {
    if constexpr (K::d_is_const()) {
#pragma unroll
      for (int a = 0; a < 9; ++a) aff[a] = m.stress_scale * tau[a] + m.mass * C[a];
    } else {
      float sD[9];
#pragma unroll
      for (int a = 0; a < 9; ++a) sD[a] = ((-Dinv[a]) * k.dt) * m.volume;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          aff[3 * r + c] = (sD[3 * r] * tau[c] + (sD[3 * r + 1] * tau[3 + c] + sD[3 * r + 2] * tau[6 + c])) +
                           m.mass * C[3 * r + c];
    }
  }

  // (d^T D^-1)_c of a particle-to-node offset d (the G2P affine factor of
  // g2p_node_contribution, TransferScheme.h:131), for the three axes'
  // offsets d[a][o] of stencil node o
  __device__ __forceinline__ void dinv_rows(const float d0, const float d1, const float d2, const StepConsts& k,
                                            float out[3]) const {
    if constexpr (K::d_is_const()) {
      out[0] = d0 * k.dinv;
      out[1] = d1 * k.dinv;
      out[2] = d2 * k.dinv;
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) out[c] = d0 * Dinv[c] + (d1 * Dinv[3 + c] + d2 * Dinv[6 + c]);
    }
  }
};

using ReferenceScheme = MlsApicScheme<QuadraticKernel>;  // mpm.cuh:24-27
using CubicScheme = MlsApicScheme<CubicKernel>;

}  // namespace mpmplug
