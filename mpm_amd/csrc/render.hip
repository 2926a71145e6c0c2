// render.hip -- the reference's debug view (include/renderer.h:34-59,
// src/renderer.cpp:80-107, shaders/cube_vertex.glsl, constant_fragment.glsl)
// without a window: the device particles drawn as points into an RGB8 image
// on the GPU.
//
// The reference's view: each particle at clip position ((x - 0.5) 1.8,
// (y - 0.5) 1.8, (z - 0.5) 1.8, 1), GL_POINTS of size 2, one colour
// (0.59, 0.54, 0.2), no depth test, over a clear colour of 0.1 grey, in an
// 800 x 600 viewport. Here: a point outside the clip cube is dropped; window
// position xw = (xn + 1) W / 2, yw = (yn + 1) H / 2; the point covers the
// pixels whose centres lie in (xw - 1, xw + 1] x (yw - 1, yw + 1]; colours are
// rounded to 8 bits, round(255 c); image rows run top (yf = H - 1) to bottom.
// All points have one colour, so the image does not depend on their order.
#include <hip/hip_runtime.h>

#include "engine_types.hpp"
#include "kernels.hpp"

namespace {

constexpr uint8_t kBack = 26;                        // round(255 * 0.1)
constexpr uint8_t kPoint[3] = {150, 138, 51};        // round(255 * (0.59, 0.54, 0.2))

__global__ void __launch_bounds__(256) k_clear_rgb(uint8_t* __restrict__ img, size_t bytes) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < bytes) img[t] = kBack;
}

__global__ void __launch_bounds__(256) k_points(ParticlesDev P, int W, int H, uint8_t* __restrict__ img) {
#pragma clang fp contract(off)
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  float c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) c[a] = (P.f[(PL_X + a) * P.stride + p] - 0.5f) * 1.8f;
  if (!(c[0] >= -1.0f && c[0] <= 1.0f && c[1] >= -1.0f && c[1] <= 1.0f && c[2] >= -1.0f && c[2] <= 1.0f)) return;
  const float xw = (c[0] + 1.0f) * (0.5f * (float)W), yw = (c[1] + 1.0f) * (0.5f * (float)H);
  // pixel centres f + 0.5 in (w - 1, w + 1]: f = floor(w - 1.5) + 1 .. floor(w + 0.5),
  // two pixels per axis
  const int x0 = (int)floorf(xw - 1.5f) + 1, x1 = (int)floorf(xw + 0.5f);
  const int y0 = (int)floorf(yw - 1.5f) + 1, y1 = (int)floorf(yw + 0.5f);
  for (int y = max(y0, 0); y <= min(y1, H - 1); ++y)
    for (int x = max(x0, 0); x <= min(x1, W - 1); ++x) {
      uint8_t* px = img + 3 * ((size_t)(H - 1 - y) * (size_t)W + (size_t)x);
      px[0] = kPoint[0];
      px[1] = kPoint[1];
      px[2] = kPoint[2];
    }
}

}  // namespace

namespace mpmk {

hipError_t render_points(const ParticlesDev& P, int W, int H, uint8_t* img, hipStream_t s) {
  const size_t bytes = (size_t)3 * (size_t)W * (size_t)H;
  hipLaunchKernelGGL(k_clear_rgb, dim3((unsigned)((bytes + 255) / 256)), dim3(256), 0, s, img, bytes);
  if (P.n > 0)
    hipLaunchKernelGGL(k_points, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, s, P, W, H, img);
  return hipGetLastError();
}

}  // namespace mpmk
