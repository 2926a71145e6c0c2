// linalg.hip -- device 3x3 SVD over a batch of matrices, for the linear
// algebra parity tests (the reference's own device tests run single matrices
// through the same routine, tests/test_linalg.cu:57-75).
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "kernels.hpp"

__global__ void __launch_bounds__(256)
    k_svd_batch(const float* __restrict__ A, float* __restrict__ U, float* __restrict__ S,
                float* __restrict__ V, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float a[9], u[9], s[3], v[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) a[t] = A[9 * i + t];
  mpmdev::svd3_robust(a, u, s, v);
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    U[9 * i + t] = u[t];
    V[9 * i + t] = v[t];
  }
#pragma unroll
  for (int t = 0; t < 3; ++t) S[3 * i + t] = s[t];
}

namespace mpmk {
// Whether the device code object of this library is loadable on the current
// device (no launch): attributes of the SVD kernel.
hipError_t probe_kernel_image(int* nregs) {
  hipFuncAttributes a;
  const hipError_t e = hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_svd_batch));
  if (e == hipSuccess && nregs) *nregs = a.numRegs;
  return e;
}

hipError_t svd_batch(const float* A, float* U, float* S, float* V, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_svd_batch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A, U, S, V, n);
  return hipGetLastError();
}
}  // namespace mpmk
