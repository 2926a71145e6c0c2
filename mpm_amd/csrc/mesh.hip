// mesh.hip -- surface mesh of the particles on the GPU: the reference's
// MeshBuilder::computeMesh (include/mesh_builder.h:181-211) with its
// distance field (fillVoxelGrid_distance, mesh_builder.h:103-149).
//
// 1. Distance field on a side^3 voxel grid (voxel_dx = N dx / side): every
//    voxel starts at the cut-off (r + 1) voxel_dx and takes the minimum of
//    |x_p - x_node| over the particles whose box of voxels, truncated as the
//    reference truncates it, reaches it. One thread per particle; the minimum
//    is an integer atomic on the float bits (all distances are >= 0), so the
//    field does not depend on the order of the particles.
// 2. The iso-surface S = r voxel_dx - d = 0 (S in FP64 as the reference
//    forms it) by marching cubes. The case table is generated at compile time
//    from the cube's faces: on each face, seen from outside, a segment runs
//    from every edge crossing where the corners go outside -> inside (counter-
//    clockwise) to the next crossing; on a face with four crossings this cuts
//    off each inside corner, and the neighbouring cube, which sees the face
//    from the other side, makes the same cut, so the surface is closed. The
//    segments chain into loops around the inside corners, and each loop is
//    cut into a fan from the first of its vertices whose diagonals do not
//    lie in a face of the cube (a diagonal in a face could meet the
//    neighbour's). Vertices lie on grid edges, linearly interpolated from the
//    edge's lower-index end, so the cubes that share an edge give the same
//    bits; the triangles wind with their normals out of the inside (S > 0).
//    The reference runs libigl's marching cubes (its own case table) on the
//    same field.
// Two passes over the cubes: count the triangles, scan, write them.
#include <hip/hip_runtime.h>

#include "engine_types.hpp"
#include "kernels.hpp"
#include "mc_table.hpp"

namespace {

__global__ void __launch_bounds__(256)
    k_sdf_scatter(ParticlesDev P, int side, float vdx, float cutoff, uint32_t* __restrict__ sdf) {
#pragma clang fp contract(off)  // the reference's separate roundings, no fused multiply-adds
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  float x[3];
  int beg[3], end[3];
  const float rg = cutoff / vdx;  // r_gridpoints, mesh_builder.h:124
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    x[a] = P.f[(PL_X + a) * P.stride + p];
    const float xg = x[a] / vdx;
    // (x - r).cast<int>() and (x + r + 1).cast<int>(): truncation, mesh_builder.h:128-129
    beg[a] = max(0, (int)(xg - rg));
    end[a] = min(side, (int)(xg + (rg + 1.0f)));
  }
  // |x - x_node| = sqrt(dx^2 + (dy^2 + dz^2)): Eigen's norm() sums the three
  // squares of a fixed-size vector as e0 + (e1 + e2); x_node = i * voxel_dx
  for (int i = beg[0]; i < end[0]; ++i) {
    const float dx = x[0] - (float)(uint32_t)i * vdx;
    const float dx2 = dx * dx;
    // rows and planes whose nearest voxel is at the cut-off or beyond change
    // nothing: rounding is monotonic, so d^2 >= fl(dx^2 + dy^2) >= dx^2 in
    // float as well
    if (sqrtf(dx2) >= cutoff) continue;
    for (int j = beg[1]; j < end[1]; ++j) {
      const float dy = x[1] - (float)(uint32_t)j * vdx;
      const float dy2 = dy * dy;
      if (sqrtf(dx2 + dy2) >= cutoff) continue;
      const size_t row = ((size_t)i * side + (size_t)j) * side;
      for (int k = beg[2]; k < end[2]; ++k) {
        const float dz = x[2] - (float)(uint32_t)k * vdx;
        const float d = sqrtf(dx2 + (dy2 + dz * dz));
        const uint32_t bits = __float_as_uint(d);
        uint32_t* const cell = sdf + row + (size_t)k;
        if (bits < *cell) atomicMin(cell, bits);
      }
    }
  }
}

// ---- the case table, generated at compile time (mc_table.hpp) ----
__constant__ mc::Table kMc = mc::make_table();
__constant__ int kMcEdge[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                                   {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};

struct Cube {
  double S[8];
  uint32_t node[8];
  int i, j, k;
};

// the cube's corners; returns the case (bit q: corner q inside, S > 0)
__device__ __forceinline__ int load_cube(const uint32_t* __restrict__ sdf, int side, double rv, int64_t cube,
                                         Cube& c) {
  const int s1 = side - 1;
  c.k = (int)(cube % s1);
  c.j = (int)((cube / s1) % s1);
  c.i = (int)(cube / ((int64_t)s1 * s1));
  int m = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint32_t n = ((uint32_t)(c.i + ((q >> 2) & 1)) * (uint32_t)side + (uint32_t)(c.j + ((q >> 1) & 1))) *
                           (uint32_t)side +
                       (uint32_t)(c.k + (q & 1));
    c.node[q] = n;
    c.S[q] = rv - (double)__uint_as_float(sdf[n]);
    m |= (c.S[q] > 0.0 ? 1 : 0) << q;
  }
  return m;
}

__global__ void __launch_bounds__(256)
    k_mc_count(const uint32_t* __restrict__ sdf, int side, double rv, uint32_t* __restrict__ count) {
  const int64_t s1 = side - 1, ncubes = s1 * s1 * s1;
  const int64_t cube = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cube >= ncubes) return;
  Cube c;
  count[cube] = kMc.ntri[load_cube(sdf, side, rv, cube, c)];
}

// grid position of node n
__device__ __forceinline__ void node_pos(uint32_t n, int side, double p[3]) {
  p[2] = (double)(n % (uint32_t)side);
  p[1] = (double)((n / (uint32_t)side) % (uint32_t)side);
  p[0] = (double)(n / ((uint32_t)side * (uint32_t)side));
}

// the iso-point on the edge between corners qa and qb, from the lower node
__device__ __forceinline__ void edge_point(const Cube& c, int qa, int qb, int side, double vdx, float* out,
                                           uint64_t* key) {
#pragma clang fp contract(off)  // one rounding per operation, as the test's restatement forms it
  if (c.node[qb] < c.node[qa]) { const int t = qa; qa = qb; qb = t; }
  const double Sa = c.S[qa], Sb = c.S[qb];
  const double t = Sa / (Sa - Sb);
  double pa[3], pb[3];
  node_pos(c.node[qa], side, pa);
  node_pos(c.node[qb], side, pb);
#pragma unroll
  for (int a = 0; a < 3; ++a) out[a] = (float)((pa[a] + t * (pb[a] - pa[a])) * vdx);
  *key = ((uint64_t)c.node[qa] << 32) | (uint64_t)c.node[qb];
}

__global__ void __launch_bounds__(256)
    k_mc_emit(const uint32_t* __restrict__ sdf, int side, double rv, double vdx, const uint32_t* __restrict__ offset,
              float* __restrict__ xyz, uint64_t* __restrict__ keys) {
  const int64_t s1 = side - 1, ncubes = s1 * s1 * s1;
  const int64_t cube = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cube >= ncubes) return;
  Cube c;
  const int m = load_cube(sdf, side, rv, cube, c);
  const int nt = kMc.ntri[m];
  const uint32_t at = offset[cube];
  for (int t = 0; t < nt; ++t)
    for (int q = 0; q < 3; ++q) {
      const int e = kMc.tri[m][3 * t + q];
      float v[3];
      uint64_t key;
      edge_point(c, kMcEdge[e][0], kMcEdge[e][1], side, vdx, v, &key);
      const size_t o = (size_t)(at + (uint32_t)t) * 3 + (size_t)q;
#pragma unroll
      for (int a = 0; a < 3; ++a) xyz[3 * o + a] = v[a];
      if (keys) keys[o] = key;
    }
}

}  // namespace

namespace mpmk {

hipError_t mesh_distance_field(const ParticlesDev& P, int side, float vdx, float cutoff, uint32_t* sdf,
                               hipStream_t s) {
  const int64_t nv = (int64_t)side * side * side;
  hipError_t e = fill_u32(sdf, nv, __builtin_bit_cast(uint32_t, cutoff), s);
  if (e != hipSuccess || P.n <= 0) return e;
  hipLaunchKernelGGL(k_sdf_scatter, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, s, P, side, vdx, cutoff,
                     sdf);
  return hipGetLastError();
}

hipError_t mesh_count(const uint32_t* sdf, int side, double rv, uint32_t* count, hipStream_t s) {
  const int64_t s1 = side - 1, ncubes = s1 * s1 * s1;
  hipLaunchKernelGGL(k_mc_count, dim3((unsigned)((ncubes + 255) / 256)), dim3(256), 0, s, sdf, side, rv, count);
  return hipGetLastError();
}

hipError_t mesh_emit(const uint32_t* sdf, int side, double rv, double vdx, const uint32_t* offset, float* xyz,
                     uint64_t* keys, hipStream_t s) {
  const int64_t s1 = side - 1, ncubes = s1 * s1 * s1;
  hipLaunchKernelGGL(k_mc_emit, dim3((unsigned)((ncubes + 255) / 256)), dim3(256), 0, s, sdf, side, rv, vdx, offset,
                     xyz, keys);
  return hipGetLastError();
}

}  // namespace mpmk
