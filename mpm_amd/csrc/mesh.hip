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
//    forms it) by marching tetrahedra: each voxel cube is cut into the 6
//    tetrahedra around its diagonal from the low to the high corner, which
//    neighbouring cubes split their shared faces the same way, so the surface
//    is closed without a case table. Vertices lie on grid edges, linearly
//    interpolated from the edge's lower-index end, so the two cubes that share
//    an edge give the same bits; each triangle is turned so that its normal
//    points from the inside (S > 0) out. The reference runs libigl's marching
//    cubes on the same field (more triangles here, the same surface).
// Two passes over the cubes: count the triangles, scan, write them.
#include <hip/hip_runtime.h>

#include "engine_types.hpp"
#include "kernels.hpp"

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

// corner c of a cube = (c >> 2 & 1, c >> 1 & 1, c & 1) along (x, y, z);
// tetrahedron t = {0, e_a, e_a + e_b, 7} for the axis order (a, b) of row t
__constant__ int kTet[6][4] = {{0, 4, 6, 7}, {0, 4, 5, 7}, {0, 2, 6, 7},
                               {0, 2, 3, 7}, {0, 1, 5, 7}, {0, 1, 3, 7}};

struct Cube {
  double S[8];
  uint32_t node[8];
  int i, j, k;
};

__device__ __forceinline__ bool load_cube(const uint32_t* __restrict__ sdf, int side, double rv,
                                          int64_t cube, Cube& c) {
  const int s1 = side - 1;
  c.k = (int)(cube % s1);
  c.j = (int)((cube / s1) % s1);
  c.i = (int)(cube / ((int64_t)s1 * s1));
  bool in = false, out = false;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint32_t n = ((uint32_t)(c.i + ((q >> 2) & 1)) * (uint32_t)side + (uint32_t)(c.j + ((q >> 1) & 1))) *
                           (uint32_t)side +
                       (uint32_t)(c.k + (q & 1));
    c.node[q] = n;
    c.S[q] = rv - (double)__uint_as_float(sdf[n]);
    in = in || c.S[q] > 0.0;
    out = out || !(c.S[q] > 0.0);
  }
  return in && out;
}

__device__ __forceinline__ int tet_triangles(const Cube& c, int t) {
  int m = 0;
#pragma unroll
  for (int v = 0; v < 4; ++v) m += c.S[kTet[t][v]] > 0.0 ? 1 : 0;
  return (m == 1 || m == 3) ? 1 : (m == 2 ? 2 : 0);
}

__global__ void __launch_bounds__(256)
    k_mt_count(const uint32_t* __restrict__ sdf, int side, double rv, uint32_t* __restrict__ count) {
  const int64_t s1 = side - 1, ncubes = s1 * s1 * s1;
  const int64_t cube = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cube >= ncubes) return;
  Cube c;
  uint32_t n = 0;
  if (load_cube(sdf, side, rv, cube, c))
    for (int t = 0; t < 6; ++t) n += (uint32_t)tet_triangles(c, t);
  count[cube] = n;
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

__device__ __forceinline__ void put_triangle(const Cube& c, const int e[3][2], const double in_c[3],
                                             const double out_c[3], int side, double vdx,
                                             float* __restrict__ xyz, uint64_t* __restrict__ keys, uint32_t at) {
  float v[3][3];
  uint64_t k[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) edge_point(c, e[q][0], e[q][1], side, vdx, v[q], &k[q]);
  // normal from inside to outside
  const double u0 = (double)v[1][0] - v[0][0], u1 = (double)v[1][1] - v[0][1], u2 = (double)v[1][2] - v[0][2];
  const double w0 = (double)v[2][0] - v[0][0], w1 = (double)v[2][1] - v[0][1], w2 = (double)v[2][2] - v[0][2];
  const double n0 = u1 * w2 - u2 * w1, n1 = u2 * w0 - u0 * w2, n2 = u0 * w1 - u1 * w0;
  const double dot = n0 * (out_c[0] - in_c[0]) + n1 * (out_c[1] - in_c[1]) + n2 * (out_c[2] - in_c[2]);
  const int order[3] = {0, dot < 0.0 ? 2 : 1, dot < 0.0 ? 1 : 2};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
#pragma unroll
    for (int a = 0; a < 3; ++a) xyz[(size_t)at * 9 + 3 * q + a] = v[order[q]][a];
    if (keys) keys[(size_t)at * 3 + q] = k[order[q]];
  }
}

__global__ void __launch_bounds__(256)
    k_mt_emit(const uint32_t* __restrict__ sdf, int side, double rv, double vdx, const uint32_t* __restrict__ offset,
              float* __restrict__ xyz, uint64_t* __restrict__ keys) {
  const int64_t s1 = side - 1, ncubes = s1 * s1 * s1;
  const int64_t cube = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cube >= ncubes) return;
  Cube c;
  if (!load_cube(sdf, side, rv, cube, c)) return;
  uint32_t at = offset[cube];
  for (int t = 0; t < 6; ++t) {
    int ins[4], outs[4], ni = 0, no = 0;
    double in_c[3] = {0.0, 0.0, 0.0}, out_c[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < 4; ++v) {
      const int q = kTet[t][v];
      double p[3];
      node_pos(c.node[q], side, p);
      if (c.S[q] > 0.0) {
        ins[ni++] = q;
        for (int a = 0; a < 3; ++a) in_c[a] += p[a];
      } else {
        outs[no++] = q;
        for (int a = 0; a < 3; ++a) out_c[a] += p[a];
      }
    }
    if (ni == 0 || no == 0) continue;
    for (int a = 0; a < 3; ++a) {
      in_c[a] /= ni;
      out_c[a] /= no;
    }
    if (ni == 1 || no == 1) {
      const int lone = ni == 1 ? ins[0] : outs[0];
      const int* other = ni == 1 ? outs : ins;
      const int e[3][2] = {{lone, other[0]}, {lone, other[1]}, {lone, other[2]}};
      put_triangle(c, e, in_c, out_c, side, vdx, xyz, keys, at++);
    } else {  // 2 in, 2 out: the quad (a c, a d, b d, b c) as two triangles
      const int a = ins[0], b = ins[1], cc = outs[0], d = outs[1];
      const int e1[3][2] = {{a, cc}, {a, d}, {b, d}};
      const int e2[3][2] = {{a, cc}, {b, d}, {b, cc}};
      put_triangle(c, e1, in_c, out_c, side, vdx, xyz, keys, at++);
      put_triangle(c, e2, in_c, out_c, side, vdx, xyz, keys, at++);
    }
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
  hipLaunchKernelGGL(k_mt_count, dim3((unsigned)((ncubes + 255) / 256)), dim3(256), 0, s, sdf, side, rv, count);
  return hipGetLastError();
}

hipError_t mesh_emit(const uint32_t* sdf, int side, double rv, double vdx, const uint32_t* offset, float* xyz,
                     uint64_t* keys, hipStream_t s) {
  const int64_t s1 = side - 1, ncubes = s1 * s1 * s1;
  hipLaunchKernelGGL(k_mt_emit, dim3((unsigned)((ncubes + 255) / 256)), dim3(256), 0, s, sdf, side, rv, vdx, offset,
                     xyz, keys);
  return hipGetLastError();
}

}  // namespace mpmk
