// mesh_post.hpp -- host-side mesh steps and file writers (mesh_post.cpp).
// Vertices are double xyz triples, faces int32 vertex-index triples.
#pragma once

#include <cstdint>
#include <string>

namespace mpm_mesh {

bool check_faces(int64_t nv, int64_t nf, const int32_t* F, std::string* why);
// one implicit cotangent-Laplacian step and the area normalisation, in place
int smooth(int64_t nv, double* V, int64_t nf, const int32_t* F, std::string* why);
// shortest-edge collapses to at most max_faces faces; U holds up to nv
// vertices, G up to nf faces
int decimate(int64_t nv, const double* V, int64_t nf, const int32_t* F, int64_t max_faces, double* U,
             int64_t* nu, int32_t* G, int64_t* ng, std::string* why);
bool write_obj(const char* path, int64_t nv, const double* V, int64_t nf, const int32_t* F);
// particles (x, v: n xyz triples) as a gzip-compressed Houdini .bgeo
bool write_bgeo(const char* path, int64_t n, const float* x, const float* v);

}  // namespace mpm_mesh
