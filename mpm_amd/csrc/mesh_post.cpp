// mesh_post.cpp -- host-side steps after the GPU mesh and the particle file
// writer: what the reference's MeshBuilder does after marching cubes
// (include/mesh_builder.h:196-246: optional Laplacian smoothing, optional
// decimation, OBJ output) and its ParticleWriter (particle_writer.h:13-44,
// Partio's .bgeo). The reference uses libigl and Partio, neither of which is
// a dependency here; these are restatements of the operations they perform, on
// plain arrays: vertices as double xyz triples, faces as int32 vertex-index
// triples.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <queue>
#include <string>
#include <vector>

#include "mesh_post.hpp"

namespace mpm_mesh {

namespace {

double tri_double_area(const double* a, const double* b, const double* c) {
  const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
  return std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
}

double dist2(const double* a, const double* b) {
  const double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// sparse symmetric matrix in compressed rows
struct Csr {
  std::vector<int64_t> row;
  std::vector<int32_t> col;
  std::vector<double> val;
  void mul(const std::vector<double>& x, std::vector<double>& y) const {
    const int64_t n = (int64_t)row.size() - 1;
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t q = row[(size_t)i]; q < row[(size_t)i + 1]; ++q) s += val[(size_t)q] * x[(size_t)col[(size_t)q]];
      y[(size_t)i] = s;
    }
  }
};

}  // namespace

bool check_faces(int64_t nv, int64_t nf, const int32_t* F, std::string* why) {
  for (int64_t t = 0; t < 3 * nf; ++t)
    if (F[t] < 0 || F[t] >= nv) {
      if (why) *why = "face index out of range";
      return false;
    }
  return true;
}

// MeshBuilder::smoothMesh (mesh_builder.h:215-246): one implicit step
// (M - 0.001 L) U = M V with L the cotangent Laplacian of (V, F)
// (igl::cotmatrix: L_jk = (cot a + cot b) / 2 over the two angles opposite
// edge jk, L_jj = -sum_k L_jk) and M the barycentric mass matrix of (V, F)
// (igl::massmatrix BARYCENTRIC: M_ii = area of the faces at i / 3), then
// U /= sqrt(area of (U, F)). The reference factors the system (sparse
// Cholesky); here preconditioned conjugate gradients (the matrix is
// symmetric positive definite for a mesh without degenerate obtuse fans) to a
// relative residual of 1e-10.
int smooth(int64_t nv, double* V, int64_t nf, const int32_t* F, std::string* why) {
  if (nv <= 0 || nf <= 0) return 0;
  // cotangent weights per face: 0.5 cot(angle opposite edge) = (b^2 + c^2 - a^2) / (4 dblA)
  std::vector<double> mass((size_t)nv, 0.0);
  std::vector<std::vector<std::pair<int32_t, double>>> off((size_t)nv);
  auto add = [&](int32_t i, int32_t j, double w) {
    for (auto& e : off[(size_t)i])
      if (e.first == j) {
        e.second += w;
        return;
      }
    off[(size_t)i].push_back({j, w});
  };
  for (int64_t t = 0; t < nf; ++t) {
    const int32_t v[3] = {F[3 * t], F[3 * t + 1], F[3 * t + 2]};
    const double* p[3] = {V + 3 * (size_t)v[0], V + 3 * (size_t)v[1], V + 3 * (size_t)v[2]};
    const double dA = tri_double_area(p[0], p[1], p[2]);
    for (int q = 0; q < 3; ++q) mass[(size_t)v[q]] += dA / 6.0;
    if (!(dA > 0.0)) continue;  // a degenerate face adds no cotangent weight
    // edge q is opposite corner q: between corners q+1 and q+2
    double l2[3];
    for (int q = 0; q < 3; ++q) l2[q] = dist2(p[(q + 1) % 3], p[(q + 2) % 3]);
    for (int q = 0; q < 3; ++q) {
      const double w = (l2[(q + 1) % 3] + l2[(q + 2) % 3] - l2[q]) / (4.0 * dA);  // cot / 2
      const int32_t a = v[(q + 1) % 3], b = v[(q + 2) % 3];
      add(a, b, w);
      add(b, a, w);
    }
  }
  // S = M - 0.001 L: off-diagonals -0.001 L_jk, diagonal M_jj + 0.001 sum_k L_jk
  Csr S;
  S.row.assign((size_t)nv + 1, 0);
  std::vector<double> diag((size_t)nv);
  for (int64_t i = 0; i < nv; ++i) {
    auto& r = off[(size_t)i];
    std::sort(r.begin(), r.end());
    double sum = 0.0;
    for (const auto& e : r) sum += e.second;
    diag[(size_t)i] = mass[(size_t)i] + 0.001 * sum;
    S.row[(size_t)i + 1] = S.row[(size_t)i] + (int64_t)r.size() + 1;
  }
  S.col.resize((size_t)S.row.back());
  S.val.resize((size_t)S.row.back());
  for (int64_t i = 0; i < nv; ++i) {
    int64_t q = S.row[(size_t)i];
    bool placed = false;
    for (const auto& e : off[(size_t)i]) {
      if (!placed && e.first > i) {
        S.col[(size_t)q] = (int32_t)i;
        S.val[(size_t)q++] = diag[(size_t)i];
        placed = true;
      }
      S.col[(size_t)q] = e.first;
      S.val[(size_t)q++] = -0.001 * e.second;
    }
    if (!placed) {
      S.col[(size_t)q] = (int32_t)i;
      S.val[(size_t)q++] = diag[(size_t)i];
    }
  }
  for (int64_t i = 0; i < nv; ++i)
    if (!(diag[(size_t)i] > 0.0)) {
      if (why) *why = "mesh_smooth: a vertex has no area (degenerate mesh)";
      return -1;
    }
  // three right-hand sides M V, solved column by column
  const size_t n = (size_t)nv;
  std::vector<double> b(n), x(n), r(n), z(n), p(n), Ap(n);
  for (int c = 0; c < 3; ++c) {
    double bn = 0.0;
    for (size_t i = 0; i < n; ++i) {
      b[i] = mass[i] * V[3 * i + c];
      x[i] = V[3 * i + c];  // start from the unsmoothed positions
      bn += b[i] * b[i];
    }
    bn = std::sqrt(bn);
    S.mul(x, Ap);
    double rz = 0.0, rn = 0.0;
    for (size_t i = 0; i < n; ++i) {
      r[i] = b[i] - Ap[i];
      z[i] = r[i] / diag[i];
      p[i] = z[i];
      rz += r[i] * z[i];
      rn += r[i] * r[i];
    }
    int it = 0;
    const int max_it = 20000;
    while (std::sqrt(rn) > 1e-10 * bn && it < max_it) {
      S.mul(p, Ap);
      double pAp = 0.0;
      for (size_t i = 0; i < n; ++i) pAp += p[i] * Ap[i];
      if (!(pAp > 0.0)) {
        if (why) *why = "mesh_smooth: the system is not positive definite";
        return -1;
      }
      const double alpha = rz / pAp;
      double rz2 = 0.0;
      rn = 0.0;
      for (size_t i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * Ap[i];
        z[i] = r[i] / diag[i];
        rz2 += r[i] * z[i];
        rn += r[i] * r[i];
      }
      const double beta = rz2 / rz;
      rz = rz2;
      for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      ++it;
    }
    if (it >= max_it) {
      if (why) *why = "mesh_smooth: the solve did not converge";
      return -1;
    }
    for (size_t i = 0; i < n; ++i) V[3 * i + c] = x[i];
  }
  double area = 0.0;
  for (int64_t t = 0; t < nf; ++t)
    area += 0.5 * tri_double_area(V + 3 * (size_t)F[3 * t], V + 3 * (size_t)F[3 * t + 1], V + 3 * (size_t)F[3 * t + 2]);
  if (!(area > 0.0)) {
    if (why) *why = "mesh_smooth: the smoothed mesh has no area";
    return -1;
  }
  const double s = std::sqrt(area);
  for (size_t i = 0; i < 3 * n; ++i) V[i] /= s;
  return 0;
}

// igl::decimate(V, F, max_m, U, G, J) with its default cost and placement
// (mesh_builder.h:203-207): edges collapse shortest first, the merged vertex
// at the edge's midpoint, until at most max_faces faces remain or no edge can
// collapse. An edge collapses only if the vertices adjacent to both of its
// ends are exactly the two opposite its faces (the link condition, which
// keeps a closed manifold closed and manifold). Ties go to the edge found
// first. The output keeps the surviving vertices and faces in input order.
int decimate(int64_t nv, const double* V, int64_t nf, const int32_t* F, int64_t max_faces, double* U,
             int64_t* nu, int32_t* G, int64_t* ng, std::string* why) {
  std::vector<double> P(V, V + 3 * (size_t)nv);
  std::vector<int32_t> T(F, F + 3 * (size_t)nf);
  std::vector<char> face_alive((size_t)nf, 1), vert_alive((size_t)nv, 1);
  std::vector<std::vector<int32_t>> vf((size_t)nv);  // faces at each vertex
  for (int64_t t = 0; t < nf; ++t)
    for (int q = 0; q < 3; ++q) vf[(size_t)T[3 * t + q]].push_back((int32_t)t);
  std::vector<uint32_t> stamp((size_t)nv, 0);
  struct Cand {
    double len;
    int64_t order;
    int32_t a, b;
    uint32_t sa, sb;
    bool operator>(const Cand& o) const { return len > o.len || (len == o.len && order > o.order); }
  };
  std::priority_queue<Cand, std::vector<Cand>, std::greater<Cand>> pq;
  int64_t order = 0;
  auto push = [&](int32_t a, int32_t b) {
    if (a > b) std::swap(a, b);
    pq.push({std::sqrt(dist2(&P[3 * (size_t)a], &P[3 * (size_t)b])), order++, a, b, stamp[(size_t)a], stamp[(size_t)b]});
  };
  for (int64_t t = 0; t < nf; ++t)
    for (int q = 0; q < 3; ++q) {
      const int32_t a = T[3 * t + q], b = T[3 * t + (q + 1) % 3];
      if (a < b) push(a, b);  // each interior edge once (its other face lists it b -> a)
    }
  auto neighbours = [&](int32_t v, std::vector<int32_t>& out) {
    out.clear();
    for (int32_t t : vf[(size_t)v])
      if (face_alive[(size_t)t])
        for (int q = 0; q < 3; ++q)
          if (T[3 * (size_t)t + q] != v) out.push_back(T[3 * (size_t)t + q]);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
  };
  int64_t faces = nf;
  std::vector<int32_t> na, nb, shared, opp;
  while (faces > max_faces && !pq.empty()) {
    const Cand c = pq.top();
    pq.pop();
    if (!vert_alive[(size_t)c.a] || !vert_alive[(size_t)c.b] || stamp[(size_t)c.a] != c.sa ||
        stamp[(size_t)c.b] != c.sb)
      continue;  // an older entry: the edge changed or went
    // the faces on the edge and their opposite vertices
    opp.clear();
    std::vector<int32_t> edge_faces;
    for (int32_t t : vf[(size_t)c.a]) {
      if (!face_alive[(size_t)t]) continue;
      const int32_t* f = &T[3 * (size_t)t];
      if (f[0] == c.b || f[1] == c.b || f[2] == c.b) {
        edge_faces.push_back(t);
        for (int q = 0; q < 3; ++q)
          if (f[q] != c.a && f[q] != c.b) opp.push_back(f[q]);
      }
    }
    if (edge_faces.size() != 2) continue;  // not an interior manifold edge
    neighbours(c.a, na);
    neighbours(c.b, nb);
    shared.clear();
    std::set_intersection(na.begin(), na.end(), nb.begin(), nb.end(), std::back_inserter(shared));
    std::sort(opp.begin(), opp.end());
    if (shared != opp) continue;  // the link condition fails
    if (na.size() + nb.size() <= 6) continue;  // the last edge of a tetrahedron
    // collapse b into a at the midpoint
    for (int q = 0; q < 3; ++q) P[3 * (size_t)c.a + q] = 0.5 * (P[3 * (size_t)c.a + q] + P[3 * (size_t)c.b + q]);
    for (int32_t t : edge_faces) face_alive[(size_t)t] = 0;
    faces -= 2;
    for (int32_t t : vf[(size_t)c.b]) {
      if (!face_alive[(size_t)t]) continue;
      for (int q = 0; q < 3; ++q)
        if (T[3 * (size_t)t + q] == c.b) T[3 * (size_t)t + q] = c.a;
      vf[(size_t)c.a].push_back(t);
    }
    vert_alive[(size_t)c.b] = 0;
    vf[(size_t)c.b].clear();
    // the edges at the merged vertex, with their new lengths (the older
    // entries of its edges fail the stamp check)
    ++stamp[(size_t)c.a];
    neighbours(c.a, na);
    for (int32_t w : na) push(c.a, w);
  }
  // compact
  std::vector<int32_t> remap((size_t)nv, -1);
  int64_t k = 0;
  for (int64_t i = 0; i < nv; ++i) {
    if (!vert_alive[(size_t)i]) continue;
    bool used = false;
    for (int32_t t : vf[(size_t)i]) used = used || face_alive[(size_t)t];
    if (!used) continue;
    remap[(size_t)i] = (int32_t)k;
    for (int q = 0; q < 3; ++q) U[3 * k + q] = P[3 * (size_t)i + q];
    ++k;
  }
  int64_t m = 0;
  for (int64_t t = 0; t < nf; ++t) {
    if (!face_alive[(size_t)t]) continue;
    for (int q = 0; q < 3; ++q) G[3 * m + q] = remap[(size_t)T[3 * t + q]];
    ++m;
  }
  *nu = k;
  *ng = m;
  (void)why;
  return 0;
}

// igl::writeOBJ: "v x y z" lines, then 1-based "f a b c" lines
bool write_obj(const char* path, int64_t nv, const double* V, int64_t nf, const int32_t* F) {
  FILE* f = std::fopen(path, "w");
  if (!f) return false;
  for (int64_t i = 0; i < nv; ++i) std::fprintf(f, "v %.17g %.17g %.17g\n", V[3 * i], V[3 * i + 1], V[3 * i + 2]);
  for (int64_t t = 0; t < nf; ++t) std::fprintf(f, "f %d %d %d\n", F[3 * t] + 1, F[3 * t + 1] + 1, F[3 * t + 2] + 1);
  return std::fclose(f) == 0;
}

namespace {
// big-endian writers into a byte buffer
void be32(std::string& s, uint32_t v) {
  const char b[4] = {(char)(v >> 24), (char)(v >> 16), (char)(v >> 8), (char)v};
  s.append(b, 4);
}
void be16(std::string& s, uint16_t v) {
  const char b[2] = {(char)(v >> 8), (char)v};
  s.append(b, 2);
}
void bef(std::string& s, float v) {
  uint32_t u;
  std::memcpy(&u, &v, 4);
  be32(s, u);
}
void attr(std::string& s, const char* name, uint16_t size, int32_t type) {
  be16(s, (uint16_t)std::strlen(name));
  s.append(name);
  be16(s, size);
  be32(s, (uint32_t)type);
  for (uint16_t q = 0; q < size; ++q) be32(s, 0u);  // default values: zero
}
}  // namespace

// ParticleWriter::writeParticles (particle_writer.h:13-44) through Partio's
// classic Houdini geometry writer, gzip-compressed (Partio::write(path, data,
// true)): big-endian, magic "Bgeo", 'V', version 5, the counts (points,
// primitives, point groups, primitive groups, point attributes, vertex
// attributes, primitive attributes, detail attributes), the point attribute
// definitions other than the position (name, size, type: 0 float, 1 int,
// 5 vector; zero defaults), then per point its homogeneous position (x, y, z,
// 1) and the attributes, then the two end bytes 0x00 0xff. The attributes are
// the reference's: id (int, the particle's index), velocity (vector), radius
// (float, 0.1).
bool write_bgeo(const char* path, int64_t n, const float* x, const float* v) {
  std::string s;
  s.append("Bgeo");
  s.push_back('V');
  be32(s, 5u);
  be32(s, (uint32_t)n);
  for (int q = 0; q < 3; ++q) be32(s, 0u);  // primitives, point groups, primitive groups
  be32(s, 3u);                              // point attributes
  for (int q = 0; q < 3; ++q) be32(s, 0u);  // vertex, primitive, detail attributes
  attr(s, "id", 1, 1);
  attr(s, "velocity", 3, 5);
  attr(s, "radius", 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    for (int q = 0; q < 3; ++q) bef(s, x[3 * i + q]);
    bef(s, 1.0f);
    be32(s, (uint32_t)(int32_t)i);
    for (int q = 0; q < 3; ++q) bef(s, v[3 * i + q]);
    bef(s, 0.1f);
  }
  s.push_back((char)0x00);
  s.push_back((char)0xff);
  gzFile f = gzopen(path, "wb");
  if (!f) return false;
  size_t at = 0;
  bool ok = true;
  while (ok && at < s.size()) {
    const unsigned chunk = (unsigned)std::min<size_t>(s.size() - at, 1u << 30);
    ok = gzwrite(f, s.data() + at, chunk) == (int)chunk;
    at += chunk;
  }
  return gzclose(f) == Z_OK && ok;
}

}  // namespace mpm_mesh
