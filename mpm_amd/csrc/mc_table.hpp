// mc_table.hpp -- the marching-cubes case table, generated at compile time
// (host and device). For each of the 256 cases (bit q: corner q inside) the
// triangles as cube edges, 3 per triangle.
//
// On each face, seen from outside, a segment runs from every edge crossing
// where the corners go outside -> inside (counter-clockwise) to the next
// crossing: on a face with four crossings this cuts off each inside corner,
// and the neighbouring cube, which sees the face from the other side, makes
// the same cut, so the surface is closed. The segments chain into loops
// around the inside corners, and each loop is cut into a fan from the first
// of its vertices whose diagonals do not lie in a face of the cube (a
// diagonal in a face could meet the neighbour's). The triangles wind with
// their normals pointing out of the inside.
#pragma once

#include <cstdint>

// corner q of a cube = (q >> 2 & 1, q >> 1 & 1, q & 1) along (x, y, z)
namespace mc {
constexpr int kEdge[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                              {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};
// the corners of each face, counter-clockwise seen from outside the cube
constexpr int kFace[6][4] = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1},
                             {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};

struct Table {
  uint8_t ntri[256];
  int8_t tri[256][15];  // 3 edges per triangle
};

constexpr int edge_of(int a, int b) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  for (int e = 0; e < 12; ++e)
    if (kEdge[e][0] == lo && kEdge[e][1] == hi) return e;
  return -1;
}

constexpr bool on_face(int e, int f) {
  for (int i = 0; i < 4; ++i)
    if (edge_of(kFace[f][i], kFace[f][(i + 1) % 4]) == e) return true;
  return false;
}

constexpr bool share_face(int e1, int e2) {
  for (int f = 0; f < 6; ++f)
    if (on_face(e1, f) && on_face(e2, f)) return true;
  return false;
}

constexpr Table make_table() {
  Table t{};
  for (int c = 0; c < 256; ++c) {
    int succ[12] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    for (int f = 0; f < 6; ++f) {
      int ce[4] = {0, 0, 0, 0};
      bool cs[4] = {false, false, false, false};  // crossing: outside -> inside
      int n = 0;
      for (int i = 0; i < 4; ++i) {
        const int a = kFace[f][i], b = kFace[f][(i + 1) % 4];
        const bool ia = (c >> a) & 1, ib = (c >> b) & 1;
        if (ia != ib) {
          ce[n] = edge_of(a, b);
          cs[n] = !ia;
          ++n;
        }
      }
      for (int i = 0; i < n; ++i)
        if (cs[i]) succ[ce[i]] = ce[(i + 1) % n];
    }
    bool seen[12] = {false, false, false, false, false, false, false, false, false, false, false, false};
    int nt = 0;
    for (int e0 = 0; e0 < 12; ++e0) {
      if (succ[e0] < 0 || seen[e0]) continue;
      int loop[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      int k = 0;
      int e = e0;
      do {
        loop[k++] = e;
        seen[e] = true;
        e = succ[e];
      } while (e != e0);
      int r = 0;
      for (; r < k; ++r) {
        bool ok = true;
        for (int q = 2; q < k - 1; ++q) ok = ok && !share_face(loop[r], loop[(r + q) % k]);
        if (ok) break;
      }
      for (int q = 1; q < k - 1; ++q) {
        t.tri[c][3 * nt] = (int8_t)loop[r];
        t.tri[c][3 * nt + 1] = (int8_t)loop[(r + q) % k];
        t.tri[c][3 * nt + 2] = (int8_t)loop[(r + q + 1) % k];
        ++nt;
      }
    }
    t.ntri[c] = (uint8_t)nt;
  }
  return t;
}
}  // namespace mc


