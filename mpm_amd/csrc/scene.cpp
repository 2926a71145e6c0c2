// scene.cpp -- the scene input path (C-ABI mpm_scene_*): scene TOML files,
// OBJ meshes, and the rejection sampler that fills a mesh with particles.
//
// Follows the reference's loader (src/main.cu:31-66), Simulation::loadMesh
// and Simulation::addParticles (src/mpm.cu:331-394):
//   * [[material]] tables: density (700), E (1.4e5), Nu (0.2), hardening (10),
//     plast_clamp_lower (0.975), plast_clamp_higher (1.0075), particle volume
//     1 / particle_count; every material is the snow model (include/mpm.cuh:26).
//     Extension: `model` = "snow" | "fixed_corotated" | "jelly" | "wc_liquid"
//     and `bulk_modulus` for the liquid.
//   * [[object]] tables: material (by name), mesh (file under <scene dir>/
//     meshes/, default sphere.obj), size (1.0, the longest bounding-box edge),
//     position (lowest corner), velocity, lifetime_begin (0), lifetime_end (max).
//   * sampling: points uniform in the mesh's bounding box, in batches of 2048,
//     three draws per point, kept where the winding number is 1, until
//     u32(u32(1 / particle volume) * bounding-box volume) points were drawn.
//   * the draws: float(rand()) / float(RAND_MAX) of the C library's rand();
//     the reference never seeds it (seed 1). This file carries its own copy of
//     glibc's generator (the TYPE_3 additive feedback generator of
//     random_r.c: x_k = x_{k-31} + x_{k-3} mod 2^32 after a 16807-LCG fill
//     and 310 discarded outputs), so a given seed gives the same particles as
//     the reference built with glibc, independent of other rand() users.
// Differences: an object naming an unknown material is an error (the
// reference silently uses material 0); a point is inside where the winding
// number rounds to 1 (the reference's library rounding is not specified).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mpm_hip.h"

namespace {

thread_local std::string g_scene_error;

int scene_fail(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_scene_error = buf;
  return MPM_ERR_IO;
}

// ------------------------------------------------------------------ TOML subset
struct TVal {
  enum Kind { STR, NUM, BOOL, ARR } kind = NUM;
  std::string s;
  double d = 0.0;
  bool b = false;
  std::vector<TVal> a;
};
using TTable = std::map<std::string, TVal>;

struct TomlDoc {
  std::map<std::string, std::vector<TTable>> arrays;  // [[name]]
  std::map<std::string, TTable> tables;               // [name]
  TTable root;
};

struct Parser {
  const std::string& src;
  size_t i = 0;
  int line = 1;
  std::string err;

  explicit Parser(const std::string& s) : src(s) {}

  bool fail(const std::string& m) {
    if (err.empty()) err = "line " + std::to_string(line) + ": " + m;
    return false;
  }
  bool eof() const { return i >= src.size(); }
  char peek() const { return eof() ? '\0' : src[i]; }
  void skip_ws() {  // spaces and tabs
    while (!eof() && (src[i] == ' ' || src[i] == '\t')) ++i;
  }
  void skip_comment() {
    if (peek() == '#')
      while (!eof() && src[i] != '\n') ++i;
  }
  // whitespace, newlines and comments (inside arrays)
  void skip_all() {
    for (;;) {
      skip_ws();
      skip_comment();
      if (peek() == '\r') { ++i; continue; }
      if (peek() == '\n') { ++i; ++line; continue; }
      return;
    }
  }
  bool end_of_line() {
    skip_ws();
    skip_comment();
    if (peek() == '\r') ++i;
    if (eof()) return true;
    if (peek() != '\n') return fail(std::string("unexpected '") + peek() + "' after value");
    ++i;
    ++line;
    return true;
  }
  bool key(std::string& k) {
    skip_ws();
    k.clear();
    if (peek() == '"') {
      TVal v;
      if (!string(v)) return false;
      k = v.s;
      return true;
    }
    while (!eof() && (std::isalnum((unsigned char)src[i]) || src[i] == '_' || src[i] == '-')) k += src[i++];
    if (k.empty()) return fail("expected a key");
    return true;
  }
  bool string(TVal& v) {
    const char q = src[i++];
    v.kind = TVal::STR;
    v.s.clear();
    while (!eof() && src[i] != q) {
      if (src[i] == '\n') return fail("newline in string");
      if (q == '"' && src[i] == '\\' && i + 1 < src.size()) {
        const char e = src[i + 1];
        const char* map = "\"\"\\\\n\nt\tr\r";
        const char* f = nullptr;
        for (const char* m = map; *m; m += 2)
          if (*m == e) f = m;
        if (!f) return fail("unsupported escape in string");
        v.s += f[1];
        i += 2;
        continue;
      }
      v.s += src[i++];
    }
    if (eof()) return fail("unterminated string");
    ++i;
    return true;
  }
  bool value(TVal& v) {
    skip_ws();
    const char c = peek();
    if (c == '"' || c == '\'') return string(v);
    if (c == '[') {
      ++i;
      v.kind = TVal::ARR;
      v.a.clear();
      for (;;) {
        skip_all();
        if (peek() == ']') { ++i; return true; }
        TVal e;
        if (!value(e)) return false;
        v.a.push_back(e);
        skip_all();
        if (peek() == ',') { ++i; continue; }
        if (peek() == ']') { ++i; return true; }
        return fail("expected ',' or ']' in array");
      }
    }
    if (src.compare(i, 4, "true") == 0) { v.kind = TVal::BOOL; v.b = true; i += 4; return true; }
    if (src.compare(i, 5, "false") == 0) { v.kind = TVal::BOOL; v.b = false; i += 5; return true; }
    std::string num;
    while (!eof() && (std::isalnum((unsigned char)src[i]) || src[i] == '+' || src[i] == '-' ||
                      src[i] == '.' || src[i] == '_'))
      if (src[i++] != '_') num += src[i - 1];
    if (num.empty()) return fail("expected a value");
    if (num == "inf" || num == "+inf") { v.kind = TVal::NUM; v.d = INFINITY; return true; }
    if (num == "-inf") { v.kind = TVal::NUM; v.d = -INFINITY; return true; }
    char* end = nullptr;
    errno = 0;
    v.d = std::strtod(num.c_str(), &end);
    if (!end || *end || errno == ERANGE) return fail("bad number '" + num + "'");
    v.kind = TVal::NUM;
    return true;
  }
  bool parse(TomlDoc& doc) {
    TTable* cur = &doc.root;
    while (!eof()) {
      skip_ws();
      skip_comment();
      if (peek() == '\r') { ++i; continue; }
      if (peek() == '\n') { ++i; ++line; continue; }
      if (eof()) break;
      if (peek() == '[') {
        const bool arr = i + 1 < src.size() && src[i + 1] == '[';
        i += arr ? 2 : 1;
        std::string name;
        if (!key(name)) return false;
        skip_ws();
        if (src.compare(i, arr ? 2 : 1, arr ? "]]" : "]") != 0) return fail("bad table header");
        i += arr ? 2 : 1;
        if (arr) {
          doc.arrays[name].emplace_back();
          cur = &doc.arrays[name].back();
        } else {
          cur = &doc.tables[name];
        }
        if (!end_of_line()) return false;
        continue;
      }
      std::string k;
      if (!key(k)) return false;
      skip_ws();
      if (peek() != '=') return fail("expected '=' after key '" + k + "'");
      ++i;
      TVal v;
      if (!value(v)) return false;
      if (cur->count(k)) return fail("duplicate key '" + k + "'");
      (*cur)[k] = v;
      if (!end_of_line()) return false;
    }
    return true;
  }
};

// ------------------------------------------------------------------ generator
// glibc's rand() (TYPE_3, degree 31, separation 3) for one seed.
struct GlibcRand {
  uint32_t ring[31];  // the 31 state words; f leads r by 3
  int f = 3, b = 0;

  explicit GlibcRand(uint32_t seed) {
    int32_t word = seed == 0 ? 1 : (int32_t)seed;
    int32_t st[31];
    st[0] = word;
    for (int i = 1; i < 31; ++i) {
      // 16807 * word mod (2^31 - 1), by Schrage's method
      const int32_t hi = word / 127773, lo = word % 127773;
      word = 16807 * lo - 2836 * hi;
      if (word < 0) word += 2147483647;
      st[i] = word;
    }
    for (int i = 0; i < 31; ++i) ring[i] = (uint32_t)st[i];
    f = 3;
    b = 0;
    for (int i = 0; i < 310; ++i) next();
  }
  // one output in [0, 2^31)
  int32_t next() {
    ring[f] += ring[b];
    const int32_t out = (int32_t)(ring[f] >> 1);
    if (++f >= 31) f = 0;
    if (++b >= 31) b = 0;
    return out;
  }
  float uniform() { return float(next()) / float(2147483647); }  // float(rand()) / float(RAND_MAX)
};

// ------------------------------------------------------------------ mesh
struct Mesh {
  std::vector<float> V;  // 3 per vertex
  std::vector<int> F;    // 3 per triangle
};

bool read_obj(const std::string& path, Mesh& m, std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = "cannot open mesh '" + path + "'";
    return false;
  }
  std::string ln;
  int lineno = 0;
  while (std::getline(in, ln)) {
    ++lineno;
    std::istringstream ss(ln);
    std::string tag;
    if (!(ss >> tag)) continue;
    if (tag == "v") {
      double x, y, z;
      if (!(ss >> x >> y >> z)) {
        err = path + ":" + std::to_string(lineno) + ": bad vertex";
        return false;
      }
      m.V.push_back((float)x);
      m.V.push_back((float)y);
      m.V.push_back((float)z);
    } else if (tag == "f") {
      std::vector<int> idx;
      std::string tok;
      while (ss >> tok) {
        const long v = std::strtol(tok.c_str(), nullptr, 10);  // "v", "v/t", "v//n", "v/t/n"
        const long nv = (long)(m.V.size() / 3);
        const long i0 = v > 0 ? v - 1 : nv + v;
        if (v == 0 || i0 < 0 || i0 >= nv) {
          err = path + ":" + std::to_string(lineno) + ": bad face index '" + tok + "'";
          return false;
        }
        idx.push_back((int)i0);
      }
      if (idx.size() < 3) {
        err = path + ":" + std::to_string(lineno) + ": face with fewer than 3 vertices";
        return false;
      }
      for (size_t t = 1; t + 1 < idx.size(); ++t) {  // fan
        m.F.push_back(idx[0]);
        m.F.push_back(idx[t]);
        m.F.push_back(idx[t + 1]);
      }
    }
  }
  if (m.V.empty() || m.F.empty()) {
    err = "mesh '" + path + "' has no faces";
    return false;
  }
  return true;
}

// Simulation::loadMesh (mpm.cu:331-345): scale so the longest bounding-box
// edge is `size`, then move the lowest corner to `position` (float math).
void place_mesh(Mesh& m, double size, const float pos[3]) {
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = m.V[a];
  for (size_t v = 0; v < m.V.size(); v += 3)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], m.V[v + a]);
      hi[a] = std::max(hi[a], m.V[v + a]);
    }
  float len = 0.0f;
  for (int a = 0; a < 3; ++a) len = std::max(len, hi[a] - lo[a]);
  const float scale = (float)(size / (double)len);
  float shift[3];
  for (int a = 0; a < 3; ++a) shift[a] = pos[a] - scale * lo[a];
  for (size_t v = 0; v < m.V.size(); v += 3)
    for (int a = 0; a < 3; ++a) m.V[v + a] = m.V[v + a] * scale + shift[a];
}

// generalized winding number of a closed triangle mesh at p: the sum of the
// triangles' signed solid angles over 4 pi (van Oosterom and Strackee 1983)
double winding_number(const Mesh& m, const double p[3]) {
  double total = 0.0;
  for (size_t t = 0; t < m.F.size(); t += 3) {
    double a[3], b[3], c[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = (double)m.V[3 * m.F[t] + k] - p[k];
      b[k] = (double)m.V[3 * m.F[t + 1] + k] - p[k];
      c[k] = (double)m.V[3 * m.F[t + 2] + k] - p[k];
    }
    const double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double lb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    const double lc = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                       a[2] * (b[0] * c[1] - b[1] * c[0]);
    const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    const double bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
    const double ca = c[0] * a[0] + c[1] * a[1] + c[2] * a[2];
    total += 2.0 * std::atan2(det, la * lb * lc + ab * lc + bc * la + ca * lb);
  }
  return total / (4.0 * M_PI);
}

struct SceneObject {
  int32_t material = 0;
  int64_t first = 0, count = 0;
  double lifetime_begin = 0.0, lifetime_end = std::numeric_limits<double>::max();
};

}  // namespace

struct mpm_scene_s {
  std::vector<mpm_material> materials;
  std::vector<SceneObject> objects;
  std::vector<float> x, v;
  std::vector<uint8_t> mat;
};

namespace {

bool get_num(const TTable& t, const char* k, double def, double& out, std::string& err) {
  auto it = t.find(k);
  if (it == t.end()) {
    out = def;
    return true;
  }
  if (it->second.kind != TVal::NUM) {
    err = std::string("'") + k + "' must be a number";
    return false;
  }
  out = it->second.d;
  return true;
}

bool get_vec3(const TTable& t, const char* k, float out[3], std::string& err) {
  auto it = t.find(k);
  if (it == t.end()) {
    err = std::string("object needs '") + k + "'";
    return false;
  }
  const TVal& v = it->second;
  if (v.kind != TVal::ARR || v.a.size() != 3) {
    err = std::string("'") + k + "' must be an array of 3 numbers";
    return false;
  }
  for (int a = 0; a < 3; ++a) {
    if (v.a[a].kind != TVal::NUM) {
      err = std::string("'") + k + "' must be an array of 3 numbers";
      return false;
    }
    out[a] = (float)v.a[a].d;
  }
  return true;
}

// Simulation::addParticles (mpm.cu:347-394)
void sample_mesh(const Mesh& m, uint32_t density, const float vel[3], uint8_t mat, GlibcRand& rng,
                 mpm_scene_s& sc) {
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = m.V[a];
  for (size_t i = 0; i < m.V.size(); i += 3)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], (double)m.V[i + a]);
      hi[a] = std::max(hi[a], (double)m.V[i + a]);
    }
  const double range[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const double vol = range[0] * range[1] * range[2];
  const uint32_t target = (uint32_t)((double)density * vol);
  constexpr int kBatch = 2048;
  std::vector<float> pts(3 * kBatch);
  uint32_t drawn = 0;
  while (drawn < target) {
    for (int i = 0; i < kBatch; ++i)
      for (int a = 0; a < 3; ++a) pts[3 * i + a] = (float)(lo[a] + (double)rng.uniform() * range[a]);
    for (int i = 0; i < kBatch; ++i) {
      ++drawn;
      const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
      if (std::lround(winding_number(m, p)) == 1) {
        for (int a = 0; a < 3; ++a) {
          sc.x.push_back(pts[3 * i + a]);
          sc.v.push_back(vel[a]);
        }
        sc.mat.push_back(mat);
      }
      if (drawn == target) return;
    }
  }
}

int load_scene(const char* path, uint32_t particle_count, uint32_t seed, mpm_scene_s& sc) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return scene_fail("mpm_scene_load: cannot open '%s'", path);
  std::stringstream buf;
  buf << in.rdbuf();
  const std::string text = buf.str();
  TomlDoc doc;
  Parser ps(text);
  if (!ps.parse(doc)) return scene_fail("mpm_scene_load: %s: %s", path, ps.err.c_str());
  std::string dir = path;
  const size_t slash = dir.find_last_of('/');
  dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
  const std::string meshes = dir + "/meshes/";

  std::map<std::string, int> index;
  std::string err;
  const double volume = 1.0 / (double)particle_count;
  for (const TTable& t : doc.arrays["material"]) {
    double density, E, nu, hard, lo, hi, bulk;
    if (!get_num(t, "density", 700.0, density, err) || !get_num(t, "E", 1.4e5, E, err) ||
        !get_num(t, "Nu", 0.2, nu, err) || !get_num(t, "hardening", 10.0, hard, err) ||
        !get_num(t, "plast_clamp_lower", 0.975, lo, err) ||
        !get_num(t, "plast_clamp_higher", 1.0075, hi, err) ||
        !get_num(t, "bulk_modulus", 0.0, bulk, err))
      return scene_fail("mpm_scene_load: %s: material %zu: %s", path, sc.materials.size(), err.c_str());
    int32_t kind = MPM_MATERIAL_SNOW;
    auto mk = t.find("model");
    if (mk != t.end()) {
      const std::string k = mk->second.s;
      if (k == "snow") kind = MPM_MATERIAL_SNOW;
      else if (k == "fixed_corotated") kind = MPM_MATERIAL_FIXED_COROTATED;
      else if (k == "jelly") kind = MPM_MATERIAL_JELLY;
      else if (k == "wc_liquid") kind = MPM_MATERIAL_WC_LIQUID;
      else return scene_fail("mpm_scene_load: %s: unknown material model '%s'", path, k.c_str());
    }
    if (sc.materials.size() >= 256) return scene_fail("mpm_scene_load: %s: more than 256 materials", path);
    mpm_material m;
    const int rc = mpm_material_init(&m, kind, volume, density, E, nu, hard, lo, hi, bulk);
    if (rc) return scene_fail("mpm_scene_load: %s: bad material %zu", path, sc.materials.size());
    auto nk = t.find("name");
    const std::string name = nk == t.end() ? std::string() : nk->second.s;
    index[name] = (int)sc.materials.size();
    sc.materials.push_back(m);
  }
  GlibcRand rng(seed);
  int oi = 0;
  for (const TTable& t : doc.arrays["object"]) {
    SceneObject o;
    auto mk = t.find("material");
    const std::string mname = mk == t.end() ? std::string() : mk->second.s;
    auto mi = index.find(mname);
    if (mi == index.end())
      return scene_fail("mpm_scene_load: %s: object %d names unknown material '%s'", path, oi, mname.c_str());
    o.material = mi->second;
    auto mesh_it = t.find("mesh");
    const std::string mesh = meshes + (mesh_it == t.end() ? std::string("sphere.obj") : mesh_it->second.s);
    double size;
    float pos[3], vel[3];
    if (!get_num(t, "size", 1.0, size, err) || !get_vec3(t, "position", pos, err) ||
        !get_vec3(t, "velocity", vel, err) ||
        !get_num(t, "lifetime_begin", 0.0, o.lifetime_begin, err) ||
        !get_num(t, "lifetime_end", std::numeric_limits<double>::max(), o.lifetime_end, err))
      return scene_fail("mpm_scene_load: %s: object %d: %s", path, oi, err.c_str());
    Mesh m;
    if (!read_obj(mesh, m, err)) return scene_fail("mpm_scene_load: %s: object %d: %s", path, oi, err.c_str());
    place_mesh(m, (double)(float)size, pos);
    const float vol_f = sc.materials[o.material].particle_volume;
    const uint32_t density = (uint32_t)(1.0 / (double)vol_f);  // u32(1.0 / material.particleVolume)
    o.first = (int64_t)sc.mat.size();
    sample_mesh(m, density, vel, (uint8_t)o.material, rng, sc);
    o.count = (int64_t)sc.mat.size() - o.first;
    sc.objects.push_back(o);
    ++oi;
  }
  return MPM_OK;
}

}  // namespace

extern "C" const char* mpm_scene_last_error(void) { return g_scene_error.c_str(); }

extern "C" int mpm_scene_load(const char* toml_path, uint32_t particle_count, uint32_t seed,
                              mpm_scene* out) {
  if (!toml_path || !out || particle_count == 0) {
    g_scene_error = "mpm_scene_load: need a path, an output and particle_count > 0";
    return MPM_ERR_INVALID;
  }
  *out = nullptr;
  std::unique_ptr<mpm_scene_s> sc(new mpm_scene_s());
  const int rc = load_scene(toml_path, particle_count, seed, *sc);
  if (rc) return rc;
  *out = sc.release();
  return MPM_OK;
}

extern "C" int mpm_scene_destroy(mpm_scene s) {
  delete s;
  return MPM_OK;
}

extern "C" int mpm_scene_info(mpm_scene s, int32_t* n_materials, int32_t* n_objects, int64_t* n_particles) {
  if (!s || !n_materials || !n_objects || !n_particles) return MPM_ERR_INVALID;
  *n_materials = (int32_t)s->materials.size();
  *n_objects = (int32_t)s->objects.size();
  *n_particles = (int64_t)s->mat.size();
  return MPM_OK;
}

extern "C" int mpm_scene_materials(mpm_scene s, mpm_material* out) {
  if (!s || (!out && !s->materials.empty())) return MPM_ERR_INVALID;
  std::copy(s->materials.begin(), s->materials.end(), out);
  return MPM_OK;
}

extern "C" int mpm_scene_object(mpm_scene s, int32_t i, int32_t* material, int64_t* first,
                                int64_t* count, double* lifetime_begin, double* lifetime_end) {
  if (!s || i < 0 || i >= (int32_t)s->objects.size() || !material || !first || !count ||
      !lifetime_begin || !lifetime_end)
    return MPM_ERR_INVALID;
  const SceneObject& o = s->objects[(size_t)i];
  *material = o.material;
  *first = o.first;
  *count = o.count;
  *lifetime_begin = o.lifetime_begin;
  *lifetime_end = o.lifetime_end;
  return MPM_OK;
}

extern "C" int mpm_scene_particles_soa(mpm_scene s, float* x, float* v, uint8_t* material) {
  if (!s) return MPM_ERR_INVALID;
  if (s->mat.empty()) return MPM_OK;
  if (!x || !v || !material) return MPM_ERR_INVALID;
  std::copy(s->x.begin(), s->x.end(), x);
  std::copy(s->v.begin(), s->v.end(), v);
  std::copy(s->mat.begin(), s->mat.end(), material);
  return MPM_OK;
}

// ---------------------------------------------------------------- one mesh
struct mpm_sampler_s {
  GlibcRand rng;
  std::vector<float> x;  // the last fill's points
  explicit mpm_sampler_s(uint32_t seed) : rng(seed) {}
};

extern "C" int mpm_sampler_create(uint32_t seed, mpm_sampler* out) {
  if (!out) return MPM_ERR_INVALID;
  *out = new mpm_sampler_s(seed);
  return MPM_OK;
}

extern "C" int mpm_sampler_destroy(mpm_sampler s) {
  delete s;
  return MPM_OK;
}

extern "C" int mpm_sampler_fill_mesh(mpm_sampler s, const char* obj_path, double size,
                                     const float position[3], float particle_volume, int64_t* count) {
  if (!s || !obj_path || !position || !count || !(particle_volume > 0.0f)) {
    g_scene_error = "mpm_sampler_fill_mesh: bad arguments";
    return MPM_ERR_INVALID;
  }
  Mesh m;
  std::string err;
  if (!read_obj(obj_path, m, err)) return scene_fail("mpm_sampler_fill_mesh: %s", err.c_str());
  place_mesh(m, size, position);
  mpm_scene_s tmp;
  const float zero[3] = {0.f, 0.f, 0.f};
  sample_mesh(m, (uint32_t)(1.0 / (double)particle_volume), zero, 0, s->rng, tmp);
  s->x.swap(tmp.x);
  *count = (int64_t)(s->x.size() / 3);
  return MPM_OK;
}

extern "C" int mpm_sampler_points(mpm_sampler s, float* x) {
  if (!s || (!x && !s->x.empty())) return MPM_ERR_INVALID;
  std::copy(s->x.begin(), s->x.end(), x);
  return MPM_OK;
}

extern "C" int mpm_debug_glibc_rand(uint32_t seed, int32_t n, int32_t* out) {
  if (!out || n < 0) return MPM_ERR_INVALID;
  GlibcRand r(seed);
  for (int32_t i = 0; i < n; ++i) out[i] = r.next();
  return MPM_OK;
}
