// mpm_simulation.hpp -- C++ facade over the C-ABI (mpm_hip.h) with the
// reference's class and method names (class Simulation, include/mpm.cuh:44-106;
// CLIOptions, include/options.h:6-56; SimObject, mpm.cuh:30-41), so a caller
// written against the reference's Simulation compiles against this header.
// Header-only; link with -lmpm_hip.
//
// Differences from the reference: errors throw mpm::Error (the reference
// ignores CUDA errors); objects whose lifetime starts or ends are brought onto
// or taken off the device at the first advance() where t crosses the bound
// (the reference adds them at the next syncDevice, mpm.cu:288-314, and never
// removes them). An object that starts is appended on the device
// (mpm_append_particles): the particles already there are not copied to the
// host and back; one that ends takes a round trip. The material list is
// copied, not referenced.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpm_hip.h"

namespace mpm {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// the library's messages name the failing entry point already
inline void check(int rc, mpm_handle h, const char* what) {
  if (rc == MPM_OK) return;
  const std::string m = mpm_last_error(h);
  throw Error(rc, m.empty() ? std::string(what) + " failed" : m);
}

// The reference's particle bytes (MLS_APIC_Particle, include/types.h:24-35 +
// TransferScheme.h:46-54): 104 bytes, F and C column-major.
struct Particle {
  uint8_t material_type;
  uint8_t pad_[3];
  float x[3], v[3];
  float F[9], C[9];
  float Jp;
  Particle() : material_type(0), pad_{0, 0, 0}, x{0, 0, 0}, v{0, 0, 0},
               F{1, 0, 0, 0, 1, 0, 0, 0, 1}, C{0, 0, 0, 0, 0, 0, 0, 0, 0}, Jp(1.0f) {}
  Particle(uint8_t m, const float* pos, const float* vel) : Particle() {
    material_type = m;
    for (int a = 0; a < 3; ++a) { x[a] = pos[a]; v[a] = vel[a]; }
  }
};
static_assert(sizeof(Particle) == 104, "the reference's particle is 104 bytes");

// include/options.h:6-56, plus the parameters the reference hard-codes
// (gravity mpm.cu:6, walls mpm.cu:92) and the engine's own options.
struct CLIOptions {
  float dt = 1e-4f;
  uint32_t N = 60;
  uint32_t particle_count = 500000;
  float dx = 1.0f / 60.0f;
  float N_real = 60.0f;
  std::string save_dir;
  std::string scene = "scenes/snowman.toml";
  uint32_t mesh_grid = 250;
  uint32_t mesh_particle_radius = 5;
  int32_t mesh_face_count = -1;
  int laplacian_smooth = 0;
  bool no_mesh = false;  // (no counterpart: the reference always writes meshes with --save-dir)
  std::string frames_dir;  // (no counterpart: the reference shows its view in a window)
  // not reference flags
  float gravity = -9.81f;
  float boundary = 0.05f;
  int32_t compat = MPM_COMPAT_REFERENCE;
  uint32_t seed = 1;
  int64_t steps = 1000;
  uint32_t sync_every = 20;

  CLIOptions() = default;
  // --name value or --name=value; the grid size also in the reference's
  // short form -N value (cxxopts 2.2 gives the one-letter option name "N",
  // options.h:26, only a short flag), and -Nvalue. Unknown flags and bad
  // values print a message and exit(1), as cxxopts' failure does in
  // options.h:48-51.
  CLIOptions(int argc, char* argv[]) {
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i], val;
      if (a.size() >= 2 && a[0] == '-' && a[1] == 'N') {  // short flag
        if (a.size() > 2) {
          val = a.substr(2);
        } else {
          if (i + 1 >= argc) die("option 'N' needs a value");
          val = argv[++i];
        }
        N = (uint32_t)to_i("N", val, 1);
        continue;
      }
      if (a.rfind("--", 0) != 0) die("unexpected argument '" + a + "'");
      a = a.substr(2);
      const size_t eq = a.find('=');
      if (eq != std::string::npos) {
        val = a.substr(eq + 1);
        a = a.substr(0, eq);
      } else if (a == "help") {
        usage();
        std::exit(0);
      } else if (a == "no-mesh") {
        no_mesh = true;
        continue;
      } else {
        if (i + 1 >= argc) die("option '" + a + "' needs a value");
        val = argv[++i];
      }
      if (a == "dt") dt = to_f(a, val);
      else if (a == "N") N = (uint32_t)to_i(a, val, 1);
      else if (a == "save-dir") save_dir = val;
      else if (a == "scene") scene = val;
      else if (a == "particle-count") particle_count = (uint32_t)to_i(a, val, 1);
      else if (a == "mesh-grid") mesh_grid = (uint32_t)to_i(a, val, 1);
      else if (a == "mesh-particle-radius") mesh_particle_radius = (uint32_t)to_i(a, val, 0);
      else if (a == "mesh-face-count") mesh_face_count = (int32_t)to_i(a, val, -1);
      else if (a == "laplacian_smooth") laplacian_smooth = (int)to_i(a, val, 0);
      else if (a == "gravity") gravity = to_f(a, val);
      else if (a == "boundary") boundary = to_f(a, val);
      else if (a == "compat") compat = val == "fixed" ? MPM_COMPAT_FIXED
                                     : val == "reference" ? MPM_COMPAT_REFERENCE
                                     : (die("compat must be 'reference' or 'fixed'"), 0);
      else if (a == "seed") seed = (uint32_t)to_i(a, val, 0);
      else if (a == "steps") steps = to_i(a, val, 0);
      else if (a == "sync-every") sync_every = (uint32_t)to_i(a, val, 1);
      else if (a == "frames-dir") frames_dir = val;
      else die("unknown option '" + a + "'");
    }
    dx = (float)(1.0 / N);  // options.h:41
    N_real = (float)N;
  }

  static void usage() {
    std::printf(
        "MPM simulation.\n"
        "  --dt arg                    Physics simulation time-step (default 1e-4)\n"
        "  -N arg                      Grid dimensions (default 60; also --N)\n"
        "  --save-dir arg              Where to save particles (default none)\n"
        "  --scene arg                 Scene file to load (default scenes/snowman.toml)\n"
        "  --particle-count arg        Particles per unit volume (default 500000)\n"
        "  --mesh-grid arg             Grid size for computing mesh (default 250)\n"
        "  --mesh-particle-radius arg  Particle radius for the mesh, in grid points (default 5)\n"
        "  --mesh-face-count arg       Approximate the mesh with this many faces (default -1: all)\n"
        "  --laplacian_smooth arg      Apply Laplacian smoothing to the mesh (default 0: off)\n"
        "  --no-mesh                   with --save-dir, write particles only\n"
        "  --steps arg                 substeps to run (default 1000)\n"
        "  --gravity arg               y acceleration (default -9.81)\n"
        "  --boundary arg              wall fraction (default 0.05)\n"
        "  --compat reference|fixed    reference quirks on or off (default reference)\n"
        "  --seed arg                  sampler seed (default 1, the reference's unseeded rand())\n"
        "  --sync-every arg            substeps between device syncs (default 20)\n"
        "  --frames-dir arg            write the debug view every 20 substeps as PPM images\n");
  }

 private:
  [[noreturn]] static void die(const std::string& m) {
    std::printf("%s\n", m.c_str());
    std::exit(1);
  }
  static float to_f(const std::string& k, const std::string& v) {
    char* e = nullptr;
    const double d = std::strtod(v.c_str(), &e);
    if (!e || *e || v.empty()) die("option '" + k + "': bad number '" + v + "'");
    return (float)d;
  }
  static long long to_i(const std::string& k, const std::string& v, long long lo) {
    char* e = nullptr;
    const long long d = std::strtoll(v.c_str(), &e, 10);
    if (!e || *e || v.empty() || d < lo) die("option '" + k + "': bad integer '" + v + "'");
    return d;
  }
};

// mpm.cuh:30-41
struct SimObject {
  explicit SimObject(const mpm_material& model) : materialModel(model) {}
  std::vector<Particle> particles;
  mpm_material materialModel;
  double lifetime_begin = 0.0;
  double lifetime_end = std::numeric_limits<double>::max();
  bool isActive(double t) const { return t >= lifetime_begin && t < lifetime_end; }
};

class Simulation {
 public:
  mpm_params par;
  double t = 0.0;
  std::vector<SimObject> objects;
  int64_t active_particle_count = 0;
  std::vector<mpm_material> material_models;

  Simulation(const CLIOptions& opts, const std::vector<mpm_material>& materials)
      : material_models(materials), seed_(opts.seed) {
    par.dt = opts.dt;
    par.N = opts.N;
    par.gravity = opts.gravity;
    par.boundary_mode = MPM_BOUNDARY_FRACTION;
    par.boundary = opts.boundary;
    par.compat = opts.compat;
    check(mpm_create(&par, &h_), nullptr, "mpm_create");
  }
  ~Simulation() {
    if (sampler_) mpm_sampler_destroy(sampler_);
    mpm_destroy(h_);
  }
  Simulation(const Simulation&) = delete;
  Simulation& operator=(const Simulation&) = delete;

  mpm_handle handle() const { return h_; }

  // Simulation::addObject (mpm.cu:237-251): fill the OBJ mesh at `filepath`,
  // scaled so its longest bounding-box edge is `size` with its lowest corner
  // at `position`, with particles of `material_model_index` moving at `velocity`.
  void addObject(const std::string& filepath, int material_model_index, float size,
                 const float position[3], const float velocity[3], double lifetime_begin = 0.0,
                 double lifetime_end = std::numeric_limits<double>::max()) {
    if (material_model_index < 0 || material_model_index >= (int)material_models.size())
      throw Error(MPM_ERR_INVALID, "addObject: no material " + std::to_string(material_model_index));
    if (!sampler_) check(mpm_sampler_create(seed_, &sampler_), nullptr, "mpm_sampler_create");
    const mpm_material& m = material_models[(size_t)material_model_index];
    int64_t n = 0;
    const int rc = mpm_sampler_fill_mesh(sampler_, filepath.c_str(), (double)size, position,
                                         m.particle_volume, &n);
    if (rc) throw Error(rc, mpm_scene_last_error());
    std::vector<float> x((size_t)n * 3);
    mpm_sampler_points(sampler_, x.data());
    SimObject o(m);
    o.lifetime_begin = lifetime_begin;
    o.lifetime_end = lifetime_end;
    for (int64_t i = 0; i < n; ++i)
      o.particles.emplace_back((uint8_t)material_model_index, &x[(size_t)(3 * i)], velocity);
    objects.push_back(std::move(o));
  }

  // An object whose particles the caller already sampled (e.g. from a scene file).
  void addObject(SimObject o) { objects.push_back(std::move(o)); }

  // Simulation::initCuda (mpm.cu:197-207)
  void initCuda() {
    check(mpm_set_materials(h_, material_models.data(), (int32_t)material_models.size()), h_,
          "mpm_set_materials");
    upload_active();
  }

  // Simulation::advance (mpm.cu:323-329): one substep, then t += dt. Objects
  // whose lifetime ended leave the device (their last state copied back into
  // the object), objects whose lifetime began join it behind the others;
  // the rest of the particles stay on the device (the reference copies every
  // particle to the host and back, particlesToHost + reallocateParticles,
  // mpm.cu:288-314).
  void advance() {
    const std::vector<char> now = active_mask(t);
    if (now != uploaded_) {
      if (active_particle_count == 0) {
        upload_active();
      } else {
        remove_ended(now);
        append_started(now);
      }
    }
    if (active_particle_count > 0) check(mpm_advance(h_, 1, nullptr), h_, "mpm_advance");
    t += par.dt;
  }

  // Simulation::syncDevice (mpm.cu:209-211, 288-306): device particles back
  // into the objects that are on the device
  void syncDevice() {
    if (active_particle_count == 0) return;
    std::vector<Particle> buf((size_t)active_particle_count);
    check(mpm_download_aos104(h_, buf.data(), nullptr), h_, "mpm_download_aos104");
    size_t i = 0;
    for (size_t k : on_device_)
      for (Particle& p : objects[k].particles) p = buf[i++];
  }

  size_t getFullParticleCount() const {
    size_t n = 0;
    for (const SimObject& o : objects) n += o.particles.size();
    return n;
  }
  std::vector<Particle>& getFullParticleList() {
    all_.clear();
    for (const SimObject& o : objects) all_.insert(all_.end(), o.particles.begin(), o.particles.end());
    return all_;
  }
  std::vector<Particle>& getActiveParticleList() {
    all_.clear();
    for (const SimObject& o : objects)
      if (o.isActive(t)) all_.insert(all_.end(), o.particles.begin(), o.particles.end());
    return all_;
  }

 private:
  std::vector<char> active_mask(double at) const {
    std::vector<char> m(objects.size());
    for (size_t k = 0; k < objects.size(); ++k) m[k] = objects[k].isActive(at) ? 1 : 0;
    return m;
  }
  void upload_active() {
    uploaded_ = active_mask(t);
    on_device_.clear();
    check(mpm_clear_particles(h_), h_, "mpm_clear_particles");
    int64_t n = 0;
    for (size_t k = 0; k < objects.size(); ++k) {
      if (!uploaded_[k]) continue;
      on_device_.push_back(k);
      if (objects[k].particles.empty()) continue;
      check(mpm_add_particles_aos104(h_, objects[k].particles.data(), (int64_t)objects[k].particles.size()),
            h_, "mpm_add_particles_aos104");
      n += (int64_t)objects[k].particles.size();
    }
    active_particle_count = n;
    if (n > 0) check(mpm_upload(h_, nullptr), h_, "mpm_upload");
  }
  // objects that stopped being active: their particles leave the device, their
  // last state back into the object
  void remove_ended(const std::vector<char>& now) {
    int64_t first = 0;
    std::vector<size_t> kept;
    for (size_t k : on_device_) {
      const int64_t n = (int64_t)objects[k].particles.size();
      if (now[k]) {
        kept.push_back(k);
        first += n;
        continue;
      }
      if (n > 0)
        check(mpm_remove_particles(h_, first, n, objects[k].particles.data(), nullptr), h_,
              "mpm_remove_particles");
      uploaded_[k] = 0;
      active_particle_count -= n;
    }
    on_device_ = kept;
  }
  // objects that became active, in object order, behind the device's particles
  void append_started(const std::vector<char>& now) {
    check(mpm_clear_particles(h_), h_, "mpm_clear_particles");
    int64_t n = 0;
    for (size_t k = 0; k < objects.size(); ++k) {
      if (uploaded_[k] || !now[k]) continue;
      on_device_.push_back(k);
      if (objects[k].particles.empty()) continue;
      check(mpm_add_particles_aos104(h_, objects[k].particles.data(), (int64_t)objects[k].particles.size()),
            h_, "mpm_add_particles_aos104");
      n += (int64_t)objects[k].particles.size();
    }
    uploaded_ = now;
    if (n > 0) check(mpm_append_particles(h_, nullptr), h_, "mpm_append_particles");
    active_particle_count += n;
  }

  mpm_handle h_ = nullptr;
  mpm_sampler sampler_ = nullptr;
  uint32_t seed_ = 1;
  std::vector<char> uploaded_;
  std::vector<size_t> on_device_;  // objects on the device, in the order of their particles there
  std::vector<Particle> all_;
};

// MeshBuilder (include/mesh_builder.h:19-35,181-211): the surface mesh of the
// simulation's particles to an OBJ file: distance field and marching cubes on
// the GPU from the device particles (mpm_mesh_build), then on the host the
// optional smoothing (--laplacian_smooth != 0, mesh_builder.h:196-200) and
// decimation (--mesh-face-count != -1, mesh_builder.h:202-207).
class MeshBuilder {
 public:
  MeshBuilder(Simulation& sim, const CLIOptions& flags, uint32_t grid_size)
      : sim_(sim), radius_(flags.mesh_particle_radius), side_(grid_size), smooth_(flags.laplacian_smooth != 0),
        faces_(flags.mesh_face_count) {}
  // the face count written; throws on failure
  int64_t computeMesh(const std::string& filename) {
    int64_t n = 0, nv = 0, nf = 0;
    check(mpm_mesh_build(sim_.handle(), (int32_t)side_, (int32_t)radius_, &n, nullptr), sim_.handle(),
          "mpm_mesh_build");
    check(mpm_mesh_indexed(sim_.handle(), &nv, &nf, nullptr, nullptr), sim_.handle(), "mpm_mesh_indexed");
    std::vector<double> V(3 * (size_t)nv);
    std::vector<int32_t> F(3 * (size_t)nf);
    check(mpm_mesh_indexed(sim_.handle(), &nv, &nf, V.data(), F.data()), sim_.handle(), "mpm_mesh_indexed");
    if (smooth_) check(mpm_mesh_smooth(nv, V.data(), nf, F.data()), nullptr, "mpm_mesh_smooth");
    if (faces_ != -1) {
      std::vector<double> U(V.size());
      std::vector<int32_t> G(F.size());
      int64_t nu = 0, ng = 0;
      check(mpm_mesh_decimate(nv, V.data(), nf, F.data(), faces_ < 0 ? 0 : faces_, U.data(), &nu, G.data(), &ng),
            nullptr, "mpm_mesh_decimate");
      V.swap(U);
      F.swap(G);
      nv = nu;
      nf = ng;
    }
    check(mpm_obj_write(filename.c_str(), nv, V.data(), nf, F.data()), nullptr, "mpm_obj_write");
    return nf;
  }

 private:
  Simulation& sim_;
  uint32_t radius_, side_;
  bool smooth_;
  int32_t faces_;
};

// Renderer (include/renderer.h:7-59): the reference's debug view of the
// device particles (mpm_render_frame), to PPM images instead of a window.
class Renderer {
 public:
  Renderer(Simulation& sim, const std::string& save_dir, uint32_t width = 800, uint32_t height = 600)
      : sim_(sim), dir_(save_dir), w_(width), h_(height), rgb_((size_t)3 * width * height) {}
  // frame_<id>.ppm in the directory; throws on failure
  void render(uint32_t frame_id) {
    check(mpm_render_frame(sim_.handle(), (int32_t)w_, (int32_t)h_, rgb_.data(), nullptr), sim_.handle(),
          "mpm_render_frame");
    char name[64];
    std::snprintf(name, sizeof(name), "/frame_%05u.ppm", frame_id);
    FILE* f = std::fopen((dir_ + name).c_str(), "wb");
    if (!f) throw Error(MPM_ERR_IO, "cannot write " + dir_ + name);
    std::fprintf(f, "P6\n%u %u\n255\n", w_, h_);
    const bool ok = std::fwrite(rgb_.data(), 1, rgb_.size(), f) == rgb_.size();
    if (std::fclose(f) != 0 || !ok) throw Error(MPM_ERR_IO, "cannot write " + dir_ + name);
  }

 private:
  Simulation& sim_;
  std::string dir_;
  uint32_t w_, h_;
  std::vector<uint8_t> rgb_;
};

// ParticleWriter (include/particle_writer.h:7-44): id, position, velocity
// and radius 0.1 of each particle to a .bgeo file (mpm_bgeo_write).
class ParticleWriter {
 public:
  void writeParticles(const std::string& filepath, const std::vector<Particle>& particles) {
    std::vector<float> x(3 * particles.size()), v(3 * particles.size());
    for (size_t i = 0; i < particles.size(); ++i)
      for (int a = 0; a < 3; ++a) {
        x[3 * i + (size_t)a] = particles[i].x[a];
        v[3 * i + (size_t)a] = particles[i].v[a];
      }
    check(mpm_bgeo_write(filepath.c_str(), (int64_t)particles.size(), x.data(), v.data()), nullptr,
          "mpm_bgeo_write");
  }
};

}  // namespace mpm
