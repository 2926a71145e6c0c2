// mpm_cli.cpp -- the `mpm` program: the reference's src/main.cu:20-116 on
// the C-ABI. Loads a scene TOML (materials, objects, meshes under
// <scene dir>/meshes/), fills the objects with particles, and advances the
// simulation, syncing every --sync-every substeps (main.cu:99-101) and, with
// --save-dir, writing 240 times per simulated second (main.cu:7,93,102-113)
// the surface mesh to <save-dir>/meshes/mesh_<frame>.obj (MeshBuilder,
// main.cu:78,104-106: --mesh-grid voxels, --mesh-particle-radius,
// --laplacian_smooth, --mesh-face-count) and the active particles to
// <save-dir>/particles/particles_<frame>.bgeo (ParticleWriter, main.cu:
// 107-111).
//
// Differences: the run stops after --steps substeps (the reference loops
// until killed, rendering to a window every 20 substeps; with --frames-dir
// the same view goes to PPM images instead, mpm_render_frame); --no-mesh
// skips the mesh.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mpm_hip.h"
#include "../../include/mpm_simulation.hpp"

namespace {

void make_dir(const std::string& d) { ::mkdir(d.c_str(), 0755); }

}  // namespace

int main(int argc, char* argv[]) {
  mpm::CLIOptions flags(argc, argv);
  std::printf("Loading scene file: %s\n", flags.scene.c_str());
  std::printf("grid: N=%u, dt=%g, %u particles per unit volume\n", flags.N, (double)flags.dt,
              flags.particle_count);
  mpm_scene sc = nullptr;
  if (mpm_scene_load(flags.scene.c_str(), flags.particle_count, flags.seed, &sc) != MPM_OK) {
    std::fprintf(stderr, "%s\n", mpm_scene_last_error());
    return 1;
  }
  int32_t nm = 0, no = 0;
  int64_t np = 0;
  mpm_scene_info(sc, &nm, &no, &np);
  std::vector<mpm_material> mats((size_t)nm);
  mpm_scene_materials(sc, mats.data());
  std::vector<float> x((size_t)np * 3), v((size_t)np * 3);
  std::vector<uint8_t> mat((size_t)np);
  mpm_scene_particles_soa(sc, x.data(), v.data(), mat.data());
  try {
    mpm::Simulation simulation(flags, mats);
    for (int32_t i = 0; i < no; ++i) {
      int32_t m;
      int64_t first, count;
      double lb, le;
      mpm_scene_object(sc, i, &m, &first, &count, &lb, &le);
      mpm::SimObject o(mats[(size_t)m]);
      o.lifetime_begin = lb;
      o.lifetime_end = le;
      for (int64_t p = first; p < first + count; ++p)
        o.particles.emplace_back((uint8_t)m, &x[(size_t)(3 * p)], &v[(size_t)(3 * p)]);
      simulation.addObject(std::move(o));
    }
    mpm_scene_destroy(sc);
    std::printf("loaded materials: %d materials, %d objects, %zu particles\n", nm, no,
                simulation.getFullParticleCount());
    simulation.initCuda();

    const bool save = !flags.save_dir.empty();
    const bool mesh = save && !flags.no_mesh;
    if (save) {
      make_dir(flags.save_dir);
      make_dir(flags.save_dir + "/particles");
      if (mesh) make_dir(flags.save_dir + "/meshes");
    }
    mpm::MeshBuilder mesher(simulation, flags, flags.mesh_grid);
    mpm::ParticleWriter writer;
    const bool frames = !flags.frames_dir.empty();
    if (frames) make_dir(flags.frames_dir);
    mpm::Renderer renderer(simulation, flags.frames_dir);
    uint32_t view_id = 0;
    if (frames && simulation.active_particle_count > 0) renderer.render(view_id++);  // main.cu:76
    const uint32_t save_every = std::max<uint32_t>(1, (uint32_t)(1.0 / 240.0 / flags.dt));
    uint32_t frame_id = 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t i = 0; i < flags.steps; ++i) {
      if (i % 10 == 0) {
        std::printf("Step %lld\r", (long long)i);
        std::fflush(stdout);
      }
      simulation.advance();
      if (i % flags.sync_every == 0) simulation.syncDevice();
      if (frames && i % 20 == 0 && simulation.active_particle_count > 0) renderer.render(view_id++);  // main.cu:99-102
      if (save && i % save_every == 0) {
        if (i % flags.sync_every != 0) simulation.syncDevice();
        char name[64];
        if (mesh && simulation.active_particle_count > 0) {
          std::snprintf(name, sizeof(name), "/meshes/mesh_%05u.obj", frame_id);
          mesher.computeMesh(flags.save_dir + name);
        }
        std::snprintf(name, sizeof(name), "/particles/particles_%u.bgeo", frame_id++);
        writer.writeParticles(flags.save_dir + name, simulation.getActiveParticleList());
      }
    }
    simulation.syncDevice();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("\ndone: %lld substeps, t = %.6f s, %lld active particles, %.3f s wall, %u frames\n",
                (long long)flags.steps, simulation.t, (long long)simulation.active_particle_count,
                secs, frame_id);
  } catch (const mpm::Error& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
