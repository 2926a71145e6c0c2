// Host-only driver for the sanitizer test (tests/test_sanitized_host.py):
// built with AddressSanitizer and UBSan together with the scene input path
// (mpm_amd/csrc/scene.cpp) and the CPU oracle (oracle/oracle_capi.cpp), it
// loads a scene, samples meshes, runs oracle substeps on the sampled
// particles and feeds malformed scene and mesh files to the loaders.
//
//   host_driver <good scene.toml> <good mesh.obj> [bad:<path> | obj:<path> ...]
//
// bad:<path> is a scene file the loader must reject; obj:<path> is a mesh
// the sampler may accept or reject (only memory errors fail). Exit status 0
// when every expectation holds; sanitizer reports abort the run.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpm_hip.h"

extern "C" {
int oracle_advance_f32(const mpm_params* p, const mpm_material* m, int nm, int64_t n, float* x, float* v,
                       float* C, float* F, float* Jp, const uint8_t* mat, int steps, int threads,
                       float* grid_out, int interp);
int oracle_advance_f64(const mpm_params* p, const mpm_material* m, int nm, int64_t n, double* x,
                       double* v, double* C, double* F, double* Jp, const uint8_t* mat, int steps,
                       int threads, double* grid_out, int interp);
int oracle_svd_f32(int64_t n, const float* A, float* U, float* s, float* V);
}

static int failures = 0;
#define EXPECT(cond, ...)                         \
  do {                                            \
    if (!(cond)) {                                \
      std::fprintf(stderr, "FAIL: " __VA_ARGS__); \
      std::fprintf(stderr, "\n");                 \
      ++failures;                                 \
    }                                             \
  } while (0)

template <typename T>
static void run_oracle(const std::vector<mpm_material>& mats, const std::vector<float>& x32,
                       const std::vector<float>& v32, const std::vector<uint8_t>& mat, int threads,
                       int interp) {
  const int64_t n = (int64_t)mat.size();
  std::vector<T> x(x32.begin(), x32.end()), v(v32.begin(), v32.end());
  std::vector<T> C(9 * n, T(0)), F(9 * n, T(0)), Jp(n, T(1));
  for (int64_t i = 0; i < n; ++i) F[9 * i] = F[9 * i + 4] = F[9 * i + 8] = T(1);
  mpm_params p{};
  p.dt = 1e-4f;
  p.N = 32;
  p.gravity = -9.81f;
  p.boundary_mode = MPM_BOUNDARY_FRACTION;
  p.boundary = 0.05f;
  p.compat = MPM_COMPAT_REFERENCE;
  std::vector<T> grid((size_t)p.N * p.N * p.N * 4);
  int rc;
  if constexpr (sizeof(T) == 4)
    rc = oracle_advance_f32(&p, mats.data(), (int)mats.size(), n, x.data(), v.data(), C.data(), F.data(),
                            Jp.data(), mat.data(), 3, threads, grid.data(), interp);
  else
    rc = oracle_advance_f64(&p, mats.data(), (int)mats.size(), n, x.data(), v.data(), C.data(), F.data(),
                            Jp.data(), mat.data(), 3, threads, grid.data(), interp);
  EXPECT(rc == 0, "oracle advance (%zu-byte floats) returned %d", sizeof(T), rc);
  bool finite = true;
  for (const T& a : x) finite = finite && std::isfinite((double)a);
  EXPECT(finite, "oracle advance produced non-finite positions");
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s scene.toml mesh.obj [bad:path | obj:path ...]\n", argv[0]);
    return 2;
  }
  // ---- the good scene: load, query, copy out, run the oracle on it
  mpm_scene sc = nullptr;
  int rc = mpm_scene_load(argv[1], 20000, 1, &sc);
  EXPECT(rc == 0 && sc, "scene load: %d %s", rc, mpm_scene_last_error());
  if (sc) {
    int32_t nm = 0, no = 0;
    int64_t np = 0;
    EXPECT(mpm_scene_info(sc, &nm, &no, &np) == 0 && nm > 0 && no > 0 && np > 0, "scene info");
    std::vector<mpm_material> mats((size_t)nm);
    EXPECT(mpm_scene_materials(sc, mats.data()) == 0, "scene materials");
    int64_t total = 0;
    for (int32_t i = 0; i < no; ++i) {
      int32_t m = -1;
      int64_t first = -1, count = -1;
      double b = 0, e = 0;
      EXPECT(mpm_scene_object(sc, i, &m, &first, &count, &b, &e) == 0, "scene object %d", i);
      EXPECT(first == total && count >= 0 && m >= 0 && m < nm, "object %d range", i);
      total += count;
    }
    EXPECT(total == np, "objects cover the particles");
    int32_t m = 0;
    int64_t first = 0, count = 0;
    double b = 0, e = 0;
    EXPECT(mpm_scene_object(sc, no, &m, &first, &count, &b, &e) != 0, "object index past the end rejected");
    EXPECT(mpm_scene_object(sc, -1, &m, &first, &count, &b, &e) != 0, "negative object index rejected");
    std::vector<float> x(3 * np), v(3 * np);
    std::vector<uint8_t> mat(np);
    EXPECT(mpm_scene_particles_soa(sc, x.data(), v.data(), mat.data()) == 0, "scene particles");
    for (int interp = 0; interp < 2; ++interp) {  // quadratic, cubic
      run_oracle<float>(mats, x, v, mat, 1, interp);
      run_oracle<float>(mats, x, v, mat, 2, interp);
      run_oracle<double>(mats, x, v, mat, 1, interp);
    }
    EXPECT(mpm_scene_destroy(sc) == 0, "scene destroy");
  }
  // ---- the sampler on the good mesh, twice on one stream
  mpm_sampler smp = nullptr;
  EXPECT(mpm_sampler_create(1, &smp) == 0 && smp, "sampler create");
  const float pos[3] = {0.3f, 0.3f, 0.3f};
  for (int k = 0; k < 2 && smp; ++k) {
    int64_t cnt = 0;
    rc = mpm_sampler_fill_mesh(smp, argv[2], 0.3, pos, 1.0f / 20000.0f, &cnt);
    EXPECT(rc == 0 && cnt > 0, "sampler fill: %d %s", rc, mpm_scene_last_error());
    std::vector<float> pts(3 * (size_t)(cnt > 0 ? cnt : 1));
    EXPECT(mpm_sampler_points(smp, pts.data()) == 0, "sampler points");
  }
  // ---- malformed inputs
  for (int i = 3; i < argc; ++i) {
    const std::string a = argv[i];
    if (a.rfind("bad:", 0) == 0) {
      mpm_scene bad = nullptr;
      rc = mpm_scene_load(a.c_str() + 4, 20000, 1, &bad);
      EXPECT(rc != 0, "scene %s accepted", a.c_str() + 4);
      std::printf("%s -> %d (%s)\n", a.c_str(), rc, mpm_scene_last_error());
      if (bad) mpm_scene_destroy(bad);
    } else if (a.rfind("obj:", 0) == 0 && smp) {
      int64_t cnt = 0;
      rc = mpm_sampler_fill_mesh(smp, a.c_str() + 4, 0.3, pos, 1.0f / 20000.0f, &cnt);
      std::printf("%s -> %d count %lld (%s)\n", a.c_str(), rc, (long long)cnt, rc ? mpm_scene_last_error() : "");
      if (rc == 0 && cnt > 0) {
        std::vector<float> pts(3 * (size_t)cnt);
        EXPECT(mpm_sampler_points(smp, pts.data()) == 0, "sampler points after %s", a.c_str());
      }
    }
  }
  if (smp) EXPECT(mpm_sampler_destroy(smp) == 0, "sampler destroy");
  // ---- the oracle's SVD on singular and ordinary matrices (the reference's polar test matrices)
  const float A[3][9] = {{1, 2, 1, 1, 3, 1, 1, 8, 1},
                         {0, 1, 0, -1, 2, -1, -1, 0.001f, -1},
                         {2, 0, 0, 0, 3, 0, 0, 0, 4}};
  float U[27], s[9], V[27];
  EXPECT(oracle_svd_f32(3, &A[0][0], U, s, V) == 0, "oracle svd");
  std::printf("host driver: %d failures\n", failures);
  return failures ? 1 : 0;
}
