// oracle_capi.cpp -- C entry points of the CPU oracle (libmpm_oracle.so).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// CPU-baseline leg of bench.py, never by the engine. Arrays use the same SoA
// layout as the engine's C-ABI (include/mpm_hip.h): x[n][3], v[n][3],
// C[n][9] and F[n][9] row-major, Jp[n], material[n].
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstdint>
#include <vector>

#include "mpm_oracle.hpp"

using namespace oracle;

namespace {

template <class S>
void load(std::vector<Particle<S>>& ps, int64_t n, const S* x, const S* v, const S* C, const S* F,
          const S* Jp, const uint8_t* mat) {
  ps.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    Particle<S>& p = ps[(size_t)i];
    for (int a = 0; a < 3; ++a) { p.x[a] = x[3 * i + a]; p.v[a] = v[3 * i + a]; }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) { p.C[r][c] = C[9 * i + 3 * r + c]; p.F[r][c] = F[9 * i + 3 * r + c]; }
    p.Jp = Jp[i];
    p.mat = mat[i];
  }
}

template <class S>
void store(const std::vector<Particle<S>>& ps, S* x, S* v, S* C, S* F, S* Jp) {
  for (size_t i = 0; i < ps.size(); ++i) {
    const Particle<S>& p = ps[i];
    for (int a = 0; a < 3; ++a) { x[3 * i + a] = p.x[a]; v[3 * i + a] = p.v[a]; }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) { C[9 * i + 3 * r + c] = p.C[r][c]; F[9 * i + 3 * r + c] = p.F[r][c]; }
    Jp[i] = p.Jp;
  }
}

template <class S>
int check_mats(const uint8_t* mat, int64_t n, int nmat) {
  for (int64_t i = 0; i < n; ++i)
    if ((int)mat[i] >= nmat) return 1;
  return 0;
}

template <class S>
int advance(const mpm_params* prm, const mpm_material* mats, int nmat, int64_t n, S* x, S* v, S* C,
            S* F, S* Jp, const uint8_t* mat, int steps, int threads, S* grid_out, int interp) {
  if (!prm || (n > 0 && (!x || !v || !C || !F || !Jp || !mat)) || prm->N < 4 || prm->N > 1024)
    return 1;
  if (interp != INTERP_QUADRATIC && interp != INTERP_CUBIC) return 1;
  if (check_mats<S>(mat, n, nmat)) return 1;
  std::vector<Particle<S>> ps;
  load<S>(ps, n, x, v, C, F, Jp, mat);
  const StepConsts<S> k = make_consts<S>(*prm, interp);
  std::vector<S> grid;
  for (int s = 0; s < steps; ++s) step<S>(ps, mats, nmat, k, grid, threads);
  store<S>(ps, x, v, C, F, Jp);
  if (grid_out && steps > 0)
    for (size_t i = 0; i < grid.size(); ++i) grid_out[i] = grid[i];
  return 0;
}

// All particles of material 0 with per-particle Lame parameters (lame: mu,
// lambda per particle): the forward of the differentiable step.
template <class S>
int advance_lame(const mpm_params* prm, const mpm_material* m, int64_t n, S* x, S* v, S* C, S* F,
                 S* Jp, const S* lame, int steps, int threads) {
  if (!prm || !m || (n > 0 && (!x || !v || !C || !F || !Jp || !lame)) || prm->N < 4 || prm->N > 1024)
    return 1;
  std::vector<uint8_t> mat((size_t)n, 0);
  std::vector<Particle<S>> ps;
  load<S>(ps, n, x, v, C, F, Jp, mat.data());
  const StepConsts<S> k = make_consts<S>(*prm);
  std::vector<S> grid;
  for (int s = 0; s < steps; ++s) step<S>(ps, m, 1, k, grid, threads, lame);
  store<S>(ps, x, v, C, F, Jp);
  return 0;
}

template <class S>
int p2g_only(const mpm_params* prm, const mpm_material* mats, int nmat, int64_t n, const S* x,
             const S* v, const S* C, const S* F, const S* Jp, const uint8_t* mat, S* grid_out,
             int interp) {
  if (!prm || !grid_out || prm->N < 4 || prm->N > 1024) return 1;
  if (interp != INTERP_QUADRATIC && interp != INTERP_CUBIC) return 1;
  if (check_mats<S>(mat, n, nmat)) return 1;
  std::vector<Particle<S>> ps;
  load<S>(ps, n, x, v, C, F, Jp, mat);
  const StepConsts<S> k = make_consts<S>(*prm, interp);
  const size_t nodes = (size_t)prm->N * prm->N * prm->N;
  for (size_t i = 0; i < 4 * nodes; ++i) grid_out[i] = S(0);
  for (int64_t pi = 0; pi < n; ++pi)
    p2g_particle<S>(ps[(size_t)pi], mats[ps[(size_t)pi].mat], k, [&](uint64_t idx, const S c4[4]) {
      for (int w = 0; w < 4; ++w) grid_out[4 * idx + w] += c4[w];
    });
  return 0;
}

template <class S>
int grid_update(const mpm_params* prm, S* grid) {
  if (!prm || !grid) return 1;
  const StepConsts<S> k = make_consts<S>(*prm);
  const uint32_t N = prm->N;
  for (uint32_t i = 0; i < N; ++i)
    for (uint32_t j = 0; j < N; ++j)
      for (uint32_t kk = 0; kk < N; ++kk)
        grid_update_node<S>(grid + 4 * ((size_t)N * N * i + (size_t)N * j + kk), i, j, kk, k);
  return 0;
}

template <class S>
int svd_many(int64_t n, const S* A, S* U, S* sig, S* V) {
  for (int64_t i = 0; i < n; ++i) {
    S a[3][3], u[3][3], s[3], vv[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) a[r][c] = A[9 * i + 3 * r + c];
    svd3<S>(a, u, s, vv);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) { U[9 * i + 3 * r + c] = u[r][c]; V[9 * i + 3 * r + c] = vv[r][c]; }
    for (int a2 = 0; a2 < 3; ++a2) sig[3 * i + a2] = s[a2];
  }
  return 0;
}

template <class S>
int stress_many(const mpm_material* m, int compat, int64_t n, const S* F, const S* Jp, S* PF) {
  for (int64_t i = 0; i < n; ++i) {
    Particle<S> p{};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) p.F[r][c] = F[9 * i + 3 * r + c];
    p.Jp = Jp[i];
    S R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (m->kind != MPM_MATERIAL_WC_LIQUID) polar_R<S>(p.F, R);
    S out[3][3];
    material_stress<S>(p, *m, compat, R, out);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) PF[9 * i + 3 * r + c] = out[r][c];
  }
  return 0;
}

template <class S>
int plasticity_many(const mpm_material* m, int64_t n, S* F, S* Jp) {
  for (int64_t i = 0; i < n; ++i) {
    Particle<S> p{};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) p.F[r][c] = F[9 * i + 3 * r + c];
    p.Jp = Jp[i];
    plasticity<S>(p, *m);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) F[9 * i + 3 * r + c] = p.F[r][c];
    Jp[i] = p.Jp;
  }
  return 0;
}

}  // namespace

extern "C" {

float oracle_rsqrt_f32(float x) { return rsqrt_rn(x); }

int oracle_svd_f32(int64_t n, const float* A, float* U, float* s, float* V) { return svd_many<float>(n, A, U, s, V); }
int oracle_svd_f64(int64_t n, const double* A, double* U, double* s, double* V) { return svd_many<double>(n, A, U, s, V); }

int oracle_stress_f32(const mpm_material* m, int compat, int64_t n, const float* F, const float* Jp, float* PF) {
  return stress_many<float>(m, compat, n, F, Jp, PF);
}
int oracle_stress_f64(const mpm_material* m, int compat, int64_t n, const double* F, const double* Jp, double* PF) {
  return stress_many<double>(m, compat, n, F, Jp, PF);
}
int oracle_plasticity_f32(const mpm_material* m, int64_t n, float* F, float* Jp) { return plasticity_many<float>(m, n, F, Jp); }
int oracle_plasticity_f64(const mpm_material* m, int64_t n, double* F, double* Jp) { return plasticity_many<double>(m, n, F, Jp); }

int oracle_p2g_f32(const mpm_params* p, const mpm_material* m, int nm, int64_t n, const float* x,
                   const float* v, const float* C, const float* F, const float* Jp, const uint8_t* mat,
                   float* grid, int interp) {
  return p2g_only<float>(p, m, nm, n, x, v, C, F, Jp, mat, grid, interp);
}
int oracle_p2g_f64(const mpm_params* p, const mpm_material* m, int nm, int64_t n, const double* x,
                   const double* v, const double* C, const double* F, const double* Jp,
                   const uint8_t* mat, double* grid, int interp) {
  return p2g_only<double>(p, m, nm, n, x, v, C, F, Jp, mat, grid, interp);
}
int oracle_grid_update_f32(const mpm_params* p, float* grid) { return grid_update<float>(p, grid); }
int oracle_grid_update_f64(const mpm_params* p, double* grid) { return grid_update<double>(p, grid); }

int oracle_advance_f32(const mpm_params* p, const mpm_material* m, int nm, int64_t n, float* x, float* v,
                       float* C, float* F, float* Jp, const uint8_t* mat, int steps, int threads,
                       float* grid_out, int interp) {
  return advance<float>(p, m, nm, n, x, v, C, F, Jp, mat, steps, threads, grid_out, interp);
}
int oracle_advance_f64(const mpm_params* p, const mpm_material* m, int nm, int64_t n, double* x,
                       double* v, double* C, double* F, double* Jp, const uint8_t* mat, int steps,
                       int threads, double* grid_out, int interp) {
  return advance<double>(p, m, nm, n, x, v, C, F, Jp, mat, steps, threads, grid_out, interp);
}

// Weights and D^-1 of the selected kernel at n positions (x[n][3]): base[n][3],
// w[n][3][4] (entries past the kernel's size are 0), Dinv[n][3][3] (generic
// inverse for the cubic kernel, the constant 4/dx^2 I for the quadratic one,
// or the generic inverse of either with generic = 1).
int oracle_kernel_f64(const mpm_params* p, int interp, int generic, int64_t n, const double* x,
                      int32_t* base, double* w, double* Dinv) {
  if (!p || p->N < 4 || p->N > 1024 || (interp != INTERP_QUADRATIC && interp != INTERP_CUBIC)) return 1;
  const StepConsts<double> k = make_consts<double>(*p, interp);
  for (int64_t i = 0; i < n; ++i) {
    int b[3];
    double ww[3][4], D[3][3];
    kernel_weights<double>(k, x + 3 * i, b, ww);
    if (generic || interp != INTERP_QUADRATIC) {
      generic_dinv<double>(x + 3 * i, b, ww, kernel_size(interp), k.dx, D);
    } else {
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) D[r][c] = (r == c) ? k.dinv : 0.0;
    }
    for (int a = 0; a < 3; ++a) {
      base[3 * i + a] = b[a];
      for (int o = 0; o < 4; ++o) w[12 * i + 4 * a + o] = ww[a][o];
      for (int c = 0; c < 3; ++c) Dinv[9 * i + 3 * a + c] = D[a][c];
    }
  }
  return 0;
}

int oracle_advance_lame_f32(const mpm_params* p, const mpm_material* m, int64_t n, float* x, float* v,
                            float* C, float* F, float* Jp, const float* lame, int steps, int threads) {
  return advance_lame<float>(p, m, n, x, v, C, F, Jp, lame, steps, threads);
}
int oracle_advance_lame_f64(const mpm_params* p, const mpm_material* m, int64_t n, double* x, double* v,
                            double* C, double* F, double* Jp, const double* lame, int steps, int threads) {
  return advance_lame<double>(p, m, n, x, v, C, F, Jp, lame, steps, threads);
}

// Wall-clock seconds of `steps` FP32 substeps (the CPU baseline leg).
double oracle_time_advance_f32(const mpm_params* p, const mpm_material* m, int nm, int64_t n, float* x,
                               float* v, float* C, float* F, float* Jp, const uint8_t* mat, int steps,
                               int threads) {
  std::vector<Particle<float>> ps;
  load<float>(ps, n, x, v, C, F, Jp, mat);
  const StepConsts<float> k = make_consts<float>(*p);
  std::vector<float> grid;
  auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < steps; ++s) step<float>(ps, m, nm, k, grid, threads);
  auto t1 = std::chrono::steady_clock::now();
  store<float>(ps, x, v, C, F, Jp);
  return std::chrono::duration<double>(t1 - t0).count();
}

int oracle_openmp_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
}
