// batch.cpp -- many independent scenes advanced together (C-ABI mpm_batch_*).
//
// No reference counterpart: the reference advances one Simulation per process
// (src/main.cu:96-101). A batch holds handles of scenes that advance in lock
// step. Their substeps run as one HIP graph: every scene's kernel chain is
// captured on one of a few streams forked from the batch's origin stream, so
// independent scenes run concurrently on the GPU, and the graph is replayed
// with one launch per capture cycle (two sort intervals of the scenes), with
// no host work per kernel. Substeps before the first cycle boundary, a
// remainder shorter than a cycle, or scenes that cannot be captured (not in
// the default fused tiled mode, timing on) run eagerly on the same streams.
// A scene whose settings, materials, particles or current buffer changed since
// the capture is captured again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpm_hip.h"
#include "engine.hpp"

using mpm_engine::Sim;

struct mpm_batch_s {
  std::mutex mu;
  std::string err;
  std::vector<Sim*> scenes;  // not owned
  std::vector<hipStream_t> streams;
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr, before = nullptr, done = nullptr;
  std::vector<hipEvent_t> joins;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> signature;  // per scene, at capture
  int64_t nstreams = 4;
  int64_t use_graph = 1;
  int64_t launches = 0;  // graph launches so far (for tests and reports)

  int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
  int hipfail(hipError_t e, const char* what) {
    return fail(MPM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    signature.clear();
  }
  ~mpm_batch_s() {
    drop_graph();
    for (auto st : streams) (void)hipStreamDestroy(st);
    for (auto ev : joins) (void)hipEventDestroy(ev);
    if (origin) (void)hipStreamDestroy(origin);
    for (auto ev : {fork, before, done})
      if (ev) (void)hipEventDestroy(ev);
  }
};

#define BCHK(expr, what)                            \
  do {                                              \
    hipError_t e_ = (expr);                         \
    if (e_ != hipSuccess) return b->hipfail(e_, what); \
  } while (0)

namespace {

int ensure_streams(mpm_batch_s* b) {
  const size_t want = (size_t)std::max<int64_t>(1, b->nstreams);
  if (!b->origin) {
    BCHK(hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking), "hipStreamCreate");
    BCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
    BCHK(hipEventCreateWithFlags(&b->before, hipEventDisableTiming), "hipEventCreate");
    BCHK(hipEventCreateWithFlags(&b->done, hipEventDisableTiming), "hipEventCreate");
  }
  while (b->streams.size() < want) {
    hipStream_t st;
    hipEvent_t ev;
    BCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    BCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    b->streams.push_back(st);
    b->joins.push_back(ev);
  }
  return MPM_OK;
}

// Run `n` substeps of every scene on the batch streams, forked from and
// joined back into the origin stream (eagerly, or recorded when capturing).
int run_scenes(mpm_batch_s* b, int32_t n) {
  const size_t ns = std::min(b->streams.size(), b->scenes.size());
  BCHK(hipEventRecord(b->fork, b->origin), "hipEventRecord");
  for (size_t k = 0; k < ns; ++k) BCHK(hipStreamWaitEvent(b->streams[k], b->fork, 0), "hipStreamWaitEvent");
  for (size_t j = 0; j < b->scenes.size(); ++j) {
    Sim* s = b->scenes[j];
    std::lock_guard<std::mutex> lock(s->mutex());
    const int rc = s->advance(n, b->streams[j % ns]);
    if (rc) return b->fail(rc, "scene %zu: %s", j, s->last_error().c_str());
  }
  for (size_t k = 0; k < ns; ++k) {
    BCHK(hipEventRecord(b->joins[k], b->streams[k]), "hipEventRecord");
    BCHK(hipStreamWaitEvent(b->origin, b->joins[k], 0), "hipStreamWaitEvent");
  }
  return MPM_OK;
}

bool all_at_cycle_start(const mpm_batch_s* b, int* cycle) {
  *cycle = 0;
  for (Sim* s : b->scenes) {
    if (!s->at_cycle_start()) return false;
    if (*cycle == 0) *cycle = s->cycle_substeps();
    if (s->cycle_substeps() != *cycle) return false;
  }
  return *cycle > 0;
}

bool signature_matches(const mpm_batch_s* b) {
  if (!b->exec || b->signature.size() != b->scenes.size()) return false;
  for (size_t j = 0; j < b->scenes.size(); ++j)
    if (b->scenes[j]->capture_signature() != b->signature[j]) return false;
  return true;
}

int capture_and_launch(mpm_batch_s* b, int32_t cycle) {
  b->drop_graph();
  std::vector<uint64_t> sig;
  for (Sim* s : b->scenes) sig.push_back(s->capture_signature());
  BCHK(hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
  const int rc = run_scenes(b, cycle);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->origin, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return b->hipfail(e, "hipStreamEndCapture");
  b->graph = g;
  BCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  b->signature = sig;  // the capture ran the scenes' host logic for one cycle
  BCHK(hipGraphLaunch(b->exec, b->origin), "hipGraphLaunch");
  ++b->launches;
  return MPM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" int mpm_batch_create(mpm_batch* out) {
  if (!out) return MPM_ERR_INVALID;
  *out = new mpm_batch_s();
  return MPM_OK;
}

extern "C" int mpm_batch_destroy(mpm_batch b) {
  delete b;
  return MPM_OK;
}

extern "C" const char* mpm_batch_last_error(mpm_batch b) { return b ? b->err.c_str() : ""; }

extern "C" int mpm_batch_add(mpm_batch b, mpm_handle h) {
  if (!b || !h) return MPM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(b->mu);
  b->scenes.push_back(reinterpret_cast<Sim*>(h));
  b->drop_graph();
  return MPM_OK;
}

extern "C" int mpm_batch_count(mpm_batch b, int32_t* out) {
  if (!b || !out) return MPM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(b->mu);
  *out = (int32_t)b->scenes.size();
  return MPM_OK;
}

extern "C" int mpm_batch_set_option(mpm_batch b, const char* name, int64_t v) {
  if (!b || !name) return MPM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(b->mu);
  const std::string n = name;
  if (n == "graph") {
    if (v != 0 && v != 1) return b->fail(MPM_ERR_INVALID, "graph must be 0 or 1");
    b->use_graph = v;
    return MPM_OK;
  }
  if (n == "streams") {
    if (v < 1 || v > 16) return b->fail(MPM_ERR_INVALID, "streams must be 1..16");
    b->nstreams = v;
    b->drop_graph();
    return MPM_OK;
  }
  return b->fail(MPM_ERR_INVALID, "mpm_batch_set_option: unknown option '%s'", name);
}

extern "C" int mpm_batch_get_option(mpm_batch b, const char* name, int64_t* v) {
  if (!b || !name || !v) return MPM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(b->mu);
  const std::string n = name;
  if (n == "graph") *v = b->use_graph;
  else if (n == "streams") *v = b->nstreams;
  else if (n == "graph_launches") *v = b->launches;
  else return b->fail(MPM_ERR_INVALID, "mpm_batch_get_option: unknown option '%s'", name);
  return MPM_OK;
}

extern "C" int mpm_batch_advance(mpm_batch b, int32_t substeps, void* stream) {
  if (!b) return MPM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(b->mu);
  if (substeps < 0) return b->fail(MPM_ERR_INVALID, "mpm_batch_advance: negative substep count");
  if (b->scenes.empty() || substeps == 0) return MPM_OK;
  int rc = ensure_streams(b);
  if (rc) return rc;
  hipStream_t user = (hipStream_t)stream;
  // order after the caller's earlier work on its stream
  BCHK(hipEventRecord(b->before, user), "hipEventRecord");
  BCHK(hipStreamWaitEvent(b->origin, b->before, 0), "hipStreamWaitEvent");
  int32_t left = substeps;
  while (left > 0) {
    int cycle = 0;
    if (b->use_graph && all_at_cycle_start(b, &cycle) && left >= cycle) {
      if (signature_matches(b)) {
        BCHK(hipGraphLaunch(b->exec, b->origin), "hipGraphLaunch");
        ++b->launches;
        for (Sim* s : b->scenes) {
          std::lock_guard<std::mutex> lock2(s->mutex());
          s->replayed(cycle);
        }
      } else {
        rc = capture_and_launch(b, cycle);
        if (rc) return rc;
      }
      left -= cycle;
    } else {
      rc = run_scenes(b, 1);
      if (rc) return rc;
      left -= 1;
    }
  }
  BCHK(hipEventRecord(b->done, b->origin), "hipEventRecord");
  BCHK(hipStreamWaitEvent(user, b->done, 0), "hipStreamWaitEvent");
  return MPM_OK;
}
