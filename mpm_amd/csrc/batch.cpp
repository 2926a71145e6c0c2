// batch.cpp -- many independent scenes advanced together (C-ABI mpm_batch_*).
//
// No reference counterpart: the reference advances one Simulation per process
// (src/main.cu:96-101). A batch holds handles of scenes that advance in lock
// step. Their substeps run as one HIP graph: every scene's kernel chain is
// captured on one of a few streams forked from the batch's origin stream, so
// independent scenes run concurrently on the GPU, and the graph is replayed
// with one launch per capture cycle (two sort intervals of the scenes), with
// no host work per kernel. Substeps before the first cycle boundary and a
// remainder shorter than a cycle run eagerly: as the same multi-scene
// launches when every scene is in the fused steady state (below), else as
// each scene's own kernel chain on the same streams (scenes that cannot be
// captured: not in the default fused tiled mode, timing on).
// A scene whose settings, materials, particles or current buffer changed since
// the capture is captured again.
//
// Fused cycles (option "fuse", default on): when every scene is in the steady
// state of the tiled fused mode, a cycle is recorded as one launch per phase
// for many scenes together instead of per-scene kernel chains: per substep a
// tile combine and reduction over the node blocks of the scenes, the due cell
// sorts (in one launch per sort kernel), and a fused G2P/P2G pass over their
// work items (one per kernel: liquid-only scenes and the others apart). The
// scenes are split into a few chains (option "chains", default 4), balanced
// by estimated cost, whose launches go to streams of their own and overlap.
// A workgroup finds its scene in a table of per-scene arguments (the same
// arguments the scene's own launch would take, so the results are the same
// bits). The tables are written into pinned host memory while the cycle is
// recorded and copied to the device once, before the graph's first launch;
// every replay reads them there. An eager fused substep writes its tables
// into one of a few pinned slots and copies them at once on the batch's stream ahead
// of its launches; a slot is written again once the copies from it have run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpm_hip.h"
#include "engine.hpp"
#include "kernels.hpp"

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
  int64_t fuse = 1;
  int64_t multi_sort = 1;  // fused substeps sort all due scenes in one launch per kernel (0: no fusion)
  int64_t chains = 4;      // chains of launches, each on a stream of its own
  int64_t item_share = 100;  // percent of each scene's estimated items given a workgroup of the pass
  int64_t launches = 0;        // graph launches so far (for tests and reports)
  int64_t fused_captures = 0;  // cycles recorded as multi-scene launches
  // argument tables of a fused cycle: pinned host copy and device copy
  uint8_t* h_tables = nullptr;
  uint8_t* d_tables = nullptr;
  size_t tables_cap = 0;
  // argument tables of eager fused substeps: pinned slots used in turn, each
  // with the event recorded after the substep that last copied from it, and
  // one device copy (the copies and launches are ordered on the origin stream)
  static constexpr int kEagerSlots = 4;
  uint8_t* h_eager[kEagerSlots] = {};
  hipEvent_t eager_ev[kEagerSlots] = {};
  bool eager_pending[kEagerSlots] = {};
  uint8_t* d_eager = nullptr;
  size_t eager_cap = 0;
  int eager_next = 0;
  int64_t eager_fused = 0;  // substeps run eagerly as multi-scene launches (for tests and reports)

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
    if (h_tables) (void)hipHostFree(h_tables);
    if (d_tables) (void)hipFree(d_tables);
    if (origin) (void)hipStreamSynchronize(origin);  // the eager copies have read their slots
    for (int i = 0; i < kEagerSlots; ++i) {
      if (h_eager[i]) (void)hipHostFree(h_eager[i]);
      if (eager_ev[i]) (void)hipEventDestroy(eager_ev[i]);
    }
    if (d_eager) (void)hipFree(d_eager);
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

// ---- fused cycles ----------------------------------------------------------
// Writes argument tables into pinned host memory as they are produced;
// identical tables (most substeps between two sorts) share one slot. The
// device copy of the buffer, at the same offsets, is the caller's: one copy of
// [0, used) once every table of a recorded cycle or an eager substep is
// written.
struct TableWriter {
  mpm_batch_s* b;
  uint8_t* host;
  uint8_t* dev;
  size_t cap;
  size_t used = 0;
  std::vector<std::pair<size_t, size_t>> slots;  // (offset, bytes)
  // returns the device address of a table holding `bytes` at `src`, or
  // nullptr (see failed())
  const void* put(const void* src, size_t bytes) {
    for (const auto& sl : slots)
      if (sl.second == bytes && std::memcmp(host + sl.first, src, bytes) == 0) return dev + sl.first;
    const size_t off = (used + 255) & ~(size_t)255;
    if (off + bytes > cap) return nullptr;
    std::memcpy(host + off, src, bytes);
    slots.push_back({off, bytes});
    used = off + bytes;
    return dev + off;
  }
  int failed() const { return b->fail(MPM_ERR_CAPACITY, "batch: argument tables overflow"); }
};

bool fusable(const mpm_batch_s* b) {
  if (!b->fuse) return false;
  for (Sim* s : b->scenes) {
    if (!s->fused_ready()) return false;
    // the fused substep's sorts are multi-scene launches, ordered after the
    // reductions that read the previous items (a scene's own sort, without
    // multi_sort, runs in its own kernel chain)
    if (!b->multi_sort || !s->multi_sort_ok()) return false;
    const int nm = s->material_count();
    if (nm < 1 || nm > MPM_MAX_MATERIALS) return false;
  }
  return true;
}

// table bytes of one substep at most: per chain (four at most) a reduce
// table, up to two pass tables and a sort table, each launch with its prefix
// (the sort: about 20 launches), every table 256-byte aligned
size_t tables_per_substep(size_t ns) {
  return ns * (sizeof(mpmk::ReduceDesc) + 2 * sizeof(mpmk::ItemDesc) + sizeof(mpmk::SortDesc)) +
         112 * (4 * (ns + 1) + 256) + 4 * 256;
}

int ensure_eager_tables(mpm_batch_s* b) {
  const size_t want = tables_per_substep(b->scenes.size()) + 4096;
  if (want <= b->eager_cap) return MPM_OK;
  // the copies queued from the old slots must have run before they go
  BCHK(hipStreamSynchronize(b->origin), "hipStreamSynchronize");
  for (int i = 0; i < mpm_batch_s::kEagerSlots; ++i) {
    if (b->h_eager[i]) (void)hipHostFree(b->h_eager[i]);
    b->h_eager[i] = nullptr;
    b->eager_pending[i] = false;
  }
  if (b->d_eager) (void)hipFree(b->d_eager);
  b->d_eager = nullptr;
  b->eager_cap = 0;
  for (int i = 0; i < mpm_batch_s::kEagerSlots; ++i) {
    BCHK(hipHostMalloc((void**)&b->h_eager[i], want, 0), "hipHostMalloc(batch tables)");
    if (!b->eager_ev[i]) BCHK(hipEventCreateWithFlags(&b->eager_ev[i], hipEventDisableTiming), "hipEventCreate");
  }
  BCHK(hipMalloc((void**)&b->d_eager, want), "hipMalloc(batch tables)");
  b->eager_cap = want;
  return MPM_OK;
}

int ensure_tables(mpm_batch_s* b, int32_t n) {
  const size_t want = tables_per_substep(b->scenes.size()) * (size_t)n + 4096;
  if (want <= b->tables_cap) return MPM_OK;
  if (b->h_tables) (void)hipHostFree(b->h_tables);
  if (b->d_tables) (void)hipFree(b->d_tables);
  b->h_tables = nullptr;
  b->d_tables = nullptr;
  b->tables_cap = 0;
  BCHK(hipHostMalloc((void**)&b->h_tables, want, 0), "hipHostMalloc(batch tables)");
  BCHK(hipMalloc((void**)&b->d_tables, want), "hipMalloc(batch tables)");
  b->tables_cap = want;
  return MPM_OK;
}

// The launches of one substep for a set of scenes, planned before any of
// them is issued: their argument tables are written (run eagerly: copied)
// first, so the launches may go to another stream after a fork.
struct Chain {
  std::vector<size_t> scenes;
  const void* rdt = nullptr;  // reduction descriptors and prefix
  const void* rpt = nullptr;
  uint32_t rtotal = 0;
  const void* sdt = nullptr;  // sort descriptors of the scenes whose sort is due
  int nsort = 0;
  struct SortLaunch { int phase, pass; const void* prefix; uint32_t total; };
  std::vector<SortLaunch> sorts;
  const void* dt[2] = {nullptr, nullptr};  // fused passes: the others, liquid-only scenes
  const void* pt[2] = {nullptr, nullptr};
  int nd[2] = {0, 0};
  uint32_t dtotal[2] = {0, 0};
};

// Plans one substep of the chain's scenes: (1) the tile reductions with the
// grid update, after the tiles of home blocks with several items were summed
// into one (small dense scenes hold dozens of items per block); (2) the cell
// sorts that are due, one launch per sort kernel for all of them (launched
// after the reductions, which read the items the sorts rebuild); (3) the
// fused passes, liquid-only scenes and the others apart. Work a scene
// enqueues while it is planned (its F planes written out) goes to `st`, the
// stream of the chain's launches.
int plan_chain(mpm_batch_s* b, TableWriter& tw, Chain& c, std::vector<mpmk::ItemDesc>& done, hipStream_t st) {
  std::vector<mpmk::ReduceDesc> rd;
  std::vector<uint32_t> rp(1, 0u);
  for (size_t j : c.scenes) {
    rd.push_back(b->scenes[j]->reduce_desc());
    rd.back().combined = 1;
    rp.push_back(rp.back() + (uint32_t)b->scenes[j]->node_blocks());
  }
  c.rdt = tw.put(rd.data(), sizeof(mpmk::ReduceDesc) * rd.size());
  c.rpt = tw.put(rp.data(), sizeof(uint32_t) * rp.size());
  if (!c.rdt || !c.rpt) return tw.failed();
  c.rtotal = rp.back();
  bool any = false;
  for (size_t j : c.scenes) any = any || b->scenes[j]->sort_due();
  if (any) {
    std::vector<mpmk::SortDesc> sd;
    std::vector<Sim*> due;
    for (size_t j : c.scenes) {
      Sim* s = b->scenes[j];
      if (!s->sort_due()) continue;
      mpmk::SortDesc d;
      const int rc = s->sort_desc(d);
      if (rc) return b->fail(rc, "scene %zu: %s", j, s->last_error().c_str());
      sd.push_back(d);
      due.push_back(s);
    }
    c.sdt = tw.put(sd.data(), sizeof(mpmk::SortDesc) * sd.size());
    if (!c.sdt) return tw.failed();
    c.nsort = (int)sd.size();
    int maxp = 0;
    for (const auto& d : sd) maxp = std::max(maxp, d.npasses);
    std::vector<std::pair<int, int>> seq = {{mpmk::SORT_FILL, 0}, {mpmk::SORT_KEYS, 0}};
    for (int p = 0; p < maxp; ++p)
      for (int ph : {(int)mpmk::SORT_HIST, (int)mpmk::SORT_RSCAN, (int)mpmk::SORT_SCATTER}) seq.push_back({ph, p});
    for (int ph = mpmk::SORT_HEADS; ph < mpmk::SORT_NUM_PHASES; ++ph) seq.push_back({ph, 0});
    std::vector<uint32_t> pre(sd.size() + 1);
    for (const auto& pp : seq) {
      pre[0] = 0;
      for (size_t j = 0; j < sd.size(); ++j)
        pre[j + 1] = pre[j] + (uint32_t)mpmk::sort_blocks(sd[j], pp.first, pp.second);
      if (pre.back() == 0) continue;
      const void* pt = tw.put(pre.data(), sizeof(uint32_t) * pre.size());
      if (!pt) return tw.failed();
      c.sorts.push_back({pp.first, pp.second, pt, pre.back()});
    }
    for (Sim* s : due) s->sort_done();
  }
  std::vector<mpmk::ItemDesc> gd[2];
  std::vector<uint32_t> gp[2] = {std::vector<uint32_t>(1, 0u), std::vector<uint32_t>(1, 0u)};
  for (size_t j : c.scenes) {
    Sim* s = b->scenes[j];
    int kinds = 0;
    int64_t launch = 0;
    const int rc = s->fused_pass(s->velocity_grid(), 1, done[j], kinds, launch, st);
    if (rc) return b->fail(rc, "scene %zu: %s", j, s->last_error().c_str());
    const int g = kinds ? 1 : 0;
    gd[g].push_back(done[j]);
    // an estimate of the scene's items, not the bound `launch` (an upload's
    // worth of occupied home blocks per scene leaves most of the bound's
    // workgroups empty, and an empty one holds its LDS through three
    // dependent loads): the pass's workgroups take further items in turn
    (void)launch;
    const int64_t wg = std::max<int64_t>(1, s->batch_workgroups() * b->item_share / 100);
    gp[g].push_back(gp[g].back() + (uint32_t)wg);
  }
  for (int g = 0; g < 2; ++g) {
    if (gd[g].empty()) continue;
    c.dt[g] = tw.put(gd[g].data(), sizeof(mpmk::ItemDesc) * gd[g].size());
    c.pt[g] = tw.put(gp[g].data(), sizeof(uint32_t) * gp[g].size());
    if (!c.dt[g] || !c.pt[g]) return tw.failed();
    c.nd[g] = (int)gd[g].size();
    c.dtotal[g] = gp[g].back();
  }
  return MPM_OK;
}

int launch_chain(mpm_batch_s* b, const Chain& c, hipStream_t st) {
  const int n = (int)c.scenes.size();
  BCHK(mpmk::tile_combine_multi((const mpmk::ReduceDesc*)c.rdt, n, st), "multi-scene tile combine");
  BCHK(mpmk::tile_reduce_multi((const mpmk::ReduceDesc*)c.rdt, (const uint32_t*)c.rpt, n, c.rtotal, st),
       "multi-scene grid update");
  for (const auto& l : c.sorts)
    BCHK(mpmk::sort_multi(l.phase, l.pass, (const mpmk::SortDesc*)c.sdt, (const uint32_t*)l.prefix, c.nsort,
                          l.total, st),
         "multi-scene sort");
  for (int g = 0; g < 2; ++g) {
    if (!c.nd[g]) continue;
    BCHK(mpmk::g2p2g_multi((const mpmk::ItemDesc*)c.dt[g], (const uint32_t*)c.pt[g], c.nd[g], c.dtotal[g], g, st),
         "multi-scene G2P/P2G");
  }
  return MPM_OK;
}

// Scenes split into `chains` chains (at most 4: the origin and 3 batch
// streams), balanced by estimated cost: the scenes largest first, each onto
// the chain of least cost so far, a particle of an SVD material costing 1.7
// of a liquid-only scene's (the fused passes' measured rates on the batched
// workload).
std::vector<std::vector<size_t>> chain_scenes(const mpm_batch_s* b) {
  const size_t nsc = b->scenes.size();
  const int64_t k = std::min<int64_t>({b->chains, (int64_t)b->streams.size() + 1, (int64_t)nsc, 4});
  std::vector<std::vector<size_t>> out;
  if (k <= 1) {
    out.emplace_back();
    for (size_t j = 0; j < nsc; ++j) out.back().push_back(j);
    return out;
  }
  std::vector<double> cost(nsc);
  for (size_t j = 0; j < nsc; ++j)
    cost[j] = (double)b->scenes[j]->count() * (b->scenes[j]->liquid_only() ? 1.0 : 1.7);
  std::vector<size_t> order(nsc);
  for (size_t j = 0; j < nsc; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return cost[a] > cost[c]; });
  out.resize((size_t)k);
  std::vector<double> load((size_t)k, 0.0);
  for (size_t j : order) {
    const size_t q = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
    out[q].push_back(j);
    load[q] += cost[j];
  }
  for (auto& c : out) std::sort(c.begin(), c.end());
  return out;
}

// `n` substeps of every scene as multi-scene launches on the origin stream:
// recorded (the stream is being captured) or, with `eager`, run at once.
// With several chains (chain_scenes) chain c > 0 runs on batch stream c - 1,
// forked from the origin: the chains touch disjoint scenes, and each fills
// the others' gaps and tails (a launch's last, partly occupied round of
// workgroups). A recorded cycle forks once and joins once, so the chains'
// substeps drift apart and one chain's grid update meets the other's fused
// pass; an eager substep joins at its end (the next one's table copy
// overwrites the device tables its launches read).
int run_scenes_fused(mpm_batch_s* b, int32_t n, bool eager, size_t* recorded_bytes = nullptr) {
  const size_t nsc = b->scenes.size();
  // recorded: the tables are copied to the device once, after the capture
  // (capture_and_launch), and stay there for every replay of the cycle
  TableWriter tw{b, b->h_tables, b->d_tables, b->tables_cap};
  const std::vector<std::vector<size_t>> groups = chain_scenes(b);
  const int nchains = (int)groups.size();
  std::vector<mpmk::ItemDesc> done(nsc);
  std::vector<hipStream_t> st((size_t)nchains);
  for (int c = 0; c < nchains; ++c) st[c] = c == 0 ? b->origin : b->streams[c - 1];
  auto fork = [&]() -> int {
    BCHK(hipEventRecord(b->fork, b->origin), "hipEventRecord");
    for (int c = 1; c < nchains; ++c) BCHK(hipStreamWaitEvent(st[c], b->fork, 0), "hipStreamWaitEvent");
    return MPM_OK;
  };
  auto join = [&]() -> int {
    for (int c = 1; c < nchains; ++c) {
      BCHK(hipEventRecord(b->joins[c - 1], st[c]), "hipEventRecord");
      BCHK(hipStreamWaitEvent(b->origin, b->joins[c - 1], 0), "hipStreamWaitEvent");
    }
    return MPM_OK;
  };
  // one substep of every chain (its errors return at once; a recording is
  // joined below whatever they are, so that its capture can end)
  auto substep = [&]() -> int {
    int slot = -1;
    if (eager) {  // the next pinned slot, once the copy queued from it ran
      slot = b->eager_next;
      b->eager_next = (slot + 1) % mpm_batch_s::kEagerSlots;
      if (b->eager_pending[slot]) BCHK(hipEventSynchronize(b->eager_ev[slot]), "hipEventSynchronize");
      b->eager_pending[slot] = false;
      tw = TableWriter{b, b->h_eager[slot], b->d_eager, b->eager_cap};
    }
    std::vector<Chain> chains((size_t)nchains);
    for (int c = 0; c < nchains; ++c) {
      chains[c].scenes = groups[c];
      const int rc = plan_chain(b, tw, chains[c], done, st[c]);
      if (rc) return rc;
    }
    // every table of the substep in one copy, ahead of its launches on the
    // origin stream (the device buffer is shared by the slots: the stream
    // orders each substep's copy after the previous substep's kernels)
    if (eager && tw.used)
      BCHK(hipMemcpyAsync(b->d_eager, b->h_eager[slot], tw.used, hipMemcpyHostToDevice, b->origin),
           "hipMemcpyAsync(batch tables)");
    if (eager && nchains > 1) {
      const int rc = fork();
      if (rc) return rc;
    }
    for (int c = 0; c < nchains; ++c) {
      const int rc = launch_chain(b, chains[c], st[c]);
      if (rc) return rc;
    }
    if (eager && nchains > 1) {
      const int rc = join();
      if (rc) return rc;
    }
    for (size_t j = 0; j < nsc; ++j) {
      b->scenes[j]->fused_pass_done(done[j]);
      b->scenes[j]->substep_done();
    }
    if (eager) {
      BCHK(hipEventRecord(b->eager_ev[slot], b->origin), "hipEventRecord");
      b->eager_pending[slot] = true;
      ++b->eager_fused;
    }
    return MPM_OK;
  };
  const bool once = !eager && nchains > 1;
  int rc = once ? fork() : MPM_OK;
  for (int32_t step = 0; step < n && !rc; ++step) rc = substep();
  if (once) {
    const int rj = join();
    if (!rc) rc = rj;
  }
  if (rc) return rc;
  if (recorded_bytes) *recorded_bytes = eager ? 0 : tw.used;
  return MPM_OK;
}

// Longest cycle recorded into one graph (substeps).
constexpr int64_t kMaxCycle = 1024;

int64_t gcd64(int64_t a, int64_t b) {
  while (b) {
    const int64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// Every scene is at the start of a capture cycle (its next substep sorts).
// Scenes may sort at different intervals: the batch's cycle is the least
// common multiple of their cycles (two sort intervals each), after which every
// scene is back in the same state with the same particle buffer current.
bool all_at_cycle_start(const mpm_batch_s* b, int* cycle) {
  *cycle = 0;
  int64_t L = 1;
  for (Sim* s : b->scenes) {
    if (!s->at_cycle_start()) return false;
    const int64_t c = s->cycle_substeps();
    if (c <= 0) return false;
    L = L / gcd64(L, c) * c;
    if (L > kMaxCycle) return false;
  }
  *cycle = (int)L;
  return L > 0;
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
  const bool fused = fusable(b);
  if (fused) {
    // the previous graph's replays and table copy may still be reading the
    // table buffers that are written (or replaced) next
    BCHK(hipStreamSynchronize(b->origin), "hipStreamSynchronize");
    int rc0 = ensure_tables(b, cycle);
    if (rc0) return rc0;
  }
  std::vector<std::unique_lock<std::mutex>> locks;
  if (fused)
    for (Sim* s : b->scenes) locks.emplace_back(s->mutex());
  BCHK(hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
  size_t table_bytes = 0;
  const int rc = fused ? run_scenes_fused(b, cycle, false, &table_bytes) : run_scenes(b, cycle);
  locks.clear();
  if (fused && !rc) ++b->fused_captures;
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
  // the cycle's argument tables, written while it was recorded: the same for
  // every replay, so one copy (ordered before the launches on the origin stream)
  if (fused && table_bytes)
    BCHK(hipMemcpyAsync(b->d_tables, b->h_tables, table_bytes, hipMemcpyHostToDevice, b->origin),
         "hipMemcpyAsync(batch tables)");
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
  Sim* sim = reinterpret_cast<Sim*>(h);
  if (std::find(b->scenes.begin(), b->scenes.end(), sim) != b->scenes.end())
    return b->fail(MPM_ERR_INVALID, "mpm_batch_add: the handle is in the batch already");
  b->scenes.push_back(sim);
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
  if (n == "fuse") {
    if (v != 0 && v != 1) return b->fail(MPM_ERR_INVALID, "fuse must be 0 or 1");
    b->fuse = v;
    b->drop_graph();
    return MPM_OK;
  }
  if (n == "multi_sort") {
    if (v != 0 && v != 1) return b->fail(MPM_ERR_INVALID, "multi_sort must be 0 or 1");
    b->multi_sort = v;
    b->drop_graph();
    return MPM_OK;
  }
  if (n == "item_share") {
    if (v < 1 || v > 100) return b->fail(MPM_ERR_INVALID, "item_share must be 1..100");
    b->item_share = v;
    b->drop_graph();
    return MPM_OK;
  }
  if (n == "chains") {
    if (v < 1 || v > 4) return b->fail(MPM_ERR_INVALID, "chains must be 1..4");
    b->chains = v;
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
  else if (n == "fuse") *v = b->fuse;
  else if (n == "multi_sort") *v = b->multi_sort;
  else if (n == "chains") *v = b->chains;
  else if (n == "item_share") *v = b->item_share;
  else if (n == "fused_captures") *v = b->fused_captures;
  else if (n == "fused_eager_substeps") *v = b->eager_fused;
  else return b->fail(MPM_ERR_INVALID, "mpm_batch_get_option: unknown option '%s'", name);
  return MPM_OK;
}

namespace {
int advance_locked(mpm_batch_s* b, int32_t substeps, void* stream);
}

extern "C" int mpm_batch_advance(mpm_batch b, int32_t substeps, void* stream) {
  if (!b) return MPM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(b->mu);
  return advance_locked(b, substeps, stream);
}

namespace {
int advance_locked(mpm_batch_s* b, int32_t substeps, void* stream) {
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
    } else if (fusable(b)) {
      rc = ensure_eager_tables(b);
      if (rc) return rc;
      std::vector<std::unique_lock<std::mutex>> locks;
      for (Sim* s : b->scenes) locks.emplace_back(s->mutex());
      rc = run_scenes_fused(b, 1, true);
      if (rc) return rc;
      left -= 1;
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
}  // namespace

// ---------------------------------------------------------------------------
// Batch tuner: candidate settings for the whole batch, each timed from the
// same snapshot of every scene through the batch itself (graph capture and
// multi-scene launches included), the fastest kept.
extern "C" int mpm_batch_autotune(mpm_batch b, int32_t trial_substeps, char* report, int32_t report_len,
                                  void* stream) {
  if (!b) return MPM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(b->mu);
  if (b->scenes.empty()) return b->fail(MPM_ERR_STATE, "mpm_batch_autotune: the batch is empty");
  if (trial_substeps < 1 || trial_substeps > 100000)
    return b->fail(MPM_ERR_INVALID, "mpm_batch_autotune: trial_substeps must be 1..100000");
  hipStream_t user = (hipStream_t)stream;
  const size_t ns = b->scenes.size();
  for (size_t j = 0; j < ns; ++j)
    if (b->scenes[j]->count() <= 0 || !b->scenes[j]->tiled())
      return b->fail(MPM_ERR_STATE,
                     "mpm_batch_autotune: scene %zu is empty or not on the tiled path (quadratic kernel, "
                     "p2g_mode 1)", j);
  // each scene's own choice (its tuner restores its state), then snapshots
  std::vector<mpm_engine::Options> own(ns), saved(ns);
  // on an error before the trials: the snapshots taken so far dropped, the
  // scenes' own settings back
  auto undo = [&](size_t upto, int code) {
    for (size_t j = 0; j < upto; ++j) {
      Sim* s = b->scenes[j];
      std::lock_guard<std::mutex> l2(s->mutex());
      s->snapshot_drop();
      s->set_options(saved[j]);
    }
    return code;
  };
  for (size_t j = 0; j < ns; ++j) {
    Sim* s = b->scenes[j];
    std::unique_lock<std::mutex> l2(s->mutex());
    saved[j] = s->options();
    const int rc = s->autotune(trial_substeps, nullptr, user);
    if (rc) {
      const int code = b->fail(rc, "scene %zu: %s", j, s->last_error().c_str());
      l2.unlock();
      return undo(j, code);
    }
    own[j] = s->options();
    own[j].radix_bits = 8;  // the batch sorts every scene with 8-bit digits (the same order)
    const int rs = s->snapshot_take(user);
    if (rs) {
      const int code = b->fail(rs, "scene %zu: %s", j, s->last_error().c_str());
      s->set_options(saved[j]);
      l2.unlock();
      return undo(j, code);
    }
  }
  struct Cand { int64_t sort_every, chunk; bool own; double us; };
  std::vector<Cand> cands = {{0, 0, true, 0.0}};
  for (int64_t se : {16, 24, 32, 48, 64, 96, 128}) cands.push_back({se, mpmk::kMaxItemParticles, false, 0.0});
  for (int64_t se : {32, 64}) cands.push_back({se, 128, false, 0.0});
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return undo(ns, b->fail(MPM_ERR_HIP, "mpm_batch_autotune: hipEventCreate"));
  }
  auto apply = [&](const Cand& c) {
    for (size_t j = 0; j < ns; ++j) {
      Sim* s = b->scenes[j];
      std::lock_guard<std::mutex> l2(s->mutex());
      mpm_engine::Options o = c.own ? own[j] : saved[j];
      if (!c.own) {
        o.sort_every = c.sort_every;
        o.chunk = c.chunk == mpmk::kMaxItemParticles ? 0 : c.chunk;
        o.radix_bits = 8;
      }
      o.p2g_mode = 1;
      o.fused = 1;
      o.timing = 0;
      s->set_options(o);
    }
  };
  auto restore_all = [&]() -> int {
    for (size_t j = 0; j < ns; ++j) {
      Sim* s = b->scenes[j];
      std::lock_guard<std::mutex> l2(s->mutex());
      const hipError_t e = s->snapshot_restore(user);
      if (e != hipSuccess) return b->hipfail(e, "mpm_batch_autotune: restore");
    }
    b->drop_graph();
    return MPM_OK;
  };
  int rc = MPM_OK;
  for (Cand& c : cands) {
    rc = restore_all();
    if (rc) break;
    apply(c);
    // the cycle: least common multiple of the scenes' two sort intervals
    int64_t L = 1;
    for (Sim* s : b->scenes) {
      const int64_t cs = s->cycle_substeps();
      L = L / gcd64(L, cs) * cs;
    }
    if (L > kMaxCycle) {
      c.us = 1e30;  // never recorded as one graph: not a candidate
      continue;
    }
    // to the first common cycle start, and through one recorded cycle
    rc = advance_locked(b, (int32_t)(L / 2 + L), user);
    if (rc) break;
    const int32_t n = (int32_t)(L * ((std::max<int64_t>(trial_substeps, 2 * L) + L - 1) / L));
    // (errors leave the loop: the scenes are restored below)
    hipError_t e = hipEventRecord(e0, user);
    if (e != hipSuccess) { rc = b->hipfail(e, "hipEventRecord"); break; }
    rc = advance_locked(b, n, user);
    if (rc) break;
    float ms = 0.0f;
    e = hipEventRecord(e1, user);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) { rc = b->hipfail(e, "mpm_batch_autotune: timing"); break; }
    c.us = 1e3 * (double)ms / n;
  }
  const int rr = restore_all();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (size_t j = 0; j < ns; ++j) {
    Sim* s = b->scenes[j];
    std::lock_guard<std::mutex> l2(s->mutex());
    s->snapshot_drop();
    if (rc || rr) s->set_options(saved[j]);
  }
  if (rc || rr) return rc ? rc : rr;
  size_t best = 0;
  for (size_t i = 1; i < cands.size(); ++i)
    if (cands[i].us < cands[best].us) best = i;
  apply(cands[best]);
  if (report && report_len > 0) {
    std::string r = "{\"choice\": ";
    char buf[160];
    auto desc = [](const Cand& c) {
      if (c.own) return std::string("\"own\"");
      char d[96];
      std::snprintf(d, sizeof(d), "{\"sort_every\": %lld, \"chunk\": %lld}", (long long)c.sort_every,
                    (long long)c.chunk);
      return std::string(d);
    };
    r += desc(cands[best]) + ", \"trials\": [";
    for (size_t i = 0; i < cands.size(); ++i) {
      r += i ? ", {\"setting\": " : "{\"setting\": ";
      r += desc(cands[i]);
      std::snprintf(buf, sizeof(buf), ", \"us_per_substep\": %.3f}", cands[i].us < 1e29 ? cands[i].us : -1.0);
      r += buf;
    }
    r += "], \"own\": [";
    for (size_t j = 0; j < ns; ++j) {
      std::snprintf(buf, sizeof(buf), "%s[%lld, %lld]", j ? ", " : "", (long long)own[j].sort_every,
                    (long long)(own[j].chunk ? own[j].chunk : mpmk::kMaxItemParticles));
      r += buf;
    }
    r += "]}";
    std::snprintf(report, (size_t)report_len, "%s", r.c_str());
  }
  return MPM_OK;
}
