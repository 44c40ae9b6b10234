// Batched sweeps of many independent grids, grouped by block size, storage
// format and arithmetic mode (one kernel instantiation each). A group runs
// either as one launch per colour pass covering the blocks of that colour in
// every grid of the group, or as one persistent launch whose workgroups take
// (sweep, colour, grid, block) items from a work queue and wait per grid for
// its previous colour pass (k_sweep_queue). The whole run is captured once
// into a HIP graph and replayed as a single graph launch. The autotuner picks
// each group's mapping (per-pass launches or the persistent launch) and its
// voxel-column variant by timing them on the group's own blocks.
//
// Each grid keeps the reference's per-grid semantics (colours 0..7 in order,
// level-0 blocks, Neumann halos, numIter rounded up to even, msbg4.cpp:60-957);
// independent grids simply share launches.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "internal.h"

using namespace msbg;

struct msbg_batch {
  struct Entry {
    msbg_grid *g;
    int chan;
    float dt;
    int c01;
  };
  struct Group {
    int bsx = 0, storage = 0;
    bool exact = false;
    bool c01 = false;  // every member clamps to [0,1]
    int tile = -1;
    bool persistent = false;  // tuned: the persistent launch
    std::vector<int> grids;  // entry indices
    DevBuf<int2> items;
    int64_t off[9] = {0};
  };
  std::vector<Entry> entries;
  std::vector<Group> groups;
  std::vector<SweepGrid> hdesc;   // descriptors as built (host copy)
  std::vector<uint64_t> topo;     // each grid's topology version at build
  DevBuf<SweepGrid> d_desc;
  // persistent runs: per entry its item counts before each colour ([e][9],
  // [e][8] = all), and the queue state: done counts per entry, the head of
  // each group, the failure flag
  DevBuf<int> d_pre;
  DevBuf<unsigned> d_qstate;
  uint64_t wait_us = 2000000;  // bound of one dependency wait
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_sweeps = -1;
  int graph_mode = -1;
  bool built = false;
  double last_ms = 0;
};

namespace {

void drop_graph(msbg_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_sweeps = -1;
  b->graph_mode = -1;
}

bool queue_ok(const msbg_batch::Group &G) {
  return (G.bsx == 16 || G.bsx == 32) && G.storage != MSBG_STORAGE_F64;
}

// queue state: done counts (one line per entry), a head per group, the flag
size_t qstate_words(size_t nEntries, size_t nGroups) {
  return nEntries * kDoneStride + nGroups + 1;
}

// a member grid changed its block map, page pool or page contents since
// the batch was built: rebuild descriptors, items and side-cars
bool stale(msbg_batch *b) {
  if (!b->built) return true;
  for (size_t e = 0; e < b->entries.size(); e++) {
    auto &E = b->entries[e];
    auto it = E.g->channels.find(E.chan);
    if (it == E.g->channels.end()) return true;
    ChanLevel &cl = it->second.lv[0];
    if ((cl.npages && !cl.side_valid) || cl.d_pages.p != b->hdesc[e].pages || cl.d_side.p != b->hdesc[e].side ||
        E.g->topo_version != b->topo[e])
      return true;
  }
  return false;
}

int build(msbg_batch *b) {
  if (!stale(b)) return 0;
  b->built = false;
  drop_graph(b);
  std::map<std::tuple<int, int, bool, bool>, int> gi;
  std::vector<SweepGrid> desc(b->entries.size());
  std::vector<std::vector<std::vector<int2>>> per;  // group -> colour -> items
  std::vector<msbg_batch::Group> groups;
  for (size_t e = 0; e < b->entries.size(); e++) {
    auto &E = b->entries[e];
    msbg_grid *g = E.g;
    auto it = g->channels.find(E.chan);
    MSBG_CHECK(it != g->channels.end() && it->second.lv[0].defined, MSBG_ESTATE,
               "batch entry %zu: channel %d has no level 0", e, E.chan);
    Channel &ch = it->second;
    ChanLevel &cl = ch.lv[0];
    MSBG_HIP(hipStreamSynchronize(g->stream));
    MSBG_TRY(ensure_nbr(g, cl, 0));
    if (cl.npages) MSBG_TRY(sweep_prepare(g, ch));
    desc[e] = sweep_desc(ch, E.dt, E.c01);
    bool exact = g->arith == MSBG_ARITH_EXACT || ch.storage == MSBG_STORAGE_F64;
    const bool c01 = E.c01 != 0;
    auto key = std::make_tuple(g->bsx0, ch.storage, exact, c01);
    if (!gi.count(key)) {
      gi[key] = (int)groups.size();
      groups.emplace_back();
      groups.back().bsx = g->bsx0;
      groups.back().storage = ch.storage;
      groups.back().exact = exact;
      groups.back().c01 = c01;
      per.emplace_back(8);
    }
    int k = gi[key];
    groups[k].grids.push_back((int)e);
    // every level-0 value block of the channel, by colour (msbg4.cpp:456-470)
    for (int64_t bid = 0; bid < g->nblocks; bid++) {
      int32_t pg = cl.h_page[bid];
      if (pg < 0 || g->h_level[bid] != 0) continue;
      int bx = (int)(bid % g->nbx), by = (int)((bid / g->nbx) % g->nby),
          bz = (int)(bid / ((int64_t)g->nbx * g->nby));
      per[k][(bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2)].push_back(make_int2((int)e, pg));
    }
  }
  for (size_t k = 0; k < groups.size(); k++) {
    // keep tiles chosen earlier for the same kernel group
    for (auto &old : b->groups)
      if (old.bsx == groups[k].bsx && old.storage == groups[k].storage &&
          old.exact == groups[k].exact)
        groups[k].tile = old.tile;
    std::vector<int2> flat;
    for (int c = 0; c < 8; c++) {
      groups[k].off[c] = (int64_t)flat.size();
      flat.insert(flat.end(), per[k][c].begin(), per[k][c].end());
    }
    groups[k].off[8] = (int64_t)flat.size();
    if (!flat.empty()) {
      MSBG_TRY(groups[k].items.alloc(flat.size()));
      MSBG_HIP(hipMemcpy(groups[k].items.p, flat.data(), sizeof(int2) * flat.size(),
                         hipMemcpyHostToDevice));
    }
  }
  // per entry: its items of each colour, as prefix counts
  std::vector<int> pre(b->entries.size() * 9, 0);
  for (size_t k = 0; k < groups.size(); k++)
    for (int c = 0; c < 8; c++)
      for (const int2 &it : per[k][c]) pre[(size_t)it.x * 9 + c + 1]++;
  for (size_t e = 0; e < b->entries.size(); e++)
    for (int c = 0; c < 8; c++) pre[e * 9 + c + 1] += pre[e * 9 + c];
  if (!pre.empty()) {
    MSBG_TRY(b->d_pre.alloc(pre.size()));
    MSBG_HIP(hipMemcpy(b->d_pre.p, pre.data(), sizeof(int) * pre.size(), hipMemcpyHostToDevice));
  }
  MSBG_TRY(b->d_qstate.ensure(qstate_words(b->entries.size(), groups.size())));
  b->groups.swap(groups);
  b->hdesc = desc;
  b->topo.clear();
  for (auto &E : b->entries) b->topo.push_back(E.g->topo_version);
  if (!desc.empty()) {
    MSBG_TRY(b->d_desc.alloc(desc.size()));
    MSBG_HIP(hipMemcpy(b->d_desc.p, desc.data(), sizeof(SweepGrid) * desc.size(),
                       hipMemcpyHostToDevice));
  }
  b->built = true;
  return 0;
}

// whether group G runs persistently in mode `mode`
bool runs_persistent(const msbg_batch::Group &G, int mode) {
  if (!queue_ok(G)) return false;
  return mode == MSBG_BATCH_PERSISTENT || (mode == MSBG_BATCH_GRAPH && G.persistent);
}

// the queue of group k for a run of `sweeps` sweeps
BatchQueue queue_of(msbg_batch *b, size_t k, int sweeps) {
  const auto &G = b->groups[k];
  const size_t ne = b->entries.size();
  BatchQueue q{};
  q.items = G.items.p;
  q.pre = b->d_pre.p;
  q.done = b->d_qstate.p;
  q.head = b->d_qstate.p + ne * kDoneStride + k;
  q.fail = b->d_qstate.p + ne * kDoneStride + b->groups.size();
  for (int c = 0; c < 9; c++) q.off[c] = (int)G.off[c];
  q.sweeps = sweeps;
  q.waitTicks = wall_ticks(b->wait_us);
  return q;
}

int enqueue(msbg_batch *b, int sweeps, int mode) {
  bool anyQ = false;
  for (auto &G : b->groups) anyQ |= runs_persistent(G, mode) && G.off[8] > 0;
  if (anyQ)
    MSBG_TRY(launch_queue_reset(b->stream, b->d_qstate.p,
                                (int)qstate_words(b->entries.size(), b->groups.size())));
  for (size_t k = 0; k < b->groups.size(); k++) {
    auto &G = b->groups[k];
    if (!runs_persistent(G, mode) || G.off[8] == 0) continue;
    MSBG_TRY(launch_sweep_queue(b->stream, G.bsx, G.storage, G.exact, G.c01, b->d_desc.p,
                                queue_of(b, k, sweeps)));
  }
  for (int it = 0; it < sweeps; it++)
    for (int pass = 0; pass < 8; pass++)
      for (auto &G : b->groups) {
        if (runs_persistent(G, mode)) continue;
        int64_t n = G.off[pass + 1] - G.off[pass];
        if (!n) continue;
        MSBG_TRY(launch_sweep_items(b->stream, G.bsx, G.storage, G.exact, G.c01, G.tile, b->d_desc.p,
                                    G.items.p + G.off[pass], n, nullptr, it));
      }
  return 0;
}

// after a run that held persistent launches: did a dependency wait give up?
int check_queue(msbg_batch *b) {
  unsigned fail = 0;
  MSBG_HIP(hipMemcpy(&fail, b->d_qstate.p + b->entries.size() * kDoneStride + b->groups.size(),
                     sizeof fail, hipMemcpyDeviceToHost));
  MSBG_CHECK(!fail, MSBG_EDEVICE,
             "persistent batch run: a dependency wait ran past its bound of %llu us; the "
             "results of this run are not valid",
             (unsigned long long)b->wait_us);
  return 0;
}

// bytes of the pages and side-cars of the grids of one group (for backup)
struct Span {
  uint8_t *p;
  size_t n;
};

std::vector<Span> group_spans(msbg_batch *b, const msbg_batch::Group &G) {
  std::vector<Span> v;
  for (int e : G.grids) {
    auto &E = b->entries[e];
    ChanLevel &cl = E.g->channels[E.chan].lv[0];
    if (cl.d_pages.p) v.push_back({cl.d_pages.p, cl.d_pages.n});
    if (cl.d_side.p) v.push_back({cl.d_side.p, cl.d_side.n});
  }
  return v;
}

}  // namespace

extern "C" {

int msbg_batch_create(msbg_batch **out) {
  MSBG_CHECK(out, MSBG_EINVAL, "msbg_batch_create: out is NULL");
  msbg_batch *b = new msbg_batch();
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess) {
    delete b;
    set_error("msbg_batch_create: stream/event creation failed");
    return MSBG_EDEVICE;
  }
  *out = b;
  return 0;
}

int msbg_batch_destroy(msbg_batch *b) {
  if (!b) return 0;
  (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  if (b->ev[0]) (void)hipEventDestroy(b->ev[0]);
  if (b->ev[1]) (void)hipEventDestroy(b->ev[1]);
  hipStream_t s = b->stream;
  delete b;
  if (s) (void)hipStreamDestroy(s);
  return 0;
}

int msbg_batch_add(msbg_batch *b, msbg_grid *g, int chan, float timeStep, int doConstrZeroOne) {
  MSBG_CHECK(b && g, MSBG_EINVAL, "msbg_batch_add: bad args");
  MSBG_CHECK(timeStep >= 0 && timeStep == timeStep, MSBG_EUNSUPP, "timeStep must be >= 0");
  auto it = g->channels.find(chan);
  MSBG_CHECK(it != g->channels.end() && it->second.lv[0].defined, MSBG_ESTATE,
             "msbg_batch_add: channel %d has no level 0", chan);
  for (auto &E : b->entries)
    MSBG_CHECK(!(E.g == g && E.chan == chan), MSBG_EINVAL,
               "msbg_batch_add: grid/channel added twice");
  b->entries.push_back({g, chan, timeStep, doConstrZeroOne});
  b->built = false;
  return 0;
}

int msbg_batch_autotune(msbg_batch *b, int passes) {
  MSBG_CHECK(b && passes >= 1, MSBG_EINVAL, "msbg_batch_autotune: bad args");
  MSBG_TRY(build(b));
  for (auto &G : b->groups) {
    if (G.exact || G.off[8] == 0) continue;
    // back up the pages the timing sweeps overwrite, restore afterwards
    std::vector<Span> spans = group_spans(b, G);
    std::vector<DevBuf<uint8_t>> backup(spans.size());
    for (size_t i = 0; i < spans.size(); i++) {
      MSBG_TRY(backup[i].alloc(spans[i].n));
      MSBG_HIP(hipMemcpyAsync(backup[i].p, spans[i].p, spans[i].n, hipMemcpyDeviceToDevice,
                              b->stream));
    }
    int best = -1;
    float bestMs = 0;
    for (int tile = 0; tile < kNumTiles; tile++) {
      // one untimed + `passes` timed launches of every colour
      for (int rep = 0; rep <= passes; rep++) {
        if (rep == 1) MSBG_HIP(hipEventRecord(b->ev[0], b->stream));
        for (int pass = 0; pass < 8; pass++) {
          int64_t n = G.off[pass + 1] - G.off[pass];
          if (n)
            MSBG_TRY(launch_sweep_items(b->stream, G.bsx, G.storage, false, G.c01, tile, b->d_desc.p,
                                        G.items.p + G.off[pass], n, nullptr, rep));
        }
      }
      MSBG_HIP(hipEventRecord(b->ev[1], b->stream));
      MSBG_HIP(hipEventSynchronize(b->ev[1]));
      float ms = 0;
      MSBG_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
      if (best < 0 || ms < bestMs) {
        best = tile;
        bestMs = ms;
      }
    }
    // the persistent launch over `passes` sweeps (one wave per block)
    bool bestQ = false;
    if (queue_ok(G)) {
      const size_t k = (size_t)(&G - b->groups.data());
      for (int rep = 0; rep < 2; rep++) {  // untimed, then timed
        MSBG_TRY(launch_queue_reset(b->stream, b->d_qstate.p,
                                    (int)qstate_words(b->entries.size(), b->groups.size())));
        if (rep == 1) MSBG_HIP(hipEventRecord(b->ev[0], b->stream));
        MSBG_TRY(launch_sweep_queue(b->stream, G.bsx, G.storage, false, G.c01, b->d_desc.p,
                                    queue_of(b, k, passes)));
      }
      MSBG_HIP(hipEventRecord(b->ev[1], b->stream));
      MSBG_HIP(hipEventSynchronize(b->ev[1]));
      MSBG_TRY(check_queue(b));
      float ms = 0;
      MSBG_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
      bestQ = ms < bestMs;
    }
    for (size_t i = 0; i < spans.size(); i++)
      MSBG_HIP(hipMemcpyAsync(spans[i].p, backup[i].p, spans[i].n, hipMemcpyDeviceToDevice,
                              b->stream));
    MSBG_HIP(hipStreamSynchronize(b->stream));
    G.persistent = bestQ;
    G.tile = best;
    for (int e : G.grids) b->entries[e].g->tile[b->entries[e].g->lg0] = best;
  }
  drop_graph(b);
  return 0;
}

int msbg_batch_get_tiles(msbg_batch *b, int *bsx, int *storage, int *tile, int cap, int *n) {
  MSBG_CHECK(b, MSBG_EINVAL, "msbg_batch_get_tiles: batch is NULL");
  MSBG_TRY(build(b));
  int k = 0;
  for (auto &G : b->groups) {
    if (k < cap) {
      if (bsx) bsx[k] = G.bsx;
      if (storage) storage[k] = G.storage;
      if (tile) {
        // the default follows the largest colour pass of the group
        int64_t nmax = 0;
        for (int c = 0; c < 8; c++) nmax = std::max<int64_t>(nmax, G.off[c + 1] - G.off[c]);
        tile[k] = G.tile < 0 ? default_tile(G.bsx, G.storage, nmax) : G.tile;
      }
    }
    k++;
  }
  if (n) *n = k;
  return 0;
}

int msbg_batch_run(msbg_batch *b, int numIter, int useGraph, double *ms) {
  MSBG_CHECK(b && numIter >= 0 && useGraph >= MSBG_BATCH_EAGER &&
                 useGraph <= MSBG_BATCH_PERSISTENT,
             MSBG_EINVAL, "msbg_batch_run: bad args");
  MSBG_TRY(build(b));
  int sweeps = numIter + (numIter % 2);  // msbg4.cpp:150-152
  for (auto &E : b->entries) MSBG_HIP(hipStreamSynchronize(E.g->stream));
  const int mode = useGraph;
  bool anyQ = false;
  for (auto &G : b->groups) anyQ |= runs_persistent(G, mode) && G.off[8] > 0 && sweeps > 0;
  if (mode != MSBG_BATCH_EAGER && sweeps > 0) {
    if (b->graph_sweeps != sweeps || b->graph_mode != mode) {
      drop_graph(b);
      MSBG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = enqueue(b, sweeps, mode);
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(b->stream, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      MSBG_HIP(e);
      b->graph = graph;
      MSBG_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_sweeps = sweeps;
      b->graph_mode = mode;
    }
    MSBG_HIP(hipEventRecord(b->ev[0], b->stream));
    MSBG_HIP(hipGraphLaunch(b->exec, b->stream));
    MSBG_HIP(hipEventRecord(b->ev[1], b->stream));
  } else {
    MSBG_HIP(hipEventRecord(b->ev[0], b->stream));
    MSBG_TRY(enqueue(b, sweeps, mode));
    MSBG_HIP(hipEventRecord(b->ev[1], b->stream));
  }
  MSBG_HIP(hipEventSynchronize(b->ev[1]));
  float t = 0;
  MSBG_HIP(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
  b->last_ms = t;
  if (ms) *ms = t;
  if (anyQ) MSBG_TRY(check_queue(b));
  return 0;
}

int msbg_batch_set_wait_limit(msbg_batch *b, unsigned microseconds) {
  MSBG_CHECK(b, MSBG_EINVAL, "msbg_batch_set_wait_limit: batch is NULL");
  if (b->wait_us != microseconds) drop_graph(b);  // the bound is a kernel argument
  b->wait_us = microseconds;
  return 0;
}

int msbg_batch_get_modes(msbg_batch *b, int *persistent, int cap, int *n) {
  MSBG_CHECK(b, MSBG_EINVAL, "msbg_batch_get_modes: batch is NULL");
  MSBG_TRY(build(b));
  int k = 0;
  for (auto &G : b->groups) {
    if (k < cap && persistent) persistent[k] = G.persistent ? 1 : 0;
    k++;
  }
  if (n) *n = k;
  return 0;
}

int msbg_set_tile(msbg_grid *g, int tile) {
  MSBG_CHECK(g && tile >= -1 && tile < kNumTiles, MSBG_EINVAL, "msbg_set_tile: bad args");
  g->tile[g->lg0] = tile;
  return 0;
}

}  // extern "C"
