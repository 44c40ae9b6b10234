// Batched sweeps of many independent grids: one launch per (colour pass,
// kernel group) covers the blocks of that colour in every grid of the group,
// and the whole run (sweeps x 8 colours x groups) is captured once into a
// HIP graph and replayed as a single graph launch. Grids are grouped by
// block size, storage format and arithmetic mode (one kernel instantiation
// each); the autotuner picks the voxel-column variant of each group by
// timing it on the group's own blocks.
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
    std::vector<int> grids;  // entry indices
    DevBuf<int2> items;
    int64_t off[9] = {0};
  };
  std::vector<Entry> entries;
  std::vector<Group> groups;
  std::vector<SweepGrid> hdesc;   // descriptors as built (host copy)
  std::vector<uint64_t> topo;     // each grid's topology version at build
  DevBuf<SweepGrid> d_desc;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_sweeps = -1;
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

int enqueue(msbg_batch *b, int sweeps) {
  for (int it = 0; it < sweeps; it++)
    for (int pass = 0; pass < 8; pass++)
      for (auto &G : b->groups) {
        int64_t n = G.off[pass + 1] - G.off[pass];
        if (!n) continue;
        MSBG_TRY(launch_sweep_items(b->stream, G.bsx, G.storage, G.exact, G.c01, G.tile, b->d_desc.p,
                                    G.items.p + G.off[pass], n, nullptr, it));
      }
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
    for (size_t i = 0; i < spans.size(); i++)
      MSBG_HIP(hipMemcpyAsync(spans[i].p, backup[i].p, spans[i].n, hipMemcpyDeviceToDevice,
                              b->stream));
    MSBG_HIP(hipStreamSynchronize(b->stream));
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
  MSBG_CHECK(b && numIter >= 0, MSBG_EINVAL, "msbg_batch_run: bad args");
  MSBG_TRY(build(b));
  int sweeps = numIter + (numIter % 2);  // msbg4.cpp:150-152
  for (auto &E : b->entries) MSBG_HIP(hipStreamSynchronize(E.g->stream));
  if (useGraph && sweeps > 0) {
    if (b->graph_sweeps != sweeps) {
      drop_graph(b);
      MSBG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = enqueue(b, sweeps);
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
    }
    MSBG_HIP(hipEventRecord(b->ev[0], b->stream));
    MSBG_HIP(hipGraphLaunch(b->exec, b->stream));
    MSBG_HIP(hipEventRecord(b->ev[1], b->stream));
  } else {
    MSBG_HIP(hipEventRecord(b->ev[0], b->stream));
    MSBG_TRY(enqueue(b, sweeps));
    MSBG_HIP(hipEventRecord(b->ev[1], b->stream));
  }
  MSBG_HIP(hipEventSynchronize(b->ev[1]));
  float t = 0;
  MSBG_HIP(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
  b->last_ms = t;
  if (ms) *ms = t;
  return 0;
}

int msbg_set_tile(msbg_grid *g, int tile) {
  MSBG_CHECK(g && tile >= -1 && tile < kNumTiles, MSBG_EINVAL, "msbg_set_tile: bad args");
  g->tile[g->lg0] = tile;
  return 0;
}

}  // extern "C"
