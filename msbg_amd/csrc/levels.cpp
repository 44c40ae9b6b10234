// Coarse-level sweeps: the 8-colour MCF sweep applied to the blocks of one
// resolution level on that level's own grid. The reference sweeps level 0
// only (msbg4.cpp:467-470); this entry point extends the same sweep to the
// coarser levels of a multiresolution channel.
//
// The halo of a level-l block follows fillHaloBlock_ with the caller's
// option (halo.cpp:267-479, halo.h:267-355):
//  - OPT_BC_NEUMANN: NULL neighbours and the outside of the domain give the
//    block's own clamped voxel (the level-0 rule);
//  - OPT_BC_COARSE_LEVEL [| OPT_INTERPOLATE]: NULL neighbours read level l+1
//    (its voxel at x/2, or interpolateGhost at (x+.5)/2), and halo voxels
//    outside the domain are clipped into it before they are read.
// Finer neighbours are read from the ghost pages prepareGhostBlocks keeps at
// level l; they are value pages there and are never swept.
//
// GPU design: the level's page pool is followed by "ghost pages" for every
// NULL block next to a swept block, each filled once per call with the
// coarse-level values of its voxels (level l+1 does not change during the
// call). The pool gets the room for them once, at set-up; a call then lends
// the pool to the working level (its own block map and halo table) and
// sweeps it in place. The halo table points at the ghost pages, so the
// unchanged sweep kernel reads coarse-level halos like neighbour pages;
// out-of-domain cells are clipped in the kernel's ring setup.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

using namespace msbg;

extern "C" {

int msbg_apply_channel_pde_level(msbg_grid *g, int chan, int level, unsigned haloOpt,
                                 const int32_t *blockList, int64_t nList, int numIter,
                                 float timeStep, int doConstrZeroOne, double *residual) {
  MSBG_CHECK(g && level >= 0 && level < g->nlevels, MSBG_EINVAL,
             "msbg_apply_channel_pde_level: bad grid or level %d", level);
  MSBG_CHECK(numIter >= 0 && timeStep >= 0.0f && nList >= 0, MSBG_EUNSUPP,
             "msbg_apply_channel_pde_level: numIter %d, timeStep %g", numIter, timeStep);
  const bool coarse = (haloOpt & MSBG_OPT_BC_COARSE_LEVEL) != 0;
  MSBG_CHECK(haloOpt == MSBG_OPT_BC_NEUMANN || haloOpt == MSBG_OPT_BC_COARSE_LEVEL ||
                 haloOpt == (MSBG_OPT_BC_COARSE_LEVEL | MSBG_OPT_INTERPOLATE),
             MSBG_EINVAL, "msbg_apply_channel_pde_level: halo option 0x%x", haloOpt);
  auto it = g->channels.find(chan);
  MSBG_CHECK(it != g->channels.end() && it->second.lv[level].defined, MSBG_ESTATE,
             "channel %d level %d not defined", chan, level);
  Channel &ch = it->second;
  ChanLevel &cl = ch.lv[level];
  MSBG_CHECK(cl.bsx == 8 || cl.bsx == 16 || cl.bsx == 32, MSBG_EUNSUPP,
             "level %d block size %d: the sweep kernel runs 8, 16 and 32", level, cl.bsx);
  MSBG_CHECK(ch.storage != MSBG_STORAGE_U8, MSBG_EUNSUPP,
             "coarse-level sweeps: UNORM8 channels are swept by applyChannelPdeFast only");
  if (coarse) {
    MSBG_CHECK(level + 1 < g->nlevels && ch.lv[level + 1].defined, MSBG_ESTATE,
               "OPT_BC_COARSE_LEVEL needs level %d of channel %d", level + 1, chan);
    MSBG_CHECK(ch.storage != MSBG_STORAGE_F64, MSBG_EUNSUPP,
               "OPT_BC_COARSE_LEVEL: FP64 channels sweep with OPT_BC_NEUMANN only");
  }
  const int nMaxIter = numIter + (numIter & 1);  // msbg4.cpp:150-152
  const size_t pageBytes = (size_t)cl.bsx * cl.bsx * cl.bsx * storage_bytes(ch.storage);

  // the set-up depends on the topology and the block list only: cached
  uint64_t key = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { key = (key ^ v) * 1099511628211ull; };
  mix(g->topo_version);
  mix(blockList ? (uint64_t)nList : ~0ull);
  if (blockList)
    for (int64_t i = 0; i < nList; i++) mix((uint32_t)blockList[i]);
  const uint64_t slot = ((uint64_t)(uint32_t)chan << 8) | ((uint64_t)level << 4) |
                        (coarse ? ((haloOpt & MSBG_OPT_INTERPOLATE) ? 2u : 1u) : 0u);
  // the cached set-up is reused only if its topology version and block
  // list equal this call's, not just their hash
  auto same_inputs = [&](const LevelWork &w) -> bool {
    if (w.key != key || w.topo != g->topo_version || w.has_list != (blockList != nullptr))
      return false;
    if (!blockList) return true;
    if ((int64_t)w.list.size() != nList) return false;
    for (int64_t i = 0; i < nList; i++)
      if (w.list[(size_t)i] != blockList[i]) return false;
    return true;
  };
  {
    // a stale set-up is destroyed in place (its device buffers freed)
    auto itw = g->level_work.find(slot);
    if (itw != g->level_work.end() && !same_inputs(itw->second)) g->level_work.erase(itw);
  }
  LevelWork &W = g->level_work[slot];
  ChanLevel &wl = W.wl;
  if (!same_inputs(W)) {
    W.topo = g->topo_version;
    W.has_list = blockList != nullptr;
    W.list.clear();
    if (blockList) W.list.assign(blockList, blockList + nList);
    // swept blocks: the list's (or every) value block at this level, by colour
    std::vector<int32_t> per[8];
    std::vector<int32_t> swept;
    auto add = [&](int64_t bid) -> int {
      MSBG_CHECK(bid >= 0 && bid < g->nblocks, MSBG_EINVAL, "block id %lld out of range",
                 (long long)bid);
      if (g->h_level[bid] != level) return 0;
      int32_t pg = cl.h_page[bid];
      MSBG_CHECK(pg >= 0, MSBG_EINVAL, "block %lld is not a value block at level %d",
                 (long long)bid, level);
      int bx = (int)(bid % g->nbx), by = (int)((bid / g->nbx) % g->nby),
          bz = (int)(bid / ((int64_t)g->nbx * g->nby));
      per[(bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2)].push_back(pg);
      swept.push_back((int32_t)bid);
      return 0;
    };
    if (blockList) {
      for (int64_t i = 0; i < nList; i++) MSBG_TRY(add(blockList[i]));
    } else {
      for (int64_t b = 0; b < g->nblocks; b++)
        if (g->h_level[b] == level && cl.h_page[b] >= 0) MSBG_TRY(add(b));
    }
    std::vector<int2> items;
    W.off[0] = 0;
    for (int c = 0; c < 8; c++) {
      for (int32_t pg : per[c]) items.push_back(make_int2(0, pg));
      W.off[c + 1] = (int64_t)items.size();
    }
    // working map: NULL blocks next to a swept block become ghost pages
    wl.defined = true;
    wl.bsx = cl.bsx;
    wl.empty_raw = cl.empty_raw;
    wl.full_raw = cl.full_raw;
    wl.h_page = cl.h_page;
    std::vector<int32_t> ghostBids;
    if (coarse) {
      for (int32_t bid : swept) {
        const int bx = bid % g->nbx, by = (bid / g->nbx) % g->nby,
                  bz = (int)(bid / ((int64_t)g->nbx * g->nby));
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              const int x = bx + dx, y = by + dy, z = bz + dz;
              if (x < 0 || y < 0 || z < 0 || x >= g->nbx || y >= g->nby || z >= g->nbz)
                continue;
              const int64_t nb = ((int64_t)z * g->nby + y) * g->nbx + x;
              if (wl.h_page[nb] != MSBG_BLOCK_NULL) continue;
              wl.h_page[nb] = (int32_t)(cl.npages + (int64_t)ghostBids.size());
              ghostBids.push_back((int32_t)nb);
            }
      }
    }
    W.nghost = (int64_t)ghostBids.size();
    // per ghost page, the directions of its swept neighbours: only the
    // layers facing them are ever read
    std::vector<uint32_t> gmask((size_t)W.nghost, 0u);
    if (W.nghost) {
      std::vector<uint8_t> isSwept((size_t)g->nblocks, 0);
      for (int32_t bid : swept) isSwept[bid] = 1;
      for (int64_t i = 0; i < W.nghost; i++) {
        const int32_t bid = ghostBids[i];
        const int bx = bid % g->nbx, by = (bid / g->nbx) % g->nby,
                  bz = (int)(bid / ((int64_t)g->nbx * g->nby));
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              const int x = bx + dx, y = by + dy, z = bz + dz;
              if (x < 0 || y < 0 || z < 0 || x >= g->nbx || y >= g->nby || z >= g->nbz)
                continue;
              if (isSwept[((int64_t)z * g->nby + y) * g->nbx + x])
                gmask[i] |= 1u << ((dz + 1) * 9 + (dy + 1) * 3 + dx + 1);
            }
      }
    }
    wl.npages = cl.npages + W.nghost;
    if (W.off[8]) {
      if (W.nghost) {
        MSBG_TRY(W.d_ghost.alloc((size_t)W.nghost));
        MSBG_TRY(W.d_gmask.alloc((size_t)W.nghost));
        MSBG_HIP(hipMemcpyAsync(W.d_ghost.p, ghostBids.data(), sizeof(int32_t) * W.nghost,
                                hipMemcpyHostToDevice, g->stream));
        MSBG_HIP(hipMemcpyAsync(W.d_gmask.p, gmask.data(), sizeof(uint32_t) * W.nghost,
                                hipMemcpyHostToDevice, g->stream));
      }
      wl.h_page_bid.assign((size_t)wl.npages, -1);
      for (int64_t b = 0; b < g->nblocks; b++)
        if (wl.h_page[b] >= 0) wl.h_page_bid[wl.h_page[b]] = (int32_t)b;
      MSBG_TRY(wl.d_page.alloc((size_t)g->nblocks));
      MSBG_TRY(wl.d_page_bid.alloc((size_t)wl.npages));
      MSBG_HIP(hipMemcpyAsync(wl.d_page.p, wl.h_page.data(), sizeof(int32_t) * g->nblocks,
                              hipMemcpyHostToDevice, g->stream));
      MSBG_HIP(hipMemcpyAsync(wl.d_page_bid.p, wl.h_page_bid.data(),
                              sizeof(int32_t) * wl.npages, hipMemcpyHostToDevice, g->stream));
      MSBG_TRY(wl.d_nbr.alloc((size_t)wl.npages * 27));
      MSBG_TRY(launch_build_nbr(g, wl, level));
      MSBG_TRY(W.d_items.alloc((size_t)W.off[8]));
      MSBG_HIP(hipMemcpyAsync(W.d_items.p, items.data(), sizeof(int2) * W.off[8],
                              hipMemcpyHostToDevice, g->stream));
      MSBG_HIP(hipStreamSynchronize(g->stream));  // the host vectors go out of scope
    }
    W.key = key;
  }
  const int64_t *off = W.off;
  const int64_t ntot = off[8];
  if (ntot == 0 || nMaxIter == 0) {
    if (residual)
      for (int i = 0; i < nMaxIter; i++) residual[i] = 0.0;
    return 0;
  }

  // room for the ghost pages behind the level's pages, kept with the pool
  // (checked per call: another operation may have replaced the pool)
  const size_t need = (size_t)wl.npages * pageBytes;
  if (cl.d_pages.n < need) {
    DevBuf<uint8_t> np;
    MSBG_TRY(np.alloc(need));
    if (cl.npages)
      MSBG_HIP(hipMemcpyAsync(np.p, cl.d_pages.p, (size_t)cl.npages * pageBytes,
                              hipMemcpyDeviceToDevice, g->stream));
    MSBG_HIP(hipStreamSynchronize(g->stream));
    std::swap(cl.d_pages.p, np.p);
    std::swap(cl.d_pages.n, np.n);
  }
  // per call: the ghosts from the current level l+1 behind the level's
  // pages, then the pool is lent to the working level for the sweeps (given
  // back on every return path) and the side-cars of both are refreshed
  // wl's side-cars of the level's own pages are still the ones the last call
  // left if nothing has written the level since. Level 0 has sweeps that
  // keep its own side-cars current without renewing the stamp, so it is
  // always rebuilt whole. With the own side-cars current, the INTERPOLATE
  // fill writes the ghost pages' side-car entries as it fills them (the
  // sweep reads no others), else the ghost pages' are rebuilt.
  const bool ownSide = level >= 1 && W.side_stamp == cl.wstamp;
  bool ghostSide = false;
  if (W.nghost)
    MSBG_TRY(launch_fill_ghost_pages(g, ch, level, haloOpt, W.d_ghost.p, W.d_gmask.p,
                                     W.nghost, cl.d_pages.p, cl.npages,
                                     ownSide ? wl.d_side.p : nullptr, wl.npages, &ghostSide));
  struct Lend {
    ChanLevel &a, &b;
    Lend(ChanLevel &x, ChanLevel &y) : a(x), b(y) { swap(); }
    ~Lend() { swap(); }
    void swap() {
      std::swap(a.d_pages.p, b.d_pages.p);
      std::swap(a.d_pages.n, b.d_pages.n);
    }
  } lend(cl, wl);
  const int64_t sideFrom = !ownSide ? 0 : (ghostSide ? wl.npages : cl.npages);
  wl.side_valid = false;
  MSBG_TRY(sweep_prepare_level(g, wl, ch.storage, sideFrom));

  SweepGrid desc{};
  desc.pages = wl.d_pages.p;
  desc.side = wl.d_side.p;
  desc.consts = wl.d_consts.p;
  desc.nbr = wl.d_nbr.p;
  desc.npages = wl.npages;
  desc.dt = timeStep;
  desc.lo = doConstrZeroOne ? 0.0f : -__builtin_inff();
  desc.hi = doConstrZeroOne ? 1.0f : __builtin_inff();
  desc.outClip = coarse ? 1 : 0;
  desc.dst_pages = desc.pages;
  desc.dst_side = desc.side;
  DevBuf<double> d_partial, d_resid;
  MSBG_TRY(W.d_desc.ensure(1));
  MSBG_HIP(hipMemcpyAsync(W.d_desc.p, &desc, sizeof desc, hipMemcpyHostToDevice, g->stream));
  if (residual) {
    MSBG_TRY(d_partial.alloc((size_t)ntot));
    MSBG_TRY(d_resid.alloc((size_t)nMaxIter));
  }
  const bool exact = g->arith == MSBG_ARITH_EXACT;
  int64_t npasses = 0;
  // the level's pages change in place from here on: its own side-cars
  // (which level-0 sweeps read) are stale, also if a launch fails
  pages_written(cl);
  MSBG_HIP(hipEventRecord(g->ev[0], g->stream));
  for (int iter = 0; iter < nMaxIter; iter++) {
    for (int pass = 0; pass < 8; pass++) {
      const int64_t n = off[pass + 1] - off[pass];
      if (!n) continue;
      MSBG_TRY(launch_sweep_items(g->stream, cl.bsx, ch.storage, exact, doConstrZeroOne != 0, -1,
                                  W.d_desc.p,
                                  W.d_items.p + off[pass], n,
                                  residual ? d_partial.p + off[pass] : nullptr, iter));
      npasses++;
    }
    if (residual) MSBG_TRY(launch_reduce_residual(g, d_partial.p, ntot, d_resid.p + iter));
  }
  MSBG_HIP(hipEventRecord(g->ev[1], g->stream));
  // the sweeps kept wl's side-cars of the pages they wrote current
  W.side_stamp = cl.wstamp;
  if (residual) {
    MSBG_HIP(hipMemcpyAsync(residual, d_resid.p, sizeof(double) * nMaxIter,
                            hipMemcpyDeviceToHost, g->stream));
  }
  MSBG_HIP(hipStreamSynchronize(g->stream));
  float ms = 0.0f;
  MSBG_HIP(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
  g->last_total_ms = ms;  // the sweep passes only, as msbg_last_sweep_timing reports
  g->last_npasses = npasses;
  g->last_avg_pass_ms = npasses ? ms / (double)npasses : 0.0;
  if (residual) {
    const double nvox = (double)ntot * cl.bsx * cl.bsx * cl.bsx;
    for (int i = 0; i < nMaxIter; i++) residual[i] = std::sqrt(residual[i] / nvox);
  }
  return 0;
}

}  // extern "C"
