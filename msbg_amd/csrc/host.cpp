// Host side of libmsbg_gpu: grid object, block tables, channels and the
// C ABI entry points that orchestrate the device kernels.
//
// Reference paths are relative to the reference's src/ directory.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "internal.h"
#include "storage.h"

namespace msbg {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &body) {
  int nt = (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 64) nt = 64;
  if (n < 4096 || nt == 1) {
    body(0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    int64_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back(body, a, b);
  }
  for (auto &t : th) t.join();
}

int storage_bytes(int storage) {
  switch (storage) {
    case MSBG_STORAGE_U16:
    case MSBG_STORAGE_F16:
    case MSBG_STORAGE_BF16: return 2;
    case MSBG_STORAGE_F32: return 4;
    case MSBG_STORAGE_F64: return 8;
    case MSBG_STORAGE_U8: return 1;
  }
  return 0;
}

static int ilog2i(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  return l;
}

// sbg.cpp:262-316: domain float bounds, beps doubling from 1e-4.
void domain_bounds(int sx, int sy, int sz, float dmin[3], float dmax[3]) {
  float s[3] = {(float)sx, (float)sy, (float)sz};
  float beps = 1e-4f;
  for (int iTry = 0; iTry < 10; iTry++) {
    bool ok = true;
    for (int k = 0; k < 3; k++) {
      dmin[k] = 0.0f + beps;
      dmax[k] = s[k] - beps;
      if (!(dmin[k] > 0.0f && dmax[k] < s[k])) ok = false;
    }
    if (ok) return;
    beps *= 2.0f;
  }
}

static Channel *find_channel(msbg_grid *g, int chan) {
  auto it = g->channels.find(chan);
  return it == g->channels.end() ? nullptr : &it->second;
}

static int level_bsx(const msbg_grid *g, int level) { return g->bsx0 >> level; }

// Make sure the halo index table of (chan, level) matches its block map.
int ensure_nbr(msbg_grid *g, ChanLevel &cl, int level) {
  if (cl.nbr_valid) return 0;
  if (cl.npages > 0) {
    MSBG_TRY(cl.d_nbr.ensure((size_t)cl.npages * 27));
    MSBG_TRY(launch_build_nbr(g, cl, level));
  }
  cl.nbr_valid = true;
  return 0;
}

// Upload the block map and page->block table after a host-side change.
static int upload_pagemap(msbg_grid *g, ChanLevel &cl) {
  MSBG_TRY(cl.d_page.ensure((size_t)g->nblocks));
  MSBG_HIP(hipMemcpyAsync(cl.d_page.p, cl.h_page.data(), sizeof(int32_t) * g->nblocks,
                          hipMemcpyHostToDevice, g->stream));
  cl.h_page_bid.assign((size_t)cl.npages, -1);
  for (int64_t b = 0; b < g->nblocks; b++)
    if (cl.h_page[b] >= 0) cl.h_page_bid[cl.h_page[b]] = (int32_t)b;
  if (cl.npages > 0) {
    MSBG_TRY(cl.d_page_bid.ensure((size_t)cl.npages));
    MSBG_HIP(hipMemcpyAsync(cl.d_page_bid.p, cl.h_page_bid.data(), sizeof(int32_t) * cl.npages,
                            hipMemcpyHostToDevice, g->stream));
  }
  MSBG_HIP(hipStreamSynchronize(g->stream));
  cl.nbr_valid = false;
  pages_written(cl);
  g->topo_version++;
  return 0;
}

static int upload_flags(msbg_grid *g) {
  MSBG_TRY(g->d_flags.ensure((size_t)g->nblocks));
  MSBG_HIP(hipMemcpyAsync(g->d_flags.p, g->h_flags.data(), sizeof(uint16_t) * g->nblocks,
                          hipMemcpyHostToDevice, g->stream));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

}  // namespace msbg

using namespace msbg;

extern "C" {

const char *msbg_last_error(void) { return g_err.c_str(); }

int msbg_create(const char *name, int sx, int sy, int sz, int blockSize0,
                unsigned options, int nLevels, msbg_grid **out) {
  MSBG_CHECK(out, MSBG_EINVAL, "msbg_create: out is NULL");
  *out = nullptr;
  if (nLevels == -1) nLevels = 3;  // msbg.cpp:137
  MSBG_CHECK(nLevels >= 1 && nLevels <= kMaxLevels, MSBG_EINVAL,
             "msbg_create: nLevels %d outside [1,%d]", nLevels, kMaxLevels);
  MSBG_CHECK(blockSize0 == 8 || blockSize0 == 16 || blockSize0 == 32 || blockSize0 == 64,
             MSBG_EINVAL, "msbg_create: blockSize0 %d not in {8,16,32,64}", blockSize0);
  MSBG_CHECK(sx > 0 && sy > 0 && sz > 0 && sx % blockSize0 == 0 && sy % blockSize0 == 0 &&
                 sz % blockSize0 == 0,
             MSBG_EINVAL, "msbg_create: domain %dx%dx%d not a multiple of block size %d", sx, sy,
             sz, blockSize0);
  MSBG_CHECK((blockSize0 >> (nLevels - 1)) >= 2, MSBG_EINVAL,
             "msbg_create: block size %d too small for %d levels", blockSize0, nLevels);
  MSBG_CHECK((options & ~MSBG_OPT_SINGLE_LEVEL) == 0, MSBG_EUNSUPP,
             "msbg_create: unsupported options 0x%x", options);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("msbg_create: no HIP device available");
    return MSBG_EDEVICE;
  }
  msbg_grid *g = new msbg_grid();
  g->name = name ? name : "";
  g->sx = sx;
  g->sy = sy;
  g->sz = sz;
  g->bsx0 = blockSize0;
  g->lg0 = ilog2i(blockSize0);
  g->nbx = sx / blockSize0;
  g->nby = sy / blockSize0;
  g->nbz = sz / blockSize0;
  g->nblocks = (int64_t)g->nbx * g->nby * g->nbz;
  g->nlevels = nLevels;
  g->options = options;
  if (g->nblocks > 2000000000LL) {  // msbg.cpp:717-718
    delete g;
    set_error("msbg_create: too many blocks");
    return MSBG_EINVAL;
  }
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&g->ev[0]) != hipSuccess || hipEventCreate(&g->ev[1]) != hipSuccess) {
    delete g;
    set_error("msbg_create: stream/event creation failed");
    return MSBG_EDEVICE;
  }
  // BlockInfo.level = nLevels-1 before any refinement map (msbg.cpp:1030-1043)
  g->h_level.assign((size_t)g->nblocks, (uint16_t)(nLevels - 1));
  g->h_flags.assign((size_t)g->nblocks, 0);
  *out = g;
  return 0;
}

int msbg_destroy(msbg_grid *g) {
  if (!g) return 0;
  (void)hipStreamSynchronize(g->stream);
  g->channels.clear();
  if (g->ev[0]) (void)hipEventDestroy(g->ev[0]);
  if (g->ev[1]) (void)hipEventDestroy(g->ev[1]);
  hipStream_t s = g->stream;
  delete g;
  if (s) (void)hipStreamDestroy(s);
  return 0;
}

int msbg_get_dims(const msbg_grid *g, int dims[8], int64_t *nblocks) {
  MSBG_CHECK(g, MSBG_EINVAL, "msbg_get_dims: grid is NULL");
  if (dims) {
    dims[0] = g->sx;
    dims[1] = g->sy;
    dims[2] = g->sz;
    dims[3] = g->bsx0;
    dims[4] = g->nbx;
    dims[5] = g->nby;
    dims[6] = g->nbz;
    dims[7] = g->nlevels;
  }
  if (nblocks) *nblocks = g->nblocks;
  return 0;
}

int msbg_set_arith(msbg_grid *g, int mode) {
  MSBG_CHECK(g, MSBG_EINVAL, "msbg_set_arith: grid is NULL");
  MSBG_CHECK(mode == MSBG_ARITH_EXACT || mode == MSBG_ARITH_FAST, MSBG_EINVAL,
             "msbg_set_arith: bad mode %d", mode);
  g->arith = mode;
  return 0;
}

int msbg_synchronize(msbg_grid *g) {
  MSBG_CHECK(g, MSBG_EINVAL, "msbg_synchronize: grid is NULL");
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// device memory for callers of the device-pointer entry points
// ---------------------------------------------------------------------------
int msbg_dev_alloc(int64_t bytes, void **out) {
  MSBG_CHECK(out && bytes >= 0, MSBG_EINVAL, "msbg_dev_alloc: bad args");
  *out = nullptr;
  if (bytes == 0) return 0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("msbg_dev_alloc: no HIP device available");
    return MSBG_EDEVICE;
  }
  if (hipMalloc(out, (size_t)bytes) != hipSuccess) {
    *out = nullptr;
    set_error("msbg_dev_alloc: %lld bytes: out of device memory", (long long)bytes);
    return MSBG_ENOMEM;
  }
  return 0;
}

int msbg_dev_free(void *p) {
  if (p) MSBG_HIP(hipFree(p));
  return 0;
}

int msbg_dev_upload(void *dst, const void *src, int64_t bytes) {
  MSBG_CHECK(bytes >= 0 && (bytes == 0 || (dst && src)), MSBG_EINVAL, "msbg_dev_upload: bad args");
  if (bytes) MSBG_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return 0;
}

int msbg_dev_download(void *dst, const void *src, int64_t bytes) {
  MSBG_CHECK(bytes >= 0 && (bytes == 0 || (dst && src)), MSBG_EINVAL,
             "msbg_dev_download: bad args");
  if (bytes) MSBG_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return 0;
}

// ---------------------------------------------------------------------------
// particles
// ---------------------------------------------------------------------------

int msbg_points_upload(msbg_grid *g, const float *xyz, int64_t n, int64_t *nkept) {
  MSBG_CHECK(g && (xyz || n == 0) && n >= 0, MSBG_EINVAL, "msbg_points_upload: bad args");
  // msbg_demo.cpp:356-360 (isInDomainRange, then trunc(pos) inRange) runs on
  // the device over the copied points
  DevBuf<float> raw;
  if (n) {
    MSBG_TRY(raw.alloc((size_t)(3 * n)));
    MSBG_HIP(hipMemcpyAsync(raw.p, xyz, sizeof(float) * 3 * n, hipMemcpyHostToDevice, g->stream));
  }
  return filter_points(g, raw.p, n, nkept);
}

int msbg_points_download(msbg_grid *g, float *xyz, int64_t cap) {
  MSBG_CHECK(g && xyz && cap >= g->npoints, MSBG_EINVAL, "msbg_points_download: bad args");
  if (g->npoints)
    MSBG_HIP(hipMemcpy(xyz, g->d_xyz.p, sizeof(float) * 3 * g->npoints, hipMemcpyDeviceToHost));
  return 0;
}

int msbg_activate_from_points(msbg_grid *g, float rScan, uint32_t *counts, int64_t *nactive) {
  MSBG_CHECK(g, MSBG_EINVAL, "msbg_activate_from_points: grid is NULL");
  MSBG_CHECK(rScan >= 0 && std::isfinite(rScan), MSBG_EINVAL, "bad rScan");
  MSBG_TRY(g->d_count.ensure((size_t)g->nblocks));
  MSBG_HIP(hipMemsetAsync(g->d_count.p, 0, sizeof(uint32_t) * g->nblocks, g->stream));
  MSBG_TRY(launch_activate(g, rScan, g->d_count.p));
  std::vector<uint32_t> tmp;
  uint32_t *dst = counts;
  if (!dst && nactive) {
    tmp.resize((size_t)g->nblocks);
    dst = tmp.data();
  }
  if (dst)
    MSBG_HIP(hipMemcpyAsync(dst, g->d_count.p, sizeof(uint32_t) * g->nblocks,
                            hipMemcpyDeviceToHost, g->stream));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  if (nactive) {
    int64_t a = 0;
    for (int64_t b = 0; b < g->nblocks; b++) a += dst[b] != 0;
    *nactive = a;
  }
  return 0;
}

int msbg_levels_from_points(msbg_grid *g, const float *rScan, int nBands, int *blockLevels) {
  MSBG_CHECK(g && rScan && nBands >= 1 && nBands <= g->nlevels - 1 + (g->nlevels == 1),
             MSBG_EINVAL, "msbg_levels_from_points: bad args");
  MSBG_TRY(g->d_count.ensure((size_t)g->nblocks));
  MSBG_TRY(g->d_tmp_levels.ensure((size_t)g->nblocks));
  // bands from the widest to the narrowest: the narrowest band wins
  for (int b = nBands - 1; b >= 0; b--) {
    MSBG_HIP(hipMemsetAsync(g->d_count.p, 0, sizeof(uint32_t) * g->nblocks, g->stream));
    MSBG_TRY(launch_activate(g, rScan[b], g->d_count.p));
    MSBG_TRY(launch_level_band(g, g->d_count.p, b, nBands, g->d_tmp_levels.p));
  }
  if (blockLevels)
    MSBG_HIP(hipMemcpyAsync(blockLevels, g->d_tmp_levels.p, sizeof(int32_t) * g->nblocks,
                            hipMemcpyDeviceToHost, g->stream));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int msbg_regularize_refinement_map(msbg_grid *g, const int *in, int *out) {
  MSBG_CHECK(g && in, MSBG_EINVAL, "msbg_regularize_refinement_map: bad args");
  MSBG_CHECK(g->nlevels >= 2 && g->nlevels <= 5, MSBG_EINVAL,
             "regularizeRefinementMap needs 2 <= nLevels <= 5 (msbg.cpp:1227)");
  if (!out) out = (int *)in;
  for (int64_t b = 0; b < g->nblocks; b++)
    MSBG_CHECK(in[b] >= 0 && in[b] < g->nlevels, MSBG_EINVAL,
               "regularize: level %d of block %lld outside [0,%d)", in[b], (long long)b,
               g->nlevels);
  if (g->nlevels == 2) return 0;  // no pass runs, out untouched (msbg.cpp:1319)
  DevBuf<int32_t> a, b;
  MSBG_TRY(a.alloc((size_t)g->nblocks));
  MSBG_TRY(b.alloc((size_t)g->nblocks));
  MSBG_HIP(hipMemcpyAsync(a.p, in, sizeof(int32_t) * g->nblocks, hipMemcpyHostToDevice,
                          g->stream));
  int32_t *src = a.p, *dst = b.p;
  for (int ll = 1; ll <= g->nlevels - 2; ll++) {
    MSBG_TRY(launch_regularize(g, src, dst, ll));
    std::swap(src, dst);
  }
  MSBG_HIP(hipMemcpyAsync(out, src, sizeof(int32_t) * g->nblocks, hipMemcpyDeviceToHost,
                          g->stream));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  // UT_ASSERT0(level2 >= ll-1) of the reference: a jump of more than one
  // level survived the passes.
  return 0;
}

int msbg_set_refinement_map(msbg_grid *g, const int *blockLevels, const uint32_t *blockFlags) {
  MSBG_CHECK(g, MSBG_EINVAL, "msbg_set_refinement_map: grid is NULL");
  MSBG_TRY(g->d_tmp_levels.ensure((size_t)g->nblocks));
  if (blockLevels) {
    for (int64_t b = 0; b < g->nblocks; b++)
      MSBG_CHECK(blockLevels[b] >= 0 && blockLevels[b] < g->nlevels, MSBG_EINVAL,
                 "setRefinementMap: level %d of block %lld outside [0,%d) (msbg.cpp:1603)",
                 blockLevels[b], (long long)b, g->nlevels);
    MSBG_HIP(hipMemcpyAsync(g->d_tmp_levels.p, blockLevels, sizeof(int32_t) * g->nblocks,
                            hipMemcpyHostToDevice, g->stream));
  }
  DevBuf<uint32_t> dfl;
  if (blockFlags) {
    MSBG_TRY(dfl.alloc((size_t)g->nblocks));
    MSBG_HIP(hipMemcpyAsync(dfl.p, blockFlags, sizeof(uint32_t) * g->nblocks,
                            hipMemcpyHostToDevice, g->stream));
  }
  MSBG_TRY(g->d_level.ensure((size_t)g->nblocks));
  MSBG_TRY(g->d_flags.ensure((size_t)g->nblocks));
  MSBG_TRY(launch_flags(g, g->d_tmp_levels.p, blockFlags ? dfl.p : nullptr));
  MSBG_HIP(hipMemcpyAsync(g->h_level.data(), g->d_level.p, sizeof(uint16_t) * g->nblocks,
                          hipMemcpyDeviceToHost, g->stream));
  MSBG_HIP(hipMemcpyAsync(g->h_flags.data(), g->d_flags.p, sizeof(uint16_t) * g->nblocks,
                          hipMemcpyDeviceToHost, g->stream));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  g->have_levels = true;
  g->topo_version++;
  for (int l = 0; l < kMaxLevels; l++) g->fc_list[l].clear();
  if (!(g->options & MSBG_OPT_SINGLE_LEVEL)) {
    for (int64_t b = 0; b < g->nblocks; b++)
      if (g->h_flags[b] & MSBG_BLK_FINE_COARSE) g->fc_list[g->h_level[b]].push_back((int32_t)b);
  }
  return 0;
}

int msbg_get_block_info(msbg_grid *g, uint16_t *levels, uint16_t *flags) {
  MSBG_CHECK(g, MSBG_EINVAL, "msbg_get_block_info: grid is NULL");
  if (levels) memcpy(levels, g->h_level.data(), sizeof(uint16_t) * g->nblocks);
  if (flags) memcpy(flags, g->h_flags.data(), sizeof(uint16_t) * g->nblocks);
  return 0;
}

int msbg_set_exists(msbg_grid *g, const uint8_t *exists) {
  MSBG_CHECK(g, MSBG_EINVAL, "msbg_set_exists: grid is NULL");
  for (int64_t b = 0; b < g->nblocks; b++) {
    bool e = exists ? exists[b] != 0 : g->h_level[b] == 0;
    g->h_flags[b] = (uint16_t)(e ? (g->h_flags[b] | MSBG_BLK_EXISTS)
                                 : (g->h_flags[b] & ~MSBG_BLK_EXISTS));
  }
  return upload_flags(g);
}

int msbg_get_fine_coarse_blocks(msbg_grid *g, int level, int32_t *out, int64_t cap, int64_t *n) {
  MSBG_CHECK(g && level >= 0 && level < g->nlevels, MSBG_EINVAL,
             "msbg_get_fine_coarse_blocks: bad args");
  const auto &v = g->fc_list[level];
  if (n) *n = (int64_t)v.size();
  if (out) memcpy(out, v.data(), sizeof(int32_t) * std::min<int64_t>(cap, (int64_t)v.size()));
  return 0;
}

// ---------------------------------------------------------------------------
// channels
// ---------------------------------------------------------------------------

static int define_level(msbg_grid *g, Channel &ch, int level, const uint8_t *kind,
                        double empty_raw, double full_raw, double fill_raw) {
  ChanLevel &cl = ch.lv[level];
  cl.defined = true;
  cl.bsx = level_bsx(g, level);
  // the const values of half-width float channels are what their pages can hold
  auto held = [&](double v) -> double {
    if (ch.storage == MSBG_STORAGE_F16) return f16_bits_to_float(f16_bits_from_float((float)v));
    if (ch.storage == MSBG_STORAGE_BF16) return bf16_bits_to_float(bf16_bits_from_float((float)v));
    return v;
  };
  cl.empty_raw = held(empty_raw);
  cl.full_raw = held(full_raw);
  cl.h_page.assign((size_t)g->nblocks, MSBG_BLOCK_NULL);
  int64_t np = 0;
  for (int64_t b = 0; b < g->nblocks; b++) {
    switch (kind ? kind[b] : MSBG_KIND_NULL) {
      case MSBG_KIND_NULL: break;
      case MSBG_KIND_EMPTY: cl.h_page[b] = MSBG_BLOCK_EMPTY; break;
      case MSBG_KIND_FULL: cl.h_page[b] = MSBG_BLOCK_FULL; break;
      case MSBG_KIND_VALUE: cl.h_page[b] = (int32_t)np++; break;
      default: set_error("channel_define: bad kind %d at block %lld", kind[b], (long long)b);
        return MSBG_EINVAL;
    }
  }
  cl.npages = np;
  size_t nv = (size_t)cl.bsx * cl.bsx * cl.bsx;
  size_t bytes = (size_t)np * nv * storage_bytes(ch.storage);
  cl.d_pages.free();
  if (bytes) MSBG_TRY(cl.d_pages.alloc(bytes));
  MSBG_TRY(upload_pagemap(g, cl));
  if (np) MSBG_TRY(launch_fill_pages(g, cl, ch.storage, fill_raw, 0, np));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int msbg_channel_define(msbg_grid *g, int chan, int storage, int level, const uint8_t *kind,
                        double empty_raw, double full_raw, double fill_raw) {
  MSBG_CHECK(g && chan > 0 && level >= 0 && level < g->nlevels && storage_bytes(storage) > 0,
             MSBG_EINVAL, "msbg_channel_define: bad args (chan %d level %d storage %d)", chan,
             level, storage);
  Channel &ch = g->channels[chan];
  if (ch.storage && ch.storage != storage) {
    for (int l = 0; l < kMaxLevels; l++) {
      MSBG_CHECK(!ch.lv[l].defined || l == level, MSBG_EINVAL,
                 "channel %d already has storage %d on level %d", chan, ch.storage, l);
    }
  }
  ch.id = chan;
  ch.storage = storage;
  return define_level(g, ch, level, kind, empty_raw, full_raw, fill_raw);
}

int msbg_channel_alloc_demo(msbg_grid *g, int chan, int storage, int32_t *activeBlocks,
                            int64_t cap, int64_t *nactive) {
  MSBG_CHECK(g && chan > 0 && storage_bytes(storage) > 0, MSBG_EINVAL,
             "msbg_channel_alloc_demo: bad args");
  std::vector<uint8_t> kind((size_t)g->nblocks);
  int64_t na = 0;
  for (int64_t b = 0; b < g->nblocks; b++) {
    bool v = g->h_level[b] == 0;
    kind[b] = v ? MSBG_KIND_VALUE : MSBG_KIND_EMPTY;
    if (v) {
      if (activeBlocks && na < cap) activeBlocks[na] = (int32_t)b;
      na++;
    }
  }
  // renderDensFromFloat_(0) / (1): UNORM16 0 / 65535, UNORM8 0 / 255; the
  // float formats 0 / 1
  double full = storage == MSBG_STORAGE_U16 ? 65535.0 : (storage == MSBG_STORAGE_U8 ? 255.0 : 1.0);
  Channel &ch = g->channels[chan];
  for (int l = 0; l < kMaxLevels; l++) ch.lv[l] = ChanLevel();
  ch.id = chan;
  ch.storage = storage;
  MSBG_TRY(define_level(g, ch, 0, kind.data(), 0.0, full, full));
  for (int64_t b = 0; b < g->nblocks; b++) {
    bool e = g->h_level[b] == 0;
    g->h_flags[b] = (uint16_t)(e ? (g->h_flags[b] | MSBG_BLK_EXISTS)
                                 : (g->h_flags[b] & ~MSBG_BLK_EXISTS));
  }
  MSBG_TRY(upload_flags(g));
  if (nactive) *nactive = na;
  return 0;
}

static int get_level(msbg_grid *g, int chan, int level, Channel **pch, ChanLevel **pcl) {
  MSBG_CHECK(g && level >= 0 && level < g->nlevels, MSBG_EINVAL, "bad grid/level");
  Channel *ch = find_channel(g, chan);
  MSBG_CHECK(ch && ch->lv[level].defined, MSBG_ESTATE, "channel %d level %d not defined", chan,
             level);
  *pch = ch;
  *pcl = &ch->lv[level];
  return 0;
}

int msbg_channel_storage(msbg_grid *g, int chan, int *storage) {
  MSBG_CHECK(g && storage, MSBG_EINVAL, "msbg_channel_storage: bad args");
  auto it = g->channels.find(chan);
  MSBG_CHECK(it != g->channels.end(), MSBG_ESTATE, "channel %d is not defined", chan);
  *storage = it->second.storage;
  return 0;
}

int msbg_channel_get_pagemap(msbg_grid *g, int chan, int level, int32_t *pagemap,
                             int64_t *npages) {
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chan, level, &ch, &cl));
  if (pagemap) memcpy(pagemap, cl->h_page.data(), sizeof(int32_t) * g->nblocks);
  if (npages) *npages = cl->npages;
  return 0;
}

int msbg_channel_download(msbg_grid *g, int chan, int level, void *pages, int64_t bytes) {
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chan, level, &ch, &cl));
  int64_t need = cl->npages * (int64_t)cl->bsx * cl->bsx * cl->bsx * storage_bytes(ch->storage);
  MSBG_CHECK(pages && bytes >= need, MSBG_EINVAL, "channel_download: need %lld bytes",
             (long long)need);
  MSBG_HIP(hipStreamSynchronize(g->stream));
  if (need) MSBG_HIP(hipMemcpy(pages, cl->d_pages.p, (size_t)need, hipMemcpyDeviceToHost));
  return 0;
}

int msbg_channel_upload(msbg_grid *g, int chan, int level, const void *pages, int64_t bytes) {
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chan, level, &ch, &cl));
  int64_t need = cl->npages * (int64_t)cl->bsx * cl->bsx * cl->bsx * storage_bytes(ch->storage);
  MSBG_CHECK(pages && bytes == need, MSBG_EINVAL, "channel_upload: need exactly %lld bytes",
             (long long)need);
  MSBG_HIP(hipStreamSynchronize(g->stream));
  if (need) MSBG_HIP(hipMemcpy(cl->d_pages.p, pages, (size_t)need, hipMemcpyHostToDevice));
  pages_written(*cl);
  return 0;
}

static int block_io(msbg_grid *g, int chan, int level, int bid, void *data, bool down) {
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chan, level, &ch, &cl));
  MSBG_CHECK(data && bid >= 0 && bid < g->nblocks, MSBG_EINVAL, "block_io: bad args");
  int32_t pg = cl->h_page[bid];
  MSBG_CHECK(pg >= 0, MSBG_EINVAL, "block %d is not a value block at level %d", bid, level);
  size_t sz = (size_t)cl->bsx * cl->bsx * cl->bsx * storage_bytes(ch->storage);
  uint8_t *dev = cl->d_pages.p + (size_t)pg * sz;
  MSBG_HIP(hipStreamSynchronize(g->stream));
  if (down) {
    MSBG_HIP(hipMemcpy(data, dev, sz, hipMemcpyDeviceToHost));
  } else {
    MSBG_HIP(hipMemcpy(dev, data, sz, hipMemcpyHostToDevice));
    pages_written(*cl);
  }
  return 0;
}

int msbg_download_block(msbg_grid *g, int chan, int level, int bid, void *data) {
  return block_io(g, chan, level, bid, data, true);
}

int msbg_upload_block(msbg_grid *g, int chan, int level, int bid, const void *data) {
  return block_io(g, chan, level, bid, (void *)data, false);
}

int msbg_channel_copy(msbg_grid *g, int chanSrc, int chanDst, int dstStorage, int level) {
  Channel *src;
  ChanLevel *cs;
  MSBG_TRY(get_level(g, chanSrc, level, &src, &cs));
  MSBG_CHECK(chanDst > 0 && chanDst != chanSrc && storage_bytes(dstStorage) > 0, MSBG_EINVAL,
             "msbg_channel_copy: bad destination %d / storage %d", chanDst, dstStorage);
  Channel &dst = g->channels[chanDst];
  src = find_channel(g, chanSrc);  // the map insert above may move nothing, re-find anyway
  cs = &src->lv[level];
  for (int l = 0; l < kMaxLevels; l++)
    MSBG_CHECK(!dst.lv[l].defined || dst.storage == dstStorage || l == level, MSBG_EINVAL,
               "channel %d already has storage %d", chanDst, dst.storage);
  dst.id = chanDst;
  dst.storage = dstStorage;
  // same block map (kinds) as the source; const values converted
  std::vector<uint8_t> kind((size_t)g->nblocks);
  for (int64_t b = 0; b < g->nblocks; b++) {
    int32_t c = cs->h_page[b];
    kind[b] = c >= 0 ? MSBG_KIND_VALUE
                     : (c == MSBG_BLOCK_EMPTY ? MSBG_KIND_EMPTY
                                              : (c == MSBG_BLOCK_FULL ? MSBG_KIND_FULL : MSBG_KIND_NULL));
  }
  auto conv = [&](double raw) {
    double v = src->storage == MSBG_STORAGE_U16 ? raw * (double)(1.0f / 65535.0f) : raw;
    if (dstStorage == MSBG_STORAGE_U16) return std::nearbyint(v * 65535.0);
    return v;
  };
  MSBG_TRY(define_level(g, dst, level, kind.data(), conv(cs->empty_raw), conv(cs->full_raw), 0.0));
  MSBG_TRY(launch_convert(g, *cs, src->storage, dst.lv[level], dstStorage));
  pages_written(dst.lv[level]);
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int msbg_channel_diff(msbg_grid *g, int chanA, int chanB, int level, double *maxAbs,
                      double *rms) {
  Channel *a, *b;
  ChanLevel *ca, *cb;
  MSBG_TRY(get_level(g, chanA, level, &a, &ca));
  MSBG_TRY(get_level(g, chanB, level, &b, &cb));
  MSBG_CHECK(ca->npages == cb->npages && ca->h_page == cb->h_page, MSBG_EINVAL,
             "msbg_channel_diff: channels %d and %d have different block maps", chanA, chanB);
  double mx = 0, ss = 0;
  MSBG_TRY(launch_diff(g, *ca, a->storage, *cb, b->storage, &mx, &ss));
  int64_t n = ca->npages * (int64_t)ca->bsx * ca->bsx * ca->bsx;
  if (maxAbs) *maxAbs = mx;
  if (rms) *rms = n ? std::sqrt(ss / (double)n) : 0.0;
  return 0;
}

int msbg_get_halo_table(msbg_grid *g, int chan, int level, int32_t *out, int64_t cap) {
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chan, level, &ch, &cl));
  MSBG_CHECK(out && cap >= cl->npages * 27, MSBG_EINVAL, "get_halo_table: need %lld ints",
             (long long)(cl->npages * 27));
  MSBG_TRY(ensure_nbr(g, *cl, level));
  if (cl->npages)
    MSBG_HIP(hipMemcpyAsync(out, cl->d_nbr.p, sizeof(int32_t) * 27 * cl->npages,
                            hipMemcpyDeviceToHost, g->stream));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// hot path
// ---------------------------------------------------------------------------

int msbg_splat_points(msbg_grid *g, int chan, float rParticle, float nbDist) {
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chan, 0, &ch, &cl));
  MSBG_CHECK(ch->storage != MSBG_STORAGE_U8, MSBG_EUNSUPP,
             "splat: UNORM8 channels are converted from the UNORM16 splat (msbg_channel_copy); "
             "the demo splats into UNORM16 (msbg_demo.cpp:451-458)");
  float rScan = rParticle + nbDist;
  MSBG_CHECK(rScan > 0 && 2 * (rScan + 0.5f) <= (float)g->bsx0, MSBG_EINVAL,
             "splat: footprint 2*(rScan+0.5)=%g exceeds block size %d "
             "(the reference's lock-free colour scheme needs it to fit)",
             2 * (rScan + 0.5f), g->bsx0);
  MSBG_TRY(launch_splat(g, *ch, rScan));
  pages_written(*cl);
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int msbg_finalize_density(msbg_grid *g, int chan, float rParticle, float nbDist) {
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chan, 0, &ch, &cl));
  MSBG_CHECK(ch->storage != MSBG_STORAGE_U8, MSBG_EUNSUPP,
             "finalize: UNORM8 channels are converted from the UNORM16 field (msbg_channel_copy)");
  MSBG_TRY(launch_finalize(g, *ch, rParticle, nbDist));
  pages_written(*cl);
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int msbg_halo_fill(msbg_grid *g, int chan, int bid, int dmax, unsigned opt, float *out) {
  MSBG_CHECK(g && out && bid >= 0 && bid < g->nblocks && dmax >= 1 && dmax <= 2, MSBG_EINVAL,
             "msbg_halo_fill: bad args");
  MSBG_CHECK((opt & ~(MSBG_OPT_BC_NEUMANN | MSBG_OPT_BC_COARSE_LEVEL | MSBG_OPT_INTERPOLATE)) == 0,
             MSBG_EUNSUPP, "msbg_halo_fill: unsupported options 0x%x", opt);
  int level = g->h_level[bid];
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chan, level, &ch, &cl));
  if (opt & MSBG_OPT_BC_COARSE_LEVEL) {
    MSBG_CHECK(level + 1 < g->nlevels && ch->lv[level + 1].defined, MSBG_ESTATE,
               "COARSE_LEVEL halo needs level %d of channel %d", level + 1, chan);
    MSBG_CHECK(!(opt & MSBG_OPT_INTERPOLATE) || dmax <= 1, MSBG_EINVAL,
               "OPT_INTERPOLATE needs dmax <= 1 (halo.cpp:301-304)");
  }
  int hsx = cl->bsx + 2 * dmax;
  DevBuf<float> d;
  MSBG_TRY(d.alloc((size_t)hsx * hsx * hsx));
  MSBG_TRY(launch_halo_fill(g, *ch, level, bid, dmax, opt, d.p));
  MSBG_HIP(hipMemcpyAsync(out, d.p, sizeof(float) * hsx * hsx * hsx, hipMemcpyDeviceToHost,
                          g->stream));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int msbg_prepare_ghost_blocks(msbg_grid *g, int chan) {
  MSBG_CHECK(g, MSBG_EINVAL, "msbg_prepare_ghost_blocks: grid is NULL");
  Channel *ch = find_channel(g, chan);
  MSBG_CHECK(ch, MSBG_ESTATE, "channel %d not defined", chan);
  MSBG_CHECK(ch->storage != MSBG_STORAGE_U16 && ch->storage != MSBG_STORAGE_F64 &&
                 ch->storage != MSBG_STORAGE_U8,
             MSBG_EUNSUPP,
             "prepareGhostBlocks supports float channels (msbg.cpp:3113-3128); storage %d",
             ch->storage);
  MSBG_CHECK(!(g->options & MSBG_OPT_SINGLE_LEVEL), MSBG_ESTATE,
             "prepareGhostBlocks needs a multi-level grid (fine-coarse lists are empty "
             "with OPT_SINGLE_LEVEL, msbg.cpp:1457)");
  for (int l = 0; l + 1 < g->nlevels; l++) {
    if (g->fc_list[l].empty()) continue;
    MSBG_CHECK(ch->lv[l].defined && ch->lv[l + 1].defined, MSBG_ESTATE,
               "prepareGhostBlocks: channel %d levels %d and %d must be defined", chan, l, l + 1);
    ChanLevel &lo = ch->lv[l + 1];
    // give every existing fine-coarse block a page at level+1; the others
    // become the const EMPTY block there (msbg.cpp:3118-3122)
    std::vector<int32_t> todo;
    int64_t extra = 0;
    for (int32_t bid : g->fc_list[l]) {
      if (!(g->h_flags[bid] & MSBG_BLK_EXISTS)) {
        lo.h_page[bid] = MSBG_BLOCK_EMPTY;
        continue;
      }
      todo.push_back(bid);
      if (lo.h_page[bid] < 0) lo.h_page[bid] = (int32_t)(lo.npages + extra++);
    }
    if (extra) {
      size_t pbytes = (size_t)lo.bsx * lo.bsx * lo.bsx * storage_bytes(ch->storage);
      DevBuf<uint8_t> np;
      MSBG_TRY(np.alloc((size_t)(lo.npages + extra) * pbytes));
      if (lo.npages)
        MSBG_HIP(hipMemcpyAsync(np.p, lo.d_pages.p, (size_t)lo.npages * pbytes,
                                hipMemcpyDeviceToDevice, g->stream));
      MSBG_HIP(hipStreamSynchronize(g->stream));
      std::swap(lo.d_pages.p, np.p);
      std::swap(lo.d_pages.n, np.n);
      lo.npages += extra;
    }
    MSBG_TRY(upload_pagemap(g, lo));
    MSBG_TRY(ensure_nbr(g, ch->lv[l], l));
    if (!todo.empty()) {
      DevBuf<int32_t> d;
      MSBG_TRY(d.alloc(todo.size()));
      MSBG_HIP(hipMemcpyAsync(d.p, todo.data(), sizeof(int32_t) * todo.size(),
                              hipMemcpyHostToDevice, g->stream));
      MSBG_TRY(launch_ghost(g, *ch, l, d.p, (int64_t)todo.size()));
      pages_written(lo);
      MSBG_HIP(hipStreamSynchronize(g->stream));
    }
  }
  return 0;
}

}  // extern "C"

namespace {

enum { MODE_COLOR8 = 0, MODE_REDBLACK = 1, MODE_JACOBI = 2 };

// The 8-colour in-place MCF sweep without constraints: the hot path, with
// cached colour lists and the fast or exact kernel (msbg4.cpp:60-957).
int mcf_colour_inplace(msbg_grid *g, int chPhi, const int32_t *blockList, int64_t nList,
                       int doConstrZeroOne, int numIter, float timeStep, double *residual) {
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chPhi, 0, &ch, &cl));
  int nMaxIter = numIter;
  if (nMaxIter % 2 == 1) nMaxIter++;  // msbg4.cpp:150-152
  // colour lists of page indices: skip level != 0 (msbg4.cpp:467-470);
  // list blocks must be value blocks (UT_ASSERT0, msbg4.cpp:444). The lists
  // are cached on the device for repeated calls with the same block list.
  uint64_t key = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { key = (key ^ v) * 1099511628211ull; };
  mix((uint64_t)chPhi);
  mix(g->topo_version);
  mix(blockList ? (uint64_t)nList : ~0ull);
  if (blockList)
    for (int64_t i = 0; i < nList; i++) mix((uint32_t)blockList[i]);
  int64_t off[9];
  if (key != g->sweep_key) {
    std::vector<int32_t> per[8];
    auto add = [&](int64_t bid) -> int {
      MSBG_CHECK(bid >= 0 && bid < g->nblocks, MSBG_EINVAL, "block id %lld out of range",
                 (long long)bid);
      int32_t pg = cl->h_page[bid];
      MSBG_CHECK(pg >= 0, MSBG_EINVAL,
                 "block %lld in blockList is not a value block (msbg4.cpp:444)", (long long)bid);
      if (g->h_level[bid] != 0) return 0;
      int bx = (int)(bid % g->nbx), by = (int)((bid / g->nbx) % g->nby),
          bz = (int)(bid / ((int64_t)g->nbx * g->nby));
      per[(bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2)].push_back(pg);
      return 0;
    };
    if (blockList) {
      for (int64_t i = 0; i < nList; i++) MSBG_TRY(add(blockList[i]));
    } else {
      for (int64_t b = 0; b < g->nblocks; b++)
        if (cl->h_page[b] >= 0) MSBG_TRY(add(b));
    }
    std::vector<int2> flat;
    off[0] = 0;
    for (int c = 0; c < 8; c++) {
      for (int32_t pg : per[c]) flat.push_back(make_int2(0, pg));
      off[c + 1] = (int64_t)flat.size();
    }
    if (off[8]) {
      MSBG_TRY(g->d_sweep_list.ensure((size_t)off[8]));
      MSBG_HIP(hipMemcpyAsync(g->d_sweep_list.p, flat.data(), sizeof(int2) * off[8],
                              hipMemcpyHostToDevice, g->stream));
      MSBG_HIP(hipStreamSynchronize(g->stream));
    }
    g->sweep_key = key;
    for (int c = 0; c < 9; c++) g->sweep_off[c] = off[c];
  } else {
    for (int c = 0; c < 9; c++) off[c] = g->sweep_off[c];
  }
  int64_t ntot = off[8];
  MSBG_TRY(ensure_nbr(g, *cl, 0));
  if (residual && nMaxIter > 0) {
    MSBG_TRY(g->d_partial.ensure((size_t)std::max<int64_t>(ntot, 1)));
    MSBG_TRY(g->d_resid.ensure((size_t)nMaxIter));
  }
  if (ntot) {
    MSBG_TRY(sweep_prepare(g, *ch));
    SweepGrid desc = sweep_desc(*ch, timeStep, doConstrZeroOne);
    MSBG_TRY(g->d_desc.ensure(1));
    MSBG_HIP(hipMemcpyAsync(g->d_desc.p, &desc, sizeof desc, hipMemcpyHostToDevice, g->stream));
    MSBG_HIP(hipStreamSynchronize(g->stream));
  }
  const bool exact = g->arith == MSBG_ARITH_EXACT;
  const int tile = g->tile[g->lg0];
  int64_t npasses = 0;
  MSBG_HIP(hipEventRecord(g->ev[0], g->stream));
  for (int it = 0; it < nMaxIter && ntot > 0; it++) {
    for (int pass = 0; pass < 8; pass++) {
      int64_t n = off[pass + 1] - off[pass];
      if (!n) continue;
      MSBG_TRY(launch_sweep_items(g->stream, g->bsx0, ch->storage, exact, doConstrZeroOne != 0,
                                  tile, g->d_desc.p,
                                  g->d_sweep_list.p + off[pass], n,
                                  residual ? g->d_partial.p + off[pass] : nullptr,
                                  (int)(g->sweep_step + (uint64_t)it)));
      npasses++;
    }
    if (residual)
      MSBG_TRY(launch_reduce_residual(g, g->d_partial.p, ntot, g->d_resid.p + it));
  }
  MSBG_HIP(hipEventRecord(g->ev[1], g->stream));
  MSBG_HIP(hipEventSynchronize(g->ev[1]));
  float ms = 0;
  MSBG_HIP(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
  g->last_total_ms = ms;
  g->last_npasses = npasses;
  g->sweep_step += (uint64_t)nMaxIter;
  g->last_avg_pass_ms = npasses ? ms / (double)npasses : 0.0;
  if (residual && nMaxIter > 0) {
    if (ntot) {
      MSBG_HIP(hipMemcpy(residual, g->d_resid.p, sizeof(double) * nMaxIter,
                         hipMemcpyDeviceToHost));
      int64_t nvox = ntot * (int64_t)cl->bsx * cl->bsx * cl->bsx;
      for (int it = 0; it < nMaxIter; it++) residual[it] = std::sqrt(residual[it] / (double)nvox);
    } else {
      for (int it = 0; it < nMaxIter; it++) residual[it] = 0.0;
    }
  }
  return 0;
}

// Every other operator / mode / constraint combination of applyChannelPdeFast
// (msbg4.cpp:60-957), exact FP32 arithmetic (FP64 for FP64 channels):
//  - colour: 8 passes by parity bits, in place;
//  - red-black: 2 passes, pass p updates blocks with (bx+by+bz)&1 != p, in
//    place (msbg4.cpp:420-425);
//  - Jacobi: through the temporary channel chTmp, which keeps chPhi's page
//    numbering; value blocks that are not swept are NULL in it, as in the
//    reference's reset temporary (msbg4.cpp:285-297); an odd numIter turns
//    the last sweep into a copy (msbg4.cpp:481-487). chTmp is removed after.
//  - hard constraints around the field at entry, snapshot in chTmp2
//    (msbg4.cpp:262-267, 813-818), removed after.
int pde_general(msbg_grid *g, int chPhi, int chTmp, int chTmp2, const int32_t *blockList,
                int64_t nList, int op, int mode, int doConstrZeroOne, bool hard, float tol1,
                float tol2, int numIter, float dt, float taubin, float alpha, double *residual) {
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chPhi, 0, &ch, &cl));
  const int storage = ch->storage;
  const int npass = mode == MODE_COLOR8 ? 8 : (mode == MODE_REDBLACK ? 2 : 1);
  std::vector<std::vector<int2>> per((size_t)npass);
  std::vector<uint8_t> swept((size_t)g->nblocks, 0);
  auto add = [&](int64_t bid) -> int {
    MSBG_CHECK(bid >= 0 && bid < g->nblocks, MSBG_EINVAL, "block id %lld out of range",
               (long long)bid);
    const int32_t pg = cl->h_page[bid];
    MSBG_CHECK(pg >= 0, MSBG_EINVAL,
               "block %lld in blockList is not a value block (msbg4.cpp:444)", (long long)bid);
    if (g->h_level[bid] != 0) return 0;  // msbg4.cpp:467-470
    const int bx = (int)(bid % g->nbx), by = (int)((bid / g->nbx) % g->nby),
              bz = (int)(bid / ((int64_t)g->nbx * g->nby));
    int pass = 0;
    if (mode == MODE_COLOR8) pass = (bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2);
    if (mode == MODE_REDBLACK) pass = ((bx + by + bz) & 1) ? 0 : 1;
    per[(size_t)pass].push_back(make_int2(0, pg));
    swept[(size_t)bid] = 1;
    return 0;
  };
  if (blockList) {
    for (int64_t i = 0; i < nList; i++) MSBG_TRY(add(blockList[i]));
  } else {
    for (int64_t b = 0; b < g->nblocks; b++)
      if (cl->h_page[b] >= 0) MSBG_TRY(add(b));
  }
  std::vector<int2> flat;
  std::vector<int64_t> off((size_t)npass + 1, 0);
  for (int p = 0; p < npass; p++) {
    off[(size_t)p] = (int64_t)flat.size();
    flat.insert(flat.end(), per[(size_t)p].begin(), per[(size_t)p].end());
  }
  off[(size_t)npass] = (int64_t)flat.size();
  const int64_t ntot = off[(size_t)npass];
  int nMaxIter = numIter;
  if (nMaxIter % 2 == 1) nMaxIter++;  // msbg4.cpp:150-152

  MSBG_TRY(ensure_nbr(g, *cl, 0));
  MSBG_TRY(sweep_prepare(g, *ch));
  Channel *tmp = nullptr, *p0 = nullptr;
  if (mode == MODE_JACOBI) {
    MSBG_CHECK(chTmp > 0 && chTmp != chPhi && (!hard || chTmp != chTmp2), MSBG_EINVAL,
               "applyChannelPdeFast: the Jacobi mode needs a temporary channel chTmp");
    std::vector<uint8_t> kind((size_t)g->nblocks);
    for (int64_t b = 0; b < g->nblocks; b++) {
      const int32_t c = cl->h_page[b];
      kind[(size_t)b] = c >= 0 ? MSBG_KIND_VALUE
                               : (c == MSBG_BLOCK_EMPTY ? MSBG_KIND_EMPTY
                                                        : (c == MSBG_BLOCK_FULL ? MSBG_KIND_FULL
                                                                                : MSBG_KIND_NULL));
    }
    g->channels.erase(chTmp);  // the reference resets the temporary (msbg4.cpp:285-289)
    Channel &t = g->channels[chTmp];
    t.id = chTmp;
    t.storage = storage;
    MSBG_TRY(define_level(g, t, 0, kind.data(), cl->empty_raw, cl->full_raw, 0.0));
    for (int64_t b = 0; b < g->nblocks; b++)
      if (cl->h_page[b] >= 0 && !swept[(size_t)b]) t.lv[0].h_page[(size_t)b] = MSBG_BLOCK_NULL;
    MSBG_TRY(upload_pagemap(g, t.lv[0]));
    MSBG_TRY(ensure_nbr(g, t.lv[0], 0));
    MSBG_TRY(sweep_prepare(g, t));
    tmp = &t;
  }
  if (hard) {
    MSBG_CHECK(chTmp2 > 0 && chTmp2 != chPhi, MSBG_EINVAL,
               "applyChannelPdeFast: hard constraints need the snapshot channel chTmp2");
    MSBG_TRY(msbg_channel_copy(g, chPhi, chTmp2, storage, 0));
    p0 = find_channel(g, chTmp2);
    MSBG_CHECK(p0 && p0->lv[0].h_page == cl->h_page, MSBG_EINVAL,
               "applyChannelPdeFast: snapshot page numbering differs");
  }
  auto desc = [&](Channel &src, Channel &dst) {
    SweepGrid d = sweep_desc(src, dt, doConstrZeroOne);
    d.dst_pages = dst.lv[0].d_pages.p;
    d.dst_side = dst.lv[0].d_side.p;
    d.phi0 = p0 ? p0->lv[0].d_pages.p : nullptr;
    d.tol1 = tol1;
    d.tol2 = tol2;
    d.taubin = op == 1 ? taubin : 0.0f;
    d.alpha = alpha;
    return d;
  };
  SweepGrid hd[2] = {tmp ? desc(*ch, *tmp) : desc(*ch, *ch), tmp ? desc(*tmp, *ch) : desc(*ch, *ch)};
  MSBG_TRY(g->d_desc.ensure(2));
  MSBG_HIP(hipMemcpyAsync(g->d_desc.p, hd, sizeof hd, hipMemcpyHostToDevice, g->stream));
  DevBuf<int2> items;
  if (ntot) {
    MSBG_TRY(items.alloc((size_t)ntot));
    MSBG_HIP(hipMemcpyAsync(items.p, flat.data(), sizeof(int2) * ntot, hipMemcpyHostToDevice,
                            g->stream));
  }
  if (residual && nMaxIter > 0) {
    MSBG_TRY(g->d_partial.ensure((size_t)std::max<int64_t>(ntot, 1)));
    MSBG_TRY(g->d_resid.ensure((size_t)nMaxIter));
  }
  constexpr int64_t kBilapChunk = 4096;
  DevBuf<uint8_t> scratch;
  if (op == 2 && ntot)
    MSBG_TRY(scratch.alloc((size_t)std::min<int64_t>(ntot, kBilapChunk) * cl->bsx * cl->bsx *
                           cl->bsx * storage_bytes(storage)));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  int64_t npasses = 0;
  MSBG_HIP(hipEventRecord(g->ev[0], g->stream));
  for (int it = 0; it < nMaxIter && ntot > 0; it++) {
    const SweepGrid *dd = g->d_desc.p + (tmp ? (it & 1) : 0);
    if (tmp && it == nMaxIter - 1 && (numIter % 2) == 1) {
      MSBG_TRY(launch_copy_items(g->stream, g->bsx0, storage, dd, items.p, ntot));
      if (residual) MSBG_HIP(hipMemsetAsync(g->d_resid.p + it, 0, sizeof(double), g->stream));
      continue;
    }
    for (int p = 0; p < npass; p++) {
      const int64_t n = off[(size_t)p + 1] - off[(size_t)p];
      if (!n) continue;
      if (op == 2) {  // through scratch pages, in chunks
        for (int64_t c0 = 0; c0 < n; c0 += kBilapChunk) {
          const int64_t nc = std::min<int64_t>(kBilapChunk, n - c0);
          MSBG_TRY(launch_bilap_items(g->stream, g->bsx0, storage, hard, dd,
                                      items.p + off[(size_t)p] + c0, nc,
                                      residual ? g->d_partial.p + off[(size_t)p] + c0 : nullptr,
                                      scratch.p));
        }
      } else {
        MSBG_TRY(launch_pde_items(g->stream, g->bsx0, storage, op, hard, dd,
                                  items.p + off[(size_t)p], n,
                                  residual ? g->d_partial.p + off[(size_t)p] : nullptr,
                                  (int)(g->sweep_step + (uint64_t)it)));
      }
      npasses++;
    }
    if (residual) MSBG_TRY(launch_reduce_residual(g, g->d_partial.p, ntot, g->d_resid.p + it));
  }
  MSBG_HIP(hipEventRecord(g->ev[1], g->stream));
  MSBG_HIP(hipEventSynchronize(g->ev[1]));
  float ms = 0;
  MSBG_HIP(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
  g->last_total_ms = ms;
  g->last_npasses = npasses;
  g->sweep_step += (uint64_t)nMaxIter;
  g->last_avg_pass_ms = npasses ? ms / (double)npasses : 0.0;
  if (residual && nMaxIter > 0) {
    if (ntot) {
      MSBG_HIP(hipMemcpy(residual, g->d_resid.p, sizeof(double) * nMaxIter,
                         hipMemcpyDeviceToHost));
      const int64_t nvox = ntot * (int64_t)cl->bsx * cl->bsx * cl->bsx;
      for (int it = 0; it < nMaxIter; it++) residual[it] = std::sqrt(residual[it] / (double)nvox);
    } else {
      for (int it = 0; it < nMaxIter; it++) residual[it] = 0.0;
    }
  }
  // the reference resets the temporaries (msbg4.cpp:949-953)
  if (tmp) g->channels.erase(chTmp);
  if (p0) g->channels.erase(chTmp2);
  return 0;
}

}  // namespace

extern "C" {

int msbg_apply_channel_pde(msbg_grid *g, int chPhi, int chTmp, int chTmp2,
                           const int32_t *blockList, int64_t nList, int laplTyp,
                           int doConstrZeroOne, float densThr, float densThrTol1,
                           float densThrTol2, int numIter, float gridSpacing, float timeStep,
                           float lambda, float taubinStrength, float anisDiffAlpha,
                           double *residual) {
  MSBG_CHECK(g, MSBG_EINVAL, "applyChannelPdeFast: grid is NULL");
  (void)lambda;  // forced to 0 by the reference (msbg4.cpp:164): no soft constraints
  // mode from the sign and the colour flag (msbg4.cpp:102-114)
  int op = laplTyp, mode = MODE_JACOBI;
  if (op < 0) {
    op = -op;
    if (op & MSBG_OPT_8_COLOR_SCHEME) {
      op &= ~MSBG_OPT_8_COLOR_SCHEME;
      mode = MODE_COLOR8;
    } else {
      mode = MODE_REDBLACK;
    }
  }
  op &= ~1024;  // the reference's halo test flag (msbg4.cpp:119-123) changes nothing here
  MSBG_CHECK(op == 1 || op == 2 || op == 3 || op == 4 || op == 14 || op == 15 || op == 17,
             MSBG_EUNSUPP,
             "applyChannelPdeFast: laplTyp %d is not implemented (1, 2, 3, 4, 14, 15, 17 are)", op);
  MSBG_CHECK(!(op == 1 || op == 3 || op == 17) || g->bsx0 == 16, MSBG_EUNSUPP,
             "applyChannelPdeFast: laplTyp %d reads neighbours at the fixed offsets of 16^3 "
             "blocks (msbg4.cpp:688-693); block size is %d", op, g->bsx0);
  MSBG_CHECK(!(op == 14 || op == 15) || anisDiffAlpha != 0.0f, MSBG_EINVAL,
             "applyChannelPdeFast: Perona-Malik needs anisDiffAlpha != 0");
  MSBG_CHECK(gridSpacing == 1.0f, MSBG_EINVAL, "gridSpacing must be 1 (msbg4.cpp:326-330)");
  MSBG_CHECK(numIter >= 0, MSBG_EUNSUPP,
             "numIter < 0 (epsComprRange) asserts in the reference (msbg4.cpp:240)");
  MSBG_CHECK(timeStep >= 0 && std::isfinite(timeStep), MSBG_EUNSUPP,
             "timeStep < 0 (2nd-order stepping) asserts in the reference (msbg4.cpp:180)");
  const bool hard = densThr > 1e-12f || densThrTol1 > 1e-12f;  // msbg4.cpp:192-193
  Channel *ch;
  ChanLevel *cl;
  MSBG_TRY(get_level(g, chPhi, 0, &ch, &cl));
  MSBG_CHECK(ch->storage != MSBG_STORAGE_F64 || op == 4, MSBG_EUNSUPP,
             "applyChannelPdeFast: FP64 channels run laplTyp 4 only");
  if (op == 4 && mode == MODE_COLOR8 && !hard)
    return mcf_colour_inplace(g, chPhi, blockList, nList, doConstrZeroOne, numIter, timeStep,
                              residual);
  return pde_general(g, chPhi, chTmp, chTmp2, blockList, nList, op, mode, doConstrZeroOne, hard,
                     densThrTol1, densThrTol2, numIter, timeStep,
                     std::fabs(taubinStrength) > 1e-20f ? taubinStrength : 0.0f, anisDiffAlpha,
                     residual);
}

int msbg_apply_channel_pde_fast(msbg_grid *g, int chPhi, int chTmp, int chTmp2,
                                const int32_t *blockList, int64_t nList, int laplTyp,
                                int doConstrZeroOne, float densThr, float densThrTol1,
                                float densThrTol2, int numIter, float gridSpacing,
                                float timeStep, double *residual) {
  return msbg_apply_channel_pde(g, chPhi, chTmp, chTmp2, blockList, nList, laplTyp,
                                doConstrZeroOne, densThr, densThrTol1, densThrTol2, numIter,
                                gridSpacing, timeStep, 0.0f, 0.0f, 0.0f, residual);
}

int msbg_last_sweep_timing(msbg_grid *g, double *total_ms, double *avg_pass_ms,
                           int64_t *npasses) {
  MSBG_CHECK(g, MSBG_EINVAL, "msbg_last_sweep_timing: grid is NULL");
  if (total_ms) *total_ms = g->last_total_ms;
  if (avg_pass_ms) *avg_pass_ms = g->last_avg_pass_ms;
  if (npasses) *npasses = g->last_npasses;
  return 0;
}

}  // extern "C"
