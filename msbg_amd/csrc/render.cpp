// Entry points of the field consumers: batched interpolation of a channel
// level and the demo's ray-march renderer (kernels in k_render.hip).

#include <hip/hip_runtime.h>

#include <cstring>

#include "internal.h"
#include "storage.h"

using namespace msbg;

namespace {

int level_vol(msbg_grid *g, int chan, int level, RenderVol *V, int *storage,
              int64_t *npages = nullptr) {
  MSBG_CHECK(g && level >= 0 && level < g->nlevels, MSBG_EINVAL, "bad grid/level");
  auto it = g->channels.find(chan);
  MSBG_CHECK(it != g->channels.end() && it->second.lv[level].defined, MSBG_ESTATE,
             "channel %d level %d not defined", chan, level);
  Channel &ch = it->second;
  ChanLevel &cl = ch.lv[level];
  RenderVol v{};
  v.page = cl.d_page.p;
  v.pages = cl.d_pages.p;
  v.sx = g->sx >> level;
  v.sy = g->sy >> level;
  v.sz = g->sz >> level;
  v.lb = 0;
  while ((1 << v.lb) < cl.bsx) v.lb++;
  v.nbx = g->nbx;
  v.nby = g->nby;
  // const blocks hold the channel's empty/full value in every voxel
  // (setConstBlockValue, sbg.h:2160-2168); as interpolant inputs
  if (ch.storage == MSBG_STORAGE_U16) {
    v.empty = (float)(uint16_t)cl.empty_raw;
    v.full = (float)(uint16_t)cl.full_raw;
  } else {
    v.empty = (float)cl.empty_raw;
    v.full = (float)cl.full_raw;
  }
  domain_bounds(v.sx, v.sy, v.sz, v.dmin, v.dmax);
  *V = v;
  *storage = ch.storage;
  if (npages) *npages = cl.npages;
  return 0;
}


int render_impl(msbg_grid *g, int chan, int width, int height, const float camPos[3],
                const float lookAt[3], float zoom, const float sunPos[3], const float color[3],
                uint32_t *rgb, double *ms, int64_t *counts) {
  MSBG_CHECK(g && camPos && lookAt && sunPos && color && rgb, MSBG_EINVAL,
             "msbg_render: NULL argument");
  MSBG_CHECK(width > 0 && height > 0 && (int64_t)width * height <= (1LL << 31), MSBG_EINVAL,
             "msbg_render: image %d x %d", width, height);
  RenderVol V;
  int storage;
  MSBG_TRY(level_vol(g, chan, 0, &V, &storage));
  RenderCam C{};
  for (int k = 0; k < 3; k++) {
    C.pos[k] = camPos[k];
    C.look[k] = lookAt[k];
    C.sun[k] = sunPos[k];
    C.color[k] = color[k];
  }
  C.focal = zoom;
  const size_t npix = (size_t)width * height;
  DevBuf<uint32_t> d_rgb;
  DevBuf<unsigned long long> d_cnt;
  MSBG_TRY(d_rgb.alloc(npix));
  if (counts) {
    MSBG_TRY(d_cnt.alloc(4));
    MSBG_HIP(hipMemsetAsync(d_cnt.p, 0, 4 * sizeof(unsigned long long), g->stream));
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = MSBG_EDEVICE;
  if (!rc && hipEventRecord(e0, g->stream) != hipSuccess) rc = MSBG_EDEVICE;
  if (!rc) rc = launch_render(g->stream, V, storage, C, width, height, d_rgb.p, d_cnt.p);
  if (!rc && hipEventRecord(e1, g->stream) != hipSuccess) rc = MSBG_EDEVICE;
  if (!rc && hipMemcpyAsync(rgb, d_rgb.p, sizeof(uint32_t) * npix, hipMemcpyDeviceToHost,
                            g->stream) != hipSuccess)
    rc = MSBG_EDEVICE;
  unsigned long long hc[4] = {0, 0, 0, 0};
  if (!rc && counts &&
      hipMemcpyAsync(hc, d_cnt.p, sizeof hc, hipMemcpyDeviceToHost, g->stream) != hipSuccess)
    rc = MSBG_EDEVICE;
  if (!rc && hipStreamSynchronize(g->stream) != hipSuccess) rc = MSBG_EDEVICE;
  float t = 0.0f;
  if (!rc && ms) {
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = MSBG_EDEVICE;
    *ms = t;
  }
  if (!rc && counts)
    for (int k = 0; k < 4; k++) counts[k] = (int64_t)hc[k];
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc == MSBG_EDEVICE) set_error("msbg_render: device error");
  return rc;
}

}  // namespace

extern "C" {

int msbg_interpolate(msbg_grid *g, int chan, int level, int kind, const float *xyz, int64_t n,
                     float *val, float *grad) {
  MSBG_CHECK(g && n >= 0 && (n == 0 || (xyz && val)), MSBG_EINVAL,
             "msbg_interpolate: bad arguments");
  MSBG_CHECK(kind == MSBG_IP_LINEAR || kind == MSBG_IP_BSPLINE3, MSBG_EINVAL,
             "msbg_interpolate: kind %d", kind);
  RenderVol V;
  int storage;
  MSBG_TRY(level_vol(g, chan, level, &V, &storage));
  if (n == 0) return 0;
  const bool wantGrad = kind == MSBG_IP_BSPLINE3 && grad;
  DevBuf<float> d_pts, d_val, d_grad;
  MSBG_TRY(d_pts.alloc((size_t)n * 3));
  MSBG_TRY(d_val.alloc((size_t)n));
  if (wantGrad) MSBG_TRY(d_grad.alloc((size_t)n * 3));
  MSBG_HIP(hipMemcpyAsync(d_pts.p, xyz, sizeof(float) * 3 * n, hipMemcpyHostToDevice,
                          g->stream));
  MSBG_TRY(launch_interp_points(g->stream, V, storage, kind, d_pts.p, n, d_val.p,
                                wantGrad ? d_grad.p : nullptr));
  MSBG_HIP(hipMemcpyAsync(val, d_val.p, sizeof(float) * n, hipMemcpyDeviceToHost, g->stream));
  if (wantGrad)
    MSBG_HIP(hipMemcpyAsync(grad, d_grad.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost,
                            g->stream));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int msbg_render(msbg_grid *g, int chan, int width, int height, const float camPos[3],
                const float lookAt[3], float zoom, const float sunPos[3], const float color[3],
                uint32_t *rgb, double *ms) {
  return render_impl(g, chan, width, height, camPos, lookAt, zoom, sunPos, color, rgb, ms,
                     nullptr);
}

int msbg_render_counts(msbg_grid *g, int chan, int width, int height, const float camPos[3],
                       const float lookAt[3], float zoom, const float sunPos[3],
                       const float color[3], uint32_t *rgb, int64_t counts[4]) {
  MSBG_CHECK(counts, MSBG_EINVAL, "msbg_render_counts: counts is NULL");
  return render_impl(g, chan, width, height, camPos, lookAt, zoom, sunPos, color, rgb, nullptr,
                     counts);
}

}  // extern "C"
