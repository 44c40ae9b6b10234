// Scene points generated on the device, straight into a grid's particle
// buffer: the objects of msbg_scene_objects (same draw), one thread per
// surface point with counter-based random numbers, the domain filter of
// msbg_points_upload (msbg_demo.cpp:356-360) and an order-preserving
// compaction. Saves the host generation and the PCIe upload of large scenes.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "internal.h"
#include "scene.h"

using namespace msbg;
using namespace msbg::scene;

namespace {

struct P3 {
  float x, y, z;
};

__device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// two standard normal deviates from one key (Box-Muller, uniforms in (0,1))
__device__ inline void normal2(uint64_t key, double &a, double &b) {
  const double u1 = ((double)(mix64(key) >> 11) + 0.5) * 0x1.0p-53;
  const double u2 = ((double)(mix64(key ^ 0x5851F42D4C957F2Dull) >> 11) + 0.5) * 0x1.0p-53;
  const double r = sqrt(-2.0 * log(u1)), t = 6.283185307179586 * u2;
  a = r * cos(t);
  b = r * sin(t);
}

__global__ void k_scene_points(const Obj *__restrict__ objs, int64_t nPts,
                               const int *__restrict__ perm, double sigma, double amp,
                               double freq, float3 dmin, float3 dmax, int3 size, int64_t total,
                               P3 *__restrict__ out, uint8_t *__restrict__ keep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t o = i / nPts, k = i % nPts;
  const Obj &ob = objs[o];
  const uint64_t base = ob.seed + (uint64_t)k * 0x9E3779B97F4A7C15ull;
  double n[3], j[3], spare;
  uint64_t c = 0;
  do {  // a direction of non-zero length
    normal2(mix64(base + 4 * c), n[0], n[1]);
    normal2(mix64(base + 4 * c + 1), n[2], spare);
    c++;
  } while (n[0] * n[0] + n[1] * n[1] + n[2] * n[2] < 1e-24);
  normal2(mix64(base + 4 * c + 2), j[0], j[1]);
  normal2(mix64(base + 4 * c + 3), j[2], spare);
  for (double &v : j) v *= sigma;
  float p[3];
  object_point(ob, perm, n, j, amp, freq, p);
  const float lo[3] = {dmin.x, dmin.y, dmin.z}, hi[3] = {dmax.x, dmax.y, dmax.z};
  const int s[3] = {size.x, size.y, size.z};
  bool ok = true;
  for (int a = 0; a < 3; a++) {
    if (p[a] != p[a] || p[a] < lo[a] || p[a] > hi[a]) ok = false;
    const int ip = (int)p[a];
    if (ip < 0 || ip > s[a] - 1) ok = false;
  }
  out[i] = P3{p[0], p[1], p[2]};
  keep[i] = ok ? 1 : 0;
}

// msbg_demo.cpp:356-360 for uploaded points: isInDomainRange, then
// trunc(pos) inRange, per coordinate
__global__ void k_points_keep(const float *__restrict__ xyz, int64_t n, float3 dmin, float3 dmax,
                              int3 size, uint8_t *__restrict__ keep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float lo[3] = {dmin.x, dmin.y, dmin.z}, hi[3] = {dmax.x, dmax.y, dmax.z};
  const int sz[3] = {size.x, size.y, size.z};
  bool ok = true;
  for (int a = 0; a < 3 && ok; a++) {
    const float p = xyz[3 * i + a];
    if (p != p || p < lo[a] || p > hi[a]) {
      ok = false;
    } else {
      const int ip = (int)p;  // truncate_to_int, p within the domain range
      ok = ip >= 0 && ip <= sz[a] - 1;
    }
  }
  keep[i] = ok ? 1 : 0;
}

}  // namespace

namespace msbg {

// The uploaded points d_raw[n][3] filtered into g->d_xyz, order kept.
int filter_points(msbg_grid *g, const float *d_raw, int64_t n, int64_t *nkept) {
  hipStream_t s = g->stream;
  float dmin[3], dmax[3];
  domain_bounds(g->sx, g->sy, g->sz, dmin, dmax);
  MSBG_CHECK(n < ((int64_t)1 << 31), MSBG_EINVAL, "points_upload: %lld points exceed 2^31",
             (long long)n);
  MSBG_TRY(g->d_xyz.ensure((size_t)(3 * std::max<int64_t>(n, 1))));
  int64_t m = 0;
  if (n) {
    DevBuf<uint8_t> keep;
    DevBuf<int> d_num;
    MSBG_TRY(keep.alloc((size_t)n));
    MSBG_TRY(d_num.alloc(1));
    hipLaunchKernelGGL(k_points_keep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_raw, n,
                       make_float3(dmin[0], dmin[1], dmin[2]),
                       make_float3(dmax[0], dmax[1], dmax[2]), make_int3(g->sx, g->sy, g->sz),
                       keep.p);
    MSBG_HIP(hipGetLastError());
    const P3 *src = (const P3 *)d_raw;
    P3 *dst = (P3 *)g->d_xyz.p;
    size_t bytes = 0;
    MSBG_HIP(hipcub::DeviceSelect::Flagged(nullptr, bytes, src, keep.p, dst, d_num.p, (int)n, s));
    DevBuf<uint8_t> work;
    MSBG_TRY(work.alloc(std::max<size_t>(bytes, 1)));
    MSBG_HIP(hipcub::DeviceSelect::Flagged(work.p, bytes, src, keep.p, dst, d_num.p, (int)n, s));
    int kept = 0;
    MSBG_HIP(hipMemcpyAsync(&kept, d_num.p, sizeof(int), hipMemcpyDeviceToHost, s));
    MSBG_HIP(hipStreamSynchronize(s));
    m = kept;
  }
  g->npoints = m;
  g->bin_valid = false;
  if (nkept) *nkept = m;
  return 0;
}

}  // namespace msbg

extern "C" int msbg_points_generate(msbg_grid *g, uint64_t seed, int nObj, float cmin, float cmax,
                                    float rmin, float rmax, int64_t nPts, float sigma,
                                    float ellipsoidFrac, float perlinAmp, float perlinFreq,
                                    int append, int64_t *nkept) {
  MSBG_CHECK(g && nObj >= 0 && nPts >= 0 && cmax >= cmin && rmax >= rmin && sigma >= 0 &&
                 perlinAmp >= 0 && perlinAmp < 1 && perlinFreq >= 0,
             MSBG_EINVAL, "msbg_points_generate: bad args");
  const int64_t total = (int64_t)nObj * nPts;
  const int64_t old = append ? g->npoints : 0;
  MSBG_CHECK(total < (int64_t)1 << 31 && old + total < (int64_t)1 << 40, MSBG_EINVAL,
             "msbg_points_generate: %lld points per call exceed 2^31", (long long)total);
  const std::vector<Obj> objs = draw_objects(seed, nObj, cmin, cmax, rmin, rmax, ellipsoidFrac);
  const Perlin noise = scene_noise(seed);
  float dmin[3], dmax[3];
  domain_bounds(g->sx, g->sy, g->sz, dmin, dmax);
  DevBuf<Obj> d_objs;
  DevBuf<int> d_perm;
  DevBuf<P3> tmp;
  DevBuf<uint8_t> keep;
  DevBuf<int> d_num;
  MSBG_TRY(d_objs.alloc(objs.size()));
  MSBG_TRY(d_perm.alloc(512));
  MSBG_TRY(tmp.alloc((size_t)total));
  MSBG_TRY(keep.alloc((size_t)total));
  MSBG_TRY(d_num.alloc(1));
  hipStream_t s = g->stream;
  if (!objs.empty())
    MSBG_HIP(hipMemcpyAsync(d_objs.p, objs.data(), sizeof(Obj) * objs.size(),
                            hipMemcpyHostToDevice, s));
  MSBG_HIP(hipMemcpyAsync(d_perm.p, noise.p, sizeof(int) * 512, hipMemcpyHostToDevice, s));
  // the kept points of earlier calls stay in front
  DevBuf<float> xyz;
  MSBG_TRY(xyz.alloc((size_t)(3 * std::max<int64_t>(old + total, 1))));
  if (old)
    MSBG_HIP(hipMemcpyAsync(xyz.p, g->d_xyz.p, sizeof(float) * 3 * old, hipMemcpyDeviceToDevice,
                            s));
  int kept = 0;
  if (total) {
    hipLaunchKernelGGL(k_scene_points, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       d_objs.p, nPts, d_perm.p, (double)sigma, (double)perlinAmp,
                       (double)perlinFreq, make_float3(dmin[0], dmin[1], dmin[2]),
                       make_float3(dmax[0], dmax[1], dmax[2]), make_int3(g->sx, g->sy, g->sz),
                       total, tmp.p, keep.p);
    MSBG_HIP(hipGetLastError());
    size_t bytes = 0;
    P3 *dst = (P3 *)(xyz.p + 3 * old);
    MSBG_HIP(hipcub::DeviceSelect::Flagged(nullptr, bytes, tmp.p, keep.p, dst, d_num.p,
                                           (int)total, s));
    DevBuf<uint8_t> work;
    MSBG_TRY(work.alloc(std::max<size_t>(bytes, 1)));
    MSBG_HIP(hipcub::DeviceSelect::Flagged(work.p, bytes, tmp.p, keep.p, dst, d_num.p,
                                           (int)total, s));
    MSBG_HIP(hipMemcpyAsync(&kept, d_num.p, sizeof(int), hipMemcpyDeviceToHost, s));
  }
  MSBG_HIP(hipStreamSynchronize(s));
  std::swap(g->d_xyz.p, xyz.p);
  std::swap(g->d_xyz.n, xyz.n);
  g->npoints = old + kept;
  g->bin_valid = false;
  if (nkept) *nkept = g->npoints;
  return 0;
}
