// Field consumers of a channel level: batched trilinear / cubic B-spline
// interpolation and the demo's ray-march renderer.
//
// Reference behaviour:
//   interpolateUint16ToFloatFast / interpolateFloatFast   sbg.h:4623-4800, 4920-5031
//   interpolateGen0<IP_LINEAR> border branch               sbg.h:3170-3245
//   interpolateWithDerivs0, IP_BSPLINE_CUBIC               sbg.h:3903-4135
//   MtCubicBSplineWeights1D                                mtool.h:344-363
//   RaymarchRenderer::render / sampleDensity / gradient    render.cpp:19-100, render.h:58-125
//
// The reference is built with -ffast-math, which lets its compiler contract
// multiply-adds into FMAs where it chooses. This file fixes one choice and
// writes every FMA explicitly with contraction off, so that its results are
// bit-identical to the CPU restatement: vectorclass mul_add is an FMA, and
// "a*b + c" / "a += b*c" fuse the first product; every other operation is
// rounded on its own.
//
// Design: one thread per pixel, 16x16 pixel tiles per workgroup (each wave
// marches a 16x4 patch, so neighbouring rays gather neighbouring voxels).
// The march runs in double as the reference does; density samples are
// FP32. The block map entry of the current block is cached in registers, so
// a step through an EMPTY block costs no page read and no map read until
// the ray leaves the block.

#include <hip/hip_runtime.h>

#include "internal.h"
#include "storage.h"

#pragma clang fp contract(off)

namespace msbg {
namespace {

template <int ST>
struct Vox {
  using R = typename Stor<ST>::raw;
  const RenderVol &V;
  // interpolant input of a raw element: the UNORM16 integer itself (the
  // reference interpolates raw values and scales afterwards), otherwise
  // the decoded value
  __device__ static inline float in(R r) {
    if constexpr (ST == MSBG_STORAGE_U16) return (float)r;
    else return (float)Stor<ST>::dec(r);
  }
  __device__ __forceinline__ int32_t code(int x, int y, int z) const {
    return V.page[((int64_t)(z >> V.lb) * V.nby + (y >> V.lb)) * V.nbx + (x >> V.lb)];
  }
  __device__ __forceinline__ float at_code(int32_t c, int x, int y, int z) const {
    if (c >= 0) {
      const int m = (1 << V.lb) - 1;
      const int64_t vid = (((int64_t)(z & m) << V.lb | (y & m)) << V.lb) | (x & m);
      return in(((const R *)V.pages)[((int64_t)c << (3 * V.lb)) + vid]);
    }
    return c == MSBG_BLOCK_FULL ? V.full : V.empty;  // NULL reads as EMPTY
  }
  __device__ __forceinline__ float at(int x, int y, int z) const { return at_code(code(x, y, z), x, y, z); }
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// interpolateGen0<IP_LINEAR> border branch: clamped indices, Interpolate1D
// p1*t + p0*(1-t) as fma(p1, t, p0*(1-t))
template <int ST>
__device__ __forceinline__ float interp_linear_border(const Vox<ST> &X, float px, float py, float pz) {
  const RenderVol &V = X.V;
  const float q[3] = {px - 0.5f, py - 0.5f, pz - 0.5f};
  int i0[3];
  float t[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    i0[k] = (int)__builtin_floorf(q[k]);
    t[k] = q[k] - (float)i0[k];
  }
  float bz[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int zz = clampi(i0[2] + k, 0, V.sz - 1);
    float by[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int yy = clampi(i0[1] + j, 0, V.sy - 1);
      const float b0 = X.at(clampi(i0[0], 0, V.sx - 1), yy, zz);
      const float b1 = X.at(clampi(i0[0] + 1, 0, V.sx - 1), yy, zz);
      by[j] = __builtin_fmaf(b1, t[0], b0 * (1.0f - t[0]));
    }
    bz[k] = __builtin_fmaf(by[1], t[1], by[0] * (1.0f - t[1]));
  }
  return __builtin_fmaf(bz[1], t[2], bz[0] * (1.0f - t[2]));
}

// cell-centred trilinear interpolation (interpolateUint16ToFloatFast);
// lerps vlerp_ps = fma(b - a, t, a) in x, then y, then z. hint: the block
// map entry of block (hb >> lb) if the caller knows it, else hb[0] = -1.
template <int ST>
__device__ __forceinline__ float interp_linear(const Vox<ST> &X, float px, float py, float pz,
                                               const int *hb = nullptr, int32_t hint = 0) {
  using R = typename Stor<ST>::raw;
  const RenderVol &V = X.V;
  const float q[3] = {px - 0.5f, py - 0.5f, pz - 0.5f};
  int i0[3];
  float t[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float f = __builtin_floorf(q[k]);
    i0[k] = (int)f;
    t[k] = q[k] - f;
  }
  if (i0[0] < 0 || i0[0] >= V.sx - 1 || i0[1] < 0 || i0[1] >= V.sy - 1 || i0[2] < 0 ||
      i0[2] >= V.sz - 1)
    return interp_linear_border(X, px, py, pz);
  float c[8];
  const int m = (1 << V.lb) - 1;
  if ((i0[0] & m) != m && (i0[1] & m) != m && (i0[2] & m) != m) {
    // the 2x2x2 stencil lies in one block: one map read (none if the
    // caller's block is this one), one page pointer, fixed corner offsets
    const bool known = hb && (hb[0] >> V.lb) == (i0[0] >> V.lb) &&
                       (hb[1] >> V.lb) == (i0[1] >> V.lb) && (hb[2] >> V.lb) == (i0[2] >> V.lb);
    const int32_t cd = known ? hint : X.code(i0[0], i0[1], i0[2]);
    if (cd >= 0) {
      const R *pg = (const R *)V.pages + ((int64_t)cd << (3 * V.lb));
      const int sy = 1 << V.lb, sz = sy << V.lb;
      const int o = (((i0[2] & m) << V.lb | (i0[1] & m)) << V.lb) | (i0[0] & m);
      c[0] = Vox<ST>::in(pg[o]);
      c[1] = Vox<ST>::in(pg[o + 1]);
      c[2] = Vox<ST>::in(pg[o + sy]);
      c[3] = Vox<ST>::in(pg[o + sy + 1]);
      c[4] = Vox<ST>::in(pg[o + sz]);
      c[5] = Vox<ST>::in(pg[o + sz + 1]);
      c[6] = Vox<ST>::in(pg[o + sz + sy]);
      c[7] = Vox<ST>::in(pg[o + sz + sy + 1]);
    } else {
      const float v = cd == MSBG_BLOCK_FULL ? V.full : V.empty;
#pragma unroll
      for (int n = 0; n < 8; n++) c[n] = v;
    }
  } else {
#pragma unroll
    for (int n = 0; n < 8; n++) c[n] = X.at(i0[0] + (n & 1), i0[1] + ((n >> 1) & 1), i0[2] + (n >> 2));
  }
  const float x0 = __builtin_fmaf(c[1] - c[0], t[0], c[0]);
  const float x1 = __builtin_fmaf(c[3] - c[2], t[0], c[2]);
  const float x2 = __builtin_fmaf(c[5] - c[4], t[0], c[4]);
  const float x3 = __builtin_fmaf(c[7] - c[6], t[0], c[6]);
  const float y0 = __builtin_fmaf(x1 - x0, t[1], x0);
  const float y1 = __builtin_fmaf(x3 - x2, t[1], x2);
  return __builtin_fmaf(y1 - y0, t[2], y0);
}

// MtCubicBSplineWeights1D, products left to right
__device__ __forceinline__ void bspline3_weights(float t, float w[4], float d[4]) {
  w[0] = (1.f / 6.f) * (1.f - t) * (1.f - t) * (1.f - t);
  w[1] = (2.f / 3.f) - 0.5f * t * t * (2.f - t);
  w[2] = (2.f / 3.f) - 0.5f * (1.f - t) * (1.f - t) * (1.f + t);
  w[3] = (1.f / 6.f) * t * t * t;
  d[0] = -(1.f - t) * (1.f - t) / 2.f;
  d[1] = 0.5f * t * t - 1.0f * (2.f - t) * t;
  d[2] = 1.f * (1.f - t) * (t + 1.f) - 0.5f * (1.f - t) * (1.f - t);
  d[3] = t * t / 2.f;
}

// interpolateWithDerivs0, IP_BSPLINE_CUBIC, options 0: 4^3 window from
// floor(p - 0.5) - 1 clamped to the domain; value and gradient accumulate
// as FMAs in z, y, x order
template <int ST>
__device__ __forceinline__ void interp_bspline3(const Vox<ST> &X, float px, float py, float pz, float *val,
                                float g[3]) {
  const RenderVol &V = X.V;
  const float p[3] = {px, py, pz};
  float wv[3][4], wd[3][4];
  int i0[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float q = p[k] - 0.5f, f = __builtin_floorf(q);
    i0[k] = (int)f - 1;
    bspline3_weights(q - f, wv[k], wd[k]);
  }
  float v = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
  for (int dz = 0; dz < 4; dz++) {
    const int zz = clampi(i0[2] + dz, 0, V.sz - 1);
#pragma unroll
    for (int dy = 0; dy < 4; dy++) {
      const int yy = clampi(i0[1] + dy, 0, V.sy - 1);
      const float wgy = wv[2][dz] * wv[1][dy];
      const float ay = wv[2][dz] * wd[1][dy], az = wd[2][dz] * wv[1][dy];
#pragma unroll
      for (int dx = 0; dx < 4; dx++) {
        const float f = X.at(clampi(i0[0] + dx, 0, V.sx - 1), yy, zz);
        v = __builtin_fmaf(wgy * wv[0][dx], f, v);
        gx = __builtin_fmaf(wgy * wd[0][dx], f, gx);
        gy = __builtin_fmaf(ay * wv[0][dx], f, gy);
        gz = __builtin_fmaf(az * wv[0][dx], f, gz);
      }
    }
  }
  *val = v;
  g[0] = gx;
  g[1] = gy;
  g[2] = gz;
}

template <int ST>
__global__ __launch_bounds__(256) void k_interp_points(RenderVol V, int kind,
                                                        const float *__restrict__ pts, int64_t n,
                                                        float *__restrict__ val,
                                                        float *__restrict__ grad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Vox<ST> X{V};
  const float px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
  if (kind == MSBG_IP_LINEAR) {
    val[i] = interp_linear(X, px, py, pz);
  } else {
    float g[3];
    interp_bspline3(X, px, py, pz, &val[i], g);
    if (grad) {
      grad[3 * i] = g[0];
      grad[3 * i + 1] = g[1];
      grad[3 * i + 2] = g[2];
    }
  }
}

// ---- ray march, in double (RDR::RealType) ----
__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
  return __builtin_fma(a[2], b[2], __builtin_fma(a[0], b[0], a[1] * b[1]));
}
__device__ __forceinline__ void normalize3(const double v[3], double o[3]) {
  const double len = __builtin_sqrt(dot3(v, v));
  if (len > (double)1e-20f) {
    o[0] = v[0] / len;
    o[1] = v[1] / len;
    o[2] = v[2] / len;
  } else {
    o[0] = 0.0;
    o[1] = 1.0;
    o[2] = 0.0;
  }
}
__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = __builtin_fma(a[1], b[2], -(a[2] * b[1]));
  o[1] = __builtin_fma(a[2], b[0], -(a[0] * b[2]));
  o[2] = __builtin_fma(a[0], b[1], -(a[1] * b[0]));
}

// Outcome of one density sample besides its value.
enum SampleKind { S_FIELD = 0, S_EMPTY = 1, S_OUT = 2 };

template <int ST>
struct Marcher {
  const Vox<ST> X;
  const float res;      // max(sx, sy, sz), as float
  const float scale;    // 1/65535 for UNORM16 (RENDER_DENS_TO_FLOAT_UI16)
  int32_t lastBid = -1; // block map entry of the last sampled block
  int32_t lastCode = MSBG_BLOCK_EMPTY;
  unsigned nSamples = 0, nInterp = 0;  // counting launch only
  // voxel position and cell of the last sample
  float vp[3];
  int vi[3];

  // sampleDensity (render.h:82-110); *kind tells why a zero came back
  __device__ __forceinline__ double sample(const double p[3], int *kind) {
    const RenderVol &V = X.V;
    nSamples++;
    vp[0] = res * (float)p[0];
    vp[1] = res * (float)p[1];
    vp[2] = res * (float)p[2];
    vi[0] = (int)__builtin_floorf(vp[0]);
    vi[1] = (int)__builtin_floorf(vp[1]);
    vi[2] = (int)__builtin_floorf(vp[2]);
    if (vi[0] < 0 || vi[0] > V.sx - 1 || vi[1] < 0 || vi[1] > V.sy - 1 || vi[2] < 0 ||
        vi[2] > V.sz - 1) {
      *kind = S_OUT;
      return 0.0;
    }
    const int32_t bid = ((vi[2] >> V.lb) * V.nby + (vi[1] >> V.lb)) * V.nbx + (vi[0] >> V.lb);
    if (bid != lastBid) {
      lastBid = bid;
      lastCode = V.page[bid];
    }
    if (lastCode == MSBG_BLOCK_EMPTY || lastCode == MSBG_BLOCK_NULL) {  // isEmptyBlock
      *kind = S_EMPTY;
      return 0.0;
    }
    *kind = S_FIELD;
    nInterp++;
    return (double)(interp_linear(X, vp[0], vp[1], vp[2], vi, lastCode) * scale);
  }

  // After a sample in an EMPTY block at vp: the number of following steps
  // whose samples certainly stay in that block (each would return 0). dv is
  // the voxel advance per step. The sampled coordinate res*(float)pos
  // differs from the exact path by far less than the 0.01-voxel margin, and
  // one more step is held back.
  __device__ __forceinline__ int steps_inside(const float dv[3]) const {
    const int bs = 1 << X.V.lb;
    const float m = 0.01f;
    float t = 1e30f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float lo = (float)((vi[k] >> X.V.lb) * bs), hi = lo + (float)bs;
      if (dv[k] > 0.0f) t = __builtin_fminf(t, (hi - m - vp[k]) / dv[k]);
      else if (dv[k] < 0.0f) t = __builtin_fminf(t, (vp[k] - lo - m) / -dv[k]);
    }
    return t > 2.0f ? (int)t - 1 : 0;
  }

  // After a sample outside the domain: true if on some axis the ray is out
  // and moving away, so it never re-enters (positions change monotonically
  // along each axis) and every later sample returns 0.
  __device__ __forceinline__ bool gone(const double dir[3]) const {
    const RenderVol &V = X.V;
    const int s[3] = {V.sx, V.sy, V.sz};
#pragma unroll
    for (int k = 0; k < 3; k++)
      if ((vi[k] < 0 && dir[k] <= 0.0) || (vi[k] > s[k] - 1 && dir[k] >= 0.0)) return true;
    return false;
  }
};

// COUNT: also accumulate {density samples taken, samples that interpolated,
// hits} into
// counts[0..2] (one atomic per wave and counter); off in the product launch
// Compiled for 6 waves per SIMD. Measured on MI355X (4K frame of a 512^3
// UNORM16 scene): 6 (80 VGPRs, the once-per-ray shading spills a little) ran
// 10% faster than the unbounded 125 VGPRs at 4 waves; 5 and 8 were in between.
constexpr int kRenderWavesPerSimd = 6;
template <int ST, bool COUNT>
__global__ __launch_bounds__(256, kRenderWavesPerSimd) void k_render(RenderVol V, RenderCam C, int W, int H,
                                                uint32_t *__restrict__ rgb,
                                                unsigned long long *__restrict__ counts) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (!COUNT && (x >= W || y >= H)) return;
  const bool live = x < W && y < H;

  // setCamera (render.h:66-77)
  double fwd[3], right[3], up[3];
  {
    const double worldUp[3] = {0.0, 1.0, 0.0};
    const double d[3] = {C.look[0] - C.pos[0], C.look[1] - C.pos[1], C.look[2] - C.pos[2]};
    double r[3];
    normalize3(d, fwd);
    cross3(fwd, worldUp, r);
    normalize3(r, right);
    cross3(right, fwd, up);
  }
  const int resi = V.sx > V.sy ? (V.sx > V.sz ? V.sx : V.sz) : (V.sy > V.sz ? V.sy : V.sz);
  Marcher<ST> M{Vox<ST>{V}, (float)resi,
                ST == MSBG_STORAGE_U16 ? 1.0f / 65535.0f : 1.0f};
  const double h = 1.0 / (double)resi;
  const double ISO = 0.5;
  const int maxSteps = (int)(2.0 / h);
  const double invW = 1.0 / (double)W, invH = 1.0 / (double)H, aspect = (double)W / (double)H;

  // primary ray (render.cpp:40-45): pixel centre in float, the rest in double
  const double ndcX = __builtin_fma((double)((float)x + 0.5f), invW, -0.5) * 2.0 * aspect;
  const double ndcY = __builtin_fma(-(double)((float)y + 0.5f), invH, 0.5) * 2.0;
  double dir0[3], dir[3];
#pragma unroll
  for (int k = 0; k < 3; k++)
    dir0[k] = __builtin_fma(up[k], ndcY, __builtin_fma(fwd[k], C.focal, right[k] * ndcX));
  normalize3(dir0, dir);

  // march (render.cpp:47-72). Two shortcuts leave the result unchanged:
  // steps that stay inside an EMPTY block only advance the position (their
  // samples are 0, so no hit can happen and prevD stays 0), and a ray that
  // has left the domain for good stops (no later sample can reach the iso
  // level). Every position is still advanced one step at a time with the
  // same FMA, so the positions that are sampled are the reference's.
  const float stepVox = M.res * (float)h;
  const float dv[3] = {(float)dir[0] * stepVox, (float)dir[1] * stepVox, (float)dir[2] * stepVox};
  double pos[3] = {C.pos[0], C.pos[1], C.pos[2]};
  double prevPos[3] = {pos[0], pos[1], pos[2]};
  int kind = S_OUT;
  double prevD = live ? M.sample(pos, &kind) : 0.0;
  double hitPos[3] = {0.0, 0.0, 0.0};
  bool hit = false;
  unsigned nAdv = 0;  // positions advanced (counting launch only)
  for (int step = 0; step < (live ? maxSteps : 0); ++step) {
#pragma unroll
    for (int k = 0; k < 3; k++) pos[k] = __builtin_fma(dir[k], h, pos[k]);
    nAdv++;
    const double d = M.sample(pos, &kind);
    if (prevD < ISO && d >= ISO) {
      const double t = (ISO - prevD) / (d - prevD);
#pragma unroll
      for (int k = 0; k < 3; k++) hitPos[k] = __builtin_fma(pos[k] - prevPos[k], t, prevPos[k]);
      hit = true;
      break;
    }
    prevD = d;
    if (kind == S_EMPTY) {
      int n = M.steps_inside(dv);
      n = n < maxSteps - 1 - step ? n : maxSteps - 1 - step;
      for (int j = 0; j < n; j++) {
#pragma unroll
        for (int k = 0; k < 3; k++) pos[k] = __builtin_fma(dir[k], h, pos[k]);
      }
      step += n;
      nAdv += n;
    } else if (kind == S_OUT && M.gone(dir)) {
      break;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) prevPos[k] = pos[k];
  }

  // shade (render.cpp:76-94)
  uint32_t out = 0x02000000u;  // BMP_MKRGB(0,0,0)
  if (hit) {
    // sampleDensityGradient (render.h:112-125)
    float gf[3] = {0.0f, 0.0f, 0.0f};
    const float p3[3] = {M.res * (float)hitPos[0], M.res * (float)hitPos[1],
                         M.res * (float)hitPos[2]};
    if (!(p3[0] < V.dmin[0] || p3[0] > V.dmax[0] || p3[1] < V.dmin[1] || p3[1] > V.dmax[1] ||
          p3[2] < V.dmin[2] || p3[2] > V.dmax[2])) {
      float val;
      interp_bspline3(M.X, p3[0], p3[1], p3[2], &val, gf);
#pragma unroll
      for (int k = 0; k < 3; k++) gf[k] *= M.scale;
    }
    double g[3], n[3], l0[3], l[3];
#pragma unroll
    for (int k = 0; k < 3; k++) g[k] = -1.0 * (double)gf[k];
    normalize3(g, n);
#pragma unroll
    for (int k = 0; k < 3; k++) l0[k] = C.sun[k] - hitPos[k];
    normalize3(l0, l);
    double diff = dot3(n, l);
    diff = diff < 0.0 ? 0.0 : diff;  // std::max(dot, 0)
    uint32_t b8[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double c = C.color[k] * diff;
      c = c < 0.0 ? 0.0 : (1.0 < c ? 1.0 : c);  // std::clamp
      b8[k] = (uint32_t)(uint8_t)__builtin_fma(c, 255.0, 0.5);
    }
    out |= (b8[2] << 16) | (b8[1] << 8) | b8[0];
  }
  if (live) rgb[(int64_t)y * W + x] = out;
  if constexpr (COUNT) {
    // whole waves reach this point (no early return when counting)
    unsigned long long c[4] = {nAdv, M.nSamples, M.nInterp, hit ? 1ull : 0ull};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      unsigned long long v = c[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if ((threadIdx.x & 63) == 0) atomicAdd(&counts[k], v);
    }
  }
}

}  // namespace

int launch_interp_points(hipStream_t s, const RenderVol &V, int storage, int kind,
                         const float *d_pts, int64_t n, float *d_val, float *d_grad) {
  if (n == 0) return 0;
  const unsigned nb = (unsigned)((n + 255) / 256);
  switch (storage) {
    case MSBG_STORAGE_U16:
      hipLaunchKernelGGL(k_interp_points<MSBG_STORAGE_U16>, dim3(nb), dim3(256), 0, s, V, kind,
                         d_pts, n, d_val, d_grad);
      break;
    case MSBG_STORAGE_F32:
      hipLaunchKernelGGL(k_interp_points<MSBG_STORAGE_F32>, dim3(nb), dim3(256), 0, s, V, kind,
                         d_pts, n, d_val, d_grad);
      break;
    case MSBG_STORAGE_F16:
      hipLaunchKernelGGL(k_interp_points<MSBG_STORAGE_F16>, dim3(nb), dim3(256), 0, s, V, kind,
                         d_pts, n, d_val, d_grad);
      break;
    case MSBG_STORAGE_BF16:
      hipLaunchKernelGGL(k_interp_points<MSBG_STORAGE_BF16>, dim3(nb), dim3(256), 0, s, V, kind,
                         d_pts, n, d_val, d_grad);
      break;
    case MSBG_STORAGE_F64:
      hipLaunchKernelGGL(k_interp_points<MSBG_STORAGE_F64>, dim3(nb), dim3(256), 0, s, V, kind,
                         d_pts, n, d_val, d_grad);
      break;
    default:
      set_error("interpolate: storage %d", storage);
      return MSBG_EINVAL;
  }
  MSBG_HIP(hipGetLastError());
  return 0;
}

template <int ST>
static void render_st(hipStream_t s, dim3 grid, const RenderVol &V, const RenderCam &C, int W,
                      int H, uint32_t *d_rgb, unsigned long long *d_counts) {
  if (d_counts)
    hipLaunchKernelGGL((k_render<ST, true>), grid, dim3(256), 0, s, V, C, W, H, d_rgb, d_counts);
  else
    hipLaunchKernelGGL((k_render<ST, false>), grid, dim3(256), 0, s, V, C, W, H, d_rgb, d_counts);
}

int launch_render(hipStream_t s, const RenderVol &V, int storage, const RenderCam &C, int W,
                  int H, uint32_t *d_rgb, unsigned long long *d_counts) {
  const dim3 grid((unsigned)((W + 15) / 16), (unsigned)((H + 15) / 16));
  switch (storage) {
    case MSBG_STORAGE_U16: render_st<MSBG_STORAGE_U16>(s, grid, V, C, W, H, d_rgb, d_counts); break;
    case MSBG_STORAGE_F32: render_st<MSBG_STORAGE_F32>(s, grid, V, C, W, H, d_rgb, d_counts); break;
    case MSBG_STORAGE_F16: render_st<MSBG_STORAGE_F16>(s, grid, V, C, W, H, d_rgb, d_counts); break;
    case MSBG_STORAGE_BF16:
      render_st<MSBG_STORAGE_BF16>(s, grid, V, C, W, H, d_rgb, d_counts);
      break;
    case MSBG_STORAGE_F64: render_st<MSBG_STORAGE_F64>(s, grid, V, C, W, H, d_rgb, d_counts); break;
    default:
      set_error("render: storage %d", storage);
      return MSBG_EINVAL;
  }
  MSBG_HIP(hipGetLastError());
  return 0;
}

}  // namespace msbg
