/*
 * msbg_oracle.c -- CPU restatement of MSBG's sparse-block stencil hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product). See
 * msbg_oracle.h for the conventions. Reference citations are relative to the
 * reference's src/ directory.
 *
 * Two builds of this one file exist:
 *  - liboracle_strict.so: -O2 -ffp-contract=off, IEEE arithmetic. This is the
 *    parity oracle: every float operation is evaluated in the reference's
 *    source order with one rounding per operation.
 *  - liboracle_fast.so: the reference's own flags (-O3 -ffast-math
 *    -fno-finite-math-only) + OpenMP, AVX2. Used only to time the CPU
 *    baseline; its bits may differ from the strict build, exactly as the
 *    reference binary's bits depend on its compiler's contraction choices.
 */
#include "msbg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

uint64_t or_version(void) { return 20261203u; }

/* ------------------------------------------------------------------------ */
/* conversions: sbg.h:225-366                                               */
/* ------------------------------------------------------------------------ */

/* renderDensToFloat_<uint16_t>: fi * (1.f/65535) (sbg.h:233-244) */
float or_u16_to_float(uint16_t u) {
  const float s = 1.0f / 65535.0f;
  return (float)u * s;
}

/* renderDensFromFloat_ -> renderDensRoundFromFloat: roundf(f*65535)
 * (sbg.h:272-315), the scalar path used by the splat and finalize. */
uint16_t or_u16_from_float_roundf(float f) {
  int fi = (int)roundf(f * 65535.0f);
  return (uint16_t)fi;
}

/* renderDensFromFloat_storeSimd8<uint16_t>: round_to_int(vf*65535), i.e.
 * cvtps2dq under the default MXCSR = round half to even (sbg.h:343-351,
 * vectorf256.h:1064-1068), used by the sweep's store. */
uint16_t or_u16_from_float_rne(float f) {
  int fi = (int)nearbyintf(f * 65535.0f);
  return (uint16_t)fi;
}

uint16_t or_f16_from_float(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 -> inf */
  if (ax < 0x38800000u) {                                   /* below 2^-14 */
    uint32_t e = ax >> 23;
    if (e < 102) return (uint16_t)sign; /* < 2^-25 rounds to zero */
    uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126u - e;          /* value / 2^-24 = m * 2^(e-126) */
    uint32_t q = m >> shift, rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t e = (ax >> 23) - 112u; /* rebias 127 -> 15 */
  uint32_t m = ax & 0x7fffffu;
  uint32_t h = (e << 10) | (m >> 13), rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)(sign | h);
}

float or_f16_to_float(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu, x;
  if (e == 0) {
    float v = (float)m * (1.0f / 16777216.0f); /* m * 2^-24, exact */
    memcpy(&x, &v, 4);
    x |= sign;
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

uint16_t or_bf16_from_float(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}

float or_bf16_to_float(uint16_t h) {
  uint32_t x = (uint32_t)h << 16;
  float f;
  memcpy(&f, &x, 4);
  return f;
}

/* renderDensToFloat_<uint8_t>: u * (1.f/255) (sbg.h:228-231) */
float or_u8_to_float(uint8_t u) {
  const float s = 1.0f / 255.0f;
  return (float)u * s;
}

/* renderDensFromFloat_storeSimd8_SR<uint8_t> (sbg.h:324-340): v = 255 f,
 * floor(v) + 1 if u < v - floor(v), else floor(v), truncated to int */
uint8_t or_u8_from_float_sr(float f, float u) {
  const float v = 255.0f * f, fl = floorf(v), fr = v - fl;
  return (uint8_t)(int)(u < fr ? fl + 1.0f : fl);
}

/* murmur3's fmix32 (the finalizer FMA_FastRandSeed mixes with,
 * fastmath.h:92-106), chained over (step, page, voxel); the float in [0,1)
 * is built like FMA_FastRand's (fastmath.h:68-73): the top 23 bits as the
 * mantissa of a number in [1,2), minus 1 */
static inline uint32_t sr_mix(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
float or_sr_uniform(uint32_t step, uint32_t page, uint32_t vid) {
  const uint32_t h = sr_mix(sr_mix(sr_mix(step ^ 0x6a09e667u) + page) + vid);
  const uint32_t b = (h >> 9) | 0x3f800000u;
  float f;
  memcpy(&f, &b, 4);
  return f - 1.0f;
}

static size_t storage_size(int st) {
  switch (st) {
    case OR_U8: return 1;
    case OR_U16: case OR_F16: case OR_BF16: return 2;
    case OR_F32: return 4;
    case OR_F64: return 8;
  }
  return 0;
}

/* raw storage value (exactly representable in double) -> float halo value.
 * transferValue<uint16_t> = renderDensToFloat_ (halo.h:484-501). */
static inline float raw_to_float(int st, double raw) {
  switch (st) {
    case OR_U16: return or_u16_to_float((uint16_t)raw);
    case OR_U8: return or_u8_to_float((uint8_t)raw);
    case OR_F64: return (float)raw;
    /* a const value of a half-width float channel is what its pages hold */
    case OR_F16: return or_f16_to_float(or_f16_from_float((float)raw));
    case OR_BF16: return or_bf16_to_float(or_bf16_from_float((float)raw));
    default: return (float)raw;
  }
}

static inline double page_raw(const or_level *L, int64_t page, int vid) {
  size_t nv = (size_t)L->bsx * L->bsx * L->bsx;
  size_t i = (size_t)page * nv + (size_t)vid;
  switch (L->storage) {
    case OR_U16: return ((const uint16_t *)L->pages)[i];
    case OR_F32: return ((const float *)L->pages)[i];
    case OR_F16: return or_f16_to_float(((const uint16_t *)L->pages)[i]);
    case OR_BF16: return or_bf16_to_float(((const uint16_t *)L->pages)[i]);
    case OR_F64: return ((const double *)L->pages)[i];
    case OR_U8: return ((const uint8_t *)L->pages)[i];
  }
  return 0;
}

/* a const value as the channel holds it: half-width float channels store
 * their empty/full values rounded to the storage precision */
static inline double const_raw(int st, double r) {
  if (st == OR_F16) return or_f16_to_float(or_f16_from_float((float)r));
  if (st == OR_BF16) return or_bf16_to_float(or_bf16_from_float((float)r));
  return r;
}

/* value of block code c (non-NULL) at voxel vid, as raw storage value */
static inline double block_raw(const or_level *L, int32_t c, int vid) {
  if (c >= 0) return page_raw(L, c, vid);
  if (c == OR_EMPTY) return const_raw(L->storage, L->empty_raw);
  return const_raw(L->storage, L->full_raw);
}

static inline int ilog2i(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  return l;
}

/* ------------------------------------------------------------------------ */
/* A1: domain bounds and point filter                                       */
/* ------------------------------------------------------------------------ */

/* sbg.cpp:262-316 (the non-PARTICLE_QUANT_POS branch): min = 0 + beps,
 * max = S - beps, beps doubling from 1e-4 until min > 0 and max < S. */
void or_domain_bounds(int sx, int sy, int sz, float dmin[3], float dmax[3]) {
  float s[3] = {(float)sx, (float)sy, (float)sz};
  float beps = 1e-4f;
  for (int iTry = 0; iTry < 10; iTry++) {
    int ok = 1;
    for (int k = 0; k < 3; k++) {
      dmin[k] = 0.0f + beps;
      dmax[k] = s[k] - beps;
      if (!(dmin[k] > 0.0f && dmax[k] < s[k])) ok = 0;
    }
    if (ok) return;
    beps *= 2.0f;
  }
}

/* msbg_demo.cpp:356-360: isInDomainRange(pos) (msbg.h:2900-2908), then
 * ipos = truncate_to_int(pos) must be inRange (sbg.h:616-620). */
int64_t or_domain_filter(int sx, int sy, int sz, const float *xyz, int64_t n,
                         uint8_t *keep) {
  float dmin[3], dmax[3];
  or_domain_bounds(sx, sy, sz, dmin, dmax);
  int s[3] = {sx, sy, sz};
  int64_t nk = 0;
  for (int64_t i = 0; i < n; i++) {
    int ok = 1;
    for (int k = 0; k < 3; k++) {
      float p = xyz[3 * i + k];
      if (p != p) { ok = 0; break; }
      if (p < dmin[k] || p > dmax[k]) { ok = 0; break; }
      int ip = (int)p; /* truncate toward zero */
      if (ip < 0 || ip > s[k] - 1) { ok = 0; break; }
    }
    keep[i] = (uint8_t)ok;
    nk += ok;
  }
  return nk;
}

/* ------------------------------------------------------------------------ */
/* A2/A3: activation and home blocks, msbg_demo.cpp:368-389                 */
/* ------------------------------------------------------------------------ */
void or_activate(int bsx, int nbx, int nby, int nbz, const float *xyz,
                 const uint8_t *keep, int64_t n, float rScan, uint32_t *count,
                 int32_t *home) {
  const float scale = 1.0f / (float)bsx; /* scale2DestBlockGrid */
  const float rScanBsx = rScan / (float)bsx;
  const int bmax[3] = {nbx - 1, nby - 1, nbz - 1};
  const int lg = ilog2i(bsx);
  for (int64_t i = 0; i < n; i++) {
    if (keep && !keep[i]) {
      if (home) home[i] = -1;
      continue;
    }
    int b1[3], b2[3];
    for (int k = 0; k < 3; k++) {
      float bp = xyz[3 * i + k] * scale;
      int lo = (int)(bp - rScanBsx), hi = (int)(bp + rScanBsx);
      b1[k] = lo > 0 ? lo : 0;
      b2[k] = hi < bmax[k] ? hi : bmax[k];
    }
    for (int bz = b1[2]; bz <= b2[2]; bz++)
      for (int by = b1[1]; by <= b2[1]; by++)
        for (int bx = b1[0]; bx <= b2[0]; bx++)
          count[((int64_t)bz * nby + by) * nbx + bx]++;
    if (home) {
      int hb[3];
      for (int k = 0; k < 3; k++) hb[k] = ((int)xyz[3 * i + k]) >> lg;
      home[i] = (int32_t)(((int64_t)hb[2] * nby + hb[1]) * nbx + hb[0]);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* A4: level map                                                            */
/* ------------------------------------------------------------------------ */

/* msbg_demo.cpp:418-428 */
void or_levels_from_activity(const uint32_t *count, int64_t nblocks,
                             int nLevels, int *levels) {
  for (int64_t i = 0; i < nblocks; i++) levels[i] = count[i] ? 0 : nLevels - 1;
}

/* msbg.cpp:1202-1423. Levels are held as int8 internally (blockmapTmp2).
 * Passes ll = 1..nLevels-2 each read the previous pass's map. With
 * nLevels == 2 there is no pass and `out` is left untouched. Returns 0, or
 * -1 where the reference would fail an always-on assertion. */
int or_regularize(int nbx, int nby, int nbz, int nLevels, const int *in,
                  int *out) {
  if (!(nLevels >= 2 && nLevels <= 5)) return -1;
  int64_t nb = (int64_t)nbx * nby * nbz;
  int8_t *tmp2 = (int8_t *)malloc((size_t)nb);
  if (!tmp2) return -1;
  for (int64_t i = 0; i < nb; i++) tmp2[i] = (int8_t)in[i];
  int rc = 0;
  for (int ll = 1; ll <= nLevels - 2; ll++) {
    int isLast = (ll == nLevels - 2);
    for (int bz = 0; bz < nbz; bz++)
      for (int by = 0; by < nby; by++)
        for (int bx = 0; bx < nbx; bx++) {
          int64_t bid = ((int64_t)bz * nby + by) * nbx + bx;
          int level = tmp2[bid];
          if (level >= ll) {
            int level2 = nLevels - 1;
            for (int dz = -1; dz <= 1; dz++)
              for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                  if (!dx && !dy && !dz) continue;
                  int x = bx + dx, y = by + dy, z = bz + dz;
                  if (x < 0 || y < 0 || z < 0 || x >= nbx || y >= nby || z >= nbz) continue;
                  int l2 = tmp2[((int64_t)z * nby + y) * nbx + x];
                  if (l2 < level2) level2 = l2;
                }
            if (level2 < ll - 1) rc = -1; /* UT_ASSERT0(level2>=ll-1) */
            if (level2 == ll - 1) level = ll;
          }
          out[bid] = level;
          if (isLast && (level < 0 || level > 4)) rc = -1;
        }
    if (!isLast)
      for (int64_t i = 0; i < nb; i++) tmp2[i] = (int8_t)out[i];
  }
  free(tmp2);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* A5: block flags, msbg.cpp:1649-1745 at levelMg 0                         */
/* ------------------------------------------------------------------------ */
void or_block_flags(int nbx, int nby, int nbz, int nLevels, const int *levels,
                    const uint32_t *inFlags, uint16_t *level_out,
                    uint16_t *flags_out) {
  for (int bz = 0; bz < nbz; bz++)
    for (int by = 0; by < nby; by++)
      for (int bx = 0; bx < nbx; bx++) {
        int64_t bid = ((int64_t)bz * nby + by) * nbx + bx;
        int level = levels[bid]; /* MAX(blockLevels[bid], levelMg=0) */
        unsigned fl = inFlags ? (uint16_t)inFlags[bid] : 0u; /* bi->flags = blockFlags[bid] */
        int border = bx == 0 || bx == nbx - 1 || by == 0 || by == nby - 1 || bz == 0 || bz == nbz - 1;
        if (border) fl |= OR_BLK_DOM_BORDER; /* FLAG_COPY */
        else fl &= ~(unsigned)OR_BLK_DOM_BORDER;
        int l2max = 0, l2min = nLevels - 1;
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              if (!dx && !dy && !dz) continue;
              int x = bx + dx, y = by + dy, z = bz + dz;
              if (x < 0 || y < 0 || z < 0 || x >= nbx || y >= nby || z >= nbz) continue;
              int l2 = levels[((int64_t)z * nby + y) * nbx + x];
              if (l2 > l2max) l2max = l2;
              if (l2 < l2min) l2min = l2;
            }
        if (l2max > level) fl |= OR_BLK_FINE_COARSE;
        if (l2min < level) fl |= OR_BLK_COARSE_FINE;
        fl &= ~(unsigned)OR_BLK_UNIFORM_EMPTY;
        if (!(fl & OR_BLK_EXISTS) && !(fl & (OR_BLK_FINE_COARSE | OR_BLK_COARSE_FINE)))
          fl |= OR_BLK_UNIFORM_EMPTY;
        level_out[bid] = (uint16_t)level;
        flags_out[bid] = (uint16_t)fl;
      }
}

/* ------------------------------------------------------------------------ */
/* A7: colours                                                              */
/* ------------------------------------------------------------------------ */
int or_color8(int bx, int by, int bz) {
  return (bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2); /* msbg.h:642-646 */
}

void or_color_lists(int nbx, int nby, const int32_t *blocks, int n,
                    int32_t *out, int32_t off[9]) {
  int cnt[8] = {0};
  for (int i = 0; i < n; i++) {
    int b = blocks[i];
    int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    cnt[or_color8(bx, by, bz)]++;
  }
  off[0] = 0;
  for (int c = 0; c < 8; c++) off[c + 1] = off[c] + cnt[c];
  int pos[8];
  for (int c = 0; c < 8; c++) pos[c] = off[c];
  for (int i = 0; i < n; i++) {
    int b = blocks[i];
    int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    out[pos[or_color8(bx, by, bz)]++] = b;
  }
}

/* ------------------------------------------------------------------------ */
/* A8: splat, msbg_demo.cpp:541-636                                         */
/* ------------------------------------------------------------------------ */
static void splat_one(const or_level *L, const float *p, float rScan,
                      float distSqMax, float distSqMaxInv) {
  const int lg = ilog2i(L->bsx), msk = L->bsx - 1;
  const int smax[3] = {L->sx - 1, L->sy - 1, L->sz - 1};
  int lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    int a = (int)ceilf(p[k] - rScan - 0.5f);
    int b = (int)floorf(p[k] + rScan - 0.5f);
    lo[k] = a > 0 ? a : 0;
    hi[k] = b < smax[k] ? b : smax[k];
  }
  const float x0 = p[0] - 0.5f, y0 = p[1] - 0.5f, z0 = p[2] - 0.5f;
  const int nv = L->bsx * L->bsx * L->bsx;
  uint16_t *pg = (uint16_t *)L->pages;
  for (int iz = lo[2]; iz <= hi[2]; iz++) {
    float dz = (float)iz - z0;
    for (int iy = lo[1]; iy <= hi[1]; iy++) {
      float dy = (float)iy - y0;
      float dzy = dz * dz + dy * dy;
      for (int ix = lo[0]; ix <= hi[0]; ix++) {
        float dx = (float)ix - x0;
        float d2 = dzy + dx * dx;
        if (d2 > distSqMax) continue;
        int64_t ib = ((int64_t)(iz >> lg) * L->nby + (iy >> lg)) * L->nbx + (ix >> lg);
        int iv = (((iz & msk) << lg) + (iy & msk)) << lg | (ix & msk);
        int32_t c = L->page[ib];
        if (c < 0) return; /* !isValueBlock -> goto loop_exit */
        uint16_t v = or_u16_from_float_roundf(d2 * distSqMaxInv);
        uint16_t *t = &pg[(int64_t)c * nv + iv];
        if (v < *t) *t = v;
      }
    }
  }
}

void or_splat_u16(const or_level *L, const float *xyz, const int32_t *home,
                  int64_t n, float rScan) {
  /* distSqMax = sqf(rParticle+nbDist) = rScan^2 (msbg_demo.cpp:524-525) */
  const float distSqMax = rScan * rScan, distSqMaxInv = 1.0f / distSqMax;
  for (int64_t i = 0; i < n; i++) {
    if (home && home[i] < 0) continue;
    splat_one(L, xyz + 3 * i, rScan, distSqMax, distSqMaxInv);
  }
}

/* ------------------------------------------------------------------------ */
/* A9: finalize, msbg_demo.cpp:655-680, MT_LINSTEPF mtool.h:400-409         */
/* ------------------------------------------------------------------------ */
/* the density f of msbg_demo.cpp:664-675 before its UNORM16 store (0 when
 * q >= 0.999); FP32 channels store it as is */
float or_finalize_density_f(uint16_t u, float rParticle, float nbDist) {
  const float rScan = rParticle + nbDist;
  const float distSqMax = rScan * rScan;
  float distSq = or_u16_to_float(u);
  if (!(distSq < 0.999f)) return 0.0f;
  distSq *= distSqMax;
  float dist = sqrtf(distSq) - rParticle;
  float a = -rParticle, b = rParticle + nbDist, x = dist, ls;
  if (x < a) ls = 0.0f;
  else if (x >= b) ls = 1.0f;
  else ls = (x - a) / (b - a);
  return 1.0f - ls;
}

uint16_t or_finalize_value(uint16_t u, float rParticle, float nbDist) {
  const float rScan = rParticle + nbDist;
  const float distSqMax = rScan * rScan;
  float distSq = or_u16_to_float(u);
  if (distSq < 0.999f) {
    distSq *= distSqMax;
    float dist = sqrtf(distSq) - rParticle;
    float a = -rParticle, b = rParticle + nbDist, x = dist, ls;
    if (x < a) ls = 0.0f;
    else if (x >= b) ls = 1.0f;
    else ls = (x - a) / (b - a);
    float f = 1.0f - ls;
    return or_u16_from_float_roundf(f);
  }
  return 0;
}

void or_finalize_u16(uint16_t *pages, int64_t nvox, float rParticle,
                     float nbDist) {
  for (int64_t i = 0; i < nvox; i++)
    pages[i] = or_finalize_value(pages[i], rParticle, nbDist);
}

/* ------------------------------------------------------------------------ */
/* A11/A13: halo fill                                                       */
/* ------------------------------------------------------------------------ */

static inline int32_t code_at(const or_level *L, int x, int y, int z) {
  int lg = ilog2i(L->bsx);
  return L->page[((int64_t)(z >> lg) * L->nby + (y >> lg)) * L->nbx + (x >> lg)];
}

static inline int vid_at(const or_level *L, int x, int y, int z) {
  int lg = ilog2i(L->bsx), m = L->bsx - 1;
  return ((((z & m) << lg) + (y & m)) << lg) + (x & m);
}

static inline int in_range(const or_level *L, int x, int y, int z) {
  return x >= 0 && x < L->sx && y >= 0 && y < L->sy && z >= 0 && z < L->sz;
}

static inline void clip3(const or_level *L, int *x, int *y, int *z) {
  if (*x < 0) *x = 0;
  if (*x > L->sx - 1) *x = L->sx - 1;
  if (*y < 0) *y = 0;
  if (*y > L->sy - 1) *y = L->sy - 1;
  if (*z < 0) *z = 0;
  if (*z > L->sz - 1) *z = L->sz - 1;
}

/* SparseGrid::getValue(int) (sbg.h:2622-2648): emptyValue for NULL */
static double get_value_raw(const or_level *L, int x, int y, int z) {
  int32_t c = code_at(L, x, y, z);
  if (c == OR_NULL) return const_raw(L->storage, L->empty_raw);
  return block_raw(L, c, vid_at(L, x, y, z));
}

/* convert a float to the raw domain the way `Data_T = float` assignment
 * does for the reference's types (uint16: truncation). */
static inline double to_raw_type(int st, float v) {
  if (st == OR_U16) return (double)(uint16_t)v;
  return (double)v;
}

/* SparseGrid<Data_T>::interpolateGhost (sbg.h:3002-3095), options 0:
 * cell-centred (po = 0.5), dmax = 1. Intermediate results are held in
 * Data_T (buf_x/buf_y/buf_z), so UNORM16 truncates at each stage. */
static double interpolate_ghost_raw(const or_level *Lo, float px, float py,
                                    float pz) {
  const float po = 0.5f;
  float x = px - po, y = py - po, z = pz - po;
  int ix = (int)floorf(x), iy = (int)floorf(y), iz = (int)floorf(z);
  float u = x - (float)ix, v = y - (float)iy, w = z - (float)iz;
  double bx[2], by[2], bzv[2];
  double val0 = Lo->empty_raw;
  int have0 = 0;
  for (int k = 0; k < 2; k++) {
    for (int j = 0; j < 2; j++) {
      for (int i = 0; i < 2; i++) {
        int X = ix + i, Y = iy + j, Z = iz + k;
        double val;
        if (!in_range(Lo, X, Y, Z)) {
          clip3(Lo, &X, &Y, &Z);
          val = get_value_raw(Lo, X, Y, Z);
        } else if (code_at(Lo, X, Y, Z) == OR_NULL) {
          if (!have0) {
            /* getValue(Vec4f p0): trunc(floor(p)), clamped (sbg.h:1170-1176) */
            int qx = (int)floorf(px), qy = (int)floorf(py), qz = (int)floorf(pz);
            clip3(Lo, &qx, &qy, &qz);
            val0 = get_value_raw(Lo, qx, qy, qz);
            have0 = 1;
          }
          val = val0;
        } else {
          val = block_raw(Lo, code_at(Lo, X, Y, Z), vid_at(Lo, X, Y, Z));
        }
        bx[i] = val;
      }
      by[j] = to_raw_type(Lo->storage, (float)bx[1] * u + (float)bx[0] * (1.0f - u));
    }
    bzv[k] = to_raw_type(Lo->storage, (float)by[1] * v + (float)by[0] * (1.0f - v));
  }
  return to_raw_type(Lo->storage, (float)bzv[1] * w + (float)bzv[0] * (1.0f - w));
}

/* interpolate_ghost_raw as an entry point (the tests' handle on it); the
 * result is in the level's raw domain. */
double or_interpolate_ghost(const or_level *Lo, float px, float py, float pz) {
  return interpolate_ghost_raw(Lo, px, py, pz);
}

/* HaloBlockSet::getOutOfBlockValue (halo.h:267-355): the value of halo voxel
 * (x,y,z) whose level-L block is NULL or which lies outside the domain. */
static float out_of_block_value(const or_level *L, const or_level *Lo,
                                int inRange, int x, int y, int z,
                                const int pos0[3], int32_t own, unsigned opt) {
  if (opt & OR_OPT_BC_NEUMANN) {
    int b = L->bsx;
    int cx = x < pos0[0] ? pos0[0] : (x > pos0[0] + b - 1 ? pos0[0] + b - 1 : x);
    int cy = y < pos0[1] ? pos0[1] : (y > pos0[1] + b - 1 ? pos0[1] + b - 1 : y);
    int cz = z < pos0[2] ? pos0[2] : (z > pos0[2] + b - 1 ? pos0[2] + b - 1 : z);
    return raw_to_float(L->storage, block_raw(L, own, vid_at(L, cx, cy, cz)));
  } else if (opt & OR_OPT_BC_COARSE_LEVEL) {
    if (!inRange) {
      clip3(L, &x, &y, &z);
      int32_t c = code_at(L, x, y, z);
      if (c != OR_NULL) return raw_to_float(L->storage, block_raw(L, c, vid_at(L, x, y, z)));
    }
    if (opt & OR_OPT_INTERPOLATE) {
      float px = ((float)x + 0.5f) / 2, py = ((float)y + 0.5f) / 2, pz = ((float)z + 0.5f) / 2;
      return raw_to_float(Lo->storage, interpolate_ghost_raw(Lo, px, py, pz));
    }
    int xl = x / 2, yl = y / 2, zl = z / 2;
    clip3(Lo, &xl, &yl, &zl);
    /* getValuePtr may return NULL here (UT_ASSERT2 only in the reference);
     * this restatement reads the coarse level's empty value then. */
    return raw_to_float(Lo->storage, get_value_raw(Lo, xl, yl, zl));
  }
  return raw_to_float(L->storage, L->empty_raw);
}

/* HaloBlockSet::fillHaloBlock_ (halo.cpp:267-479). Each halo row (hy,hz)
 * is split like the reference into the first segment (x < x0), the middle
 * segment (the block column of the block itself: one block lookup for the
 * whole row) and the last segment (x >= x0+bsx); every voxel follows the
 * same rule: in-domain non-NULL block -> its value, else
 * getOutOfBlockValue. */
void or_halo_fill(const or_level *L, const or_level *Lo, int bid, int dmax,
                  unsigned opt, float *out) {
  const int b = L->bsx, hsx = b + 2 * dmax;
  int bx = bid % L->nbx, by = (bid / L->nbx) % L->nby, bz = bid / (L->nbx * L->nby);
  const int pos0[3] = {bx * b, by * b, bz * b};
  const int32_t own = L->page[bid];
  for (int hz = 0; hz < hsx; hz++)
    for (int hy = 0; hy < hsx; hy++) {
      int y = pos0[1] - dmax + hy, z = pos0[2] - dmax + hz;
      float *row = out + (size_t)(hz * hsx + hy) * hsx;
      int inRZY = z >= 0 && z < L->sz && y >= 0 && y < L->sy;
      for (int hx = 0; hx < hsx; hx++) {
        int x = pos0[0] - dmax + hx;
        if (hx == dmax && hsx > 2 * dmax) {
          /* middle segment */
          int32_t c = inRZY ? code_at(L, x, y, z) : OR_NULL;
          if (c != OR_NULL && L->storage == OR_U16 && c >= 0) {
            const uint16_t *src = (const uint16_t *)L->pages +
                                  (size_t)c * b * b * b + (size_t)vid_at(L, x, y, z);
            const float s = 1.0f / 65535.0f;
            for (int j = 0; j < b; j++) row[hx + j] = (float)src[j] * s;
          } else if (c != OR_NULL && L->storage == OR_F32 && c >= 0) {
            const float *src = (const float *)L->pages + (size_t)c * b * b * b +
                               (size_t)vid_at(L, x, y, z);
            for (int j = 0; j < b; j++) row[hx + j] = src[j];
          } else {
            for (int j = 0; j < b; j++) {
              int xx = x + j;
              row[hx + j] = (c != OR_NULL)
                                ? raw_to_float(L->storage, block_raw(L, c, vid_at(L, xx, y, z)))
                                : out_of_block_value(L, Lo, inRZY, xx, y, z, pos0, own, opt);
            }
          }
          hx += b - 1;
          continue;
        }
        int inR = inRZY && x >= 0 && x < L->sx;
        int32_t c = inR ? code_at(L, x, y, z) : OR_NULL;
        if (inR && c != OR_NULL)
          row[hx] = raw_to_float(L->storage, block_raw(L, c, vid_at(L, x, y, z)));
        else
          row[hx] = out_of_block_value(L, Lo, inR, x, y, z, pos0, own, opt);
      }
    }
}

/* Double-precision halo for the FP64 oracle run (Neumann/EMPTY only). */
void or_halo_fill_f64(const or_level *L, int bid, int dmax, unsigned opt,
                      double *out) {
  const int b = L->bsx, hsx = b + 2 * dmax;
  int bx = bid % L->nbx, by = (bid / L->nbx) % L->nby, bz = bid / (L->nbx * L->nby);
  const int pos0[3] = {bx * b, by * b, bz * b};
  const int32_t own = L->page[bid];
  for (int hz = 0; hz < hsx; hz++)
    for (int hy = 0; hy < hsx; hy++)
      for (int hx = 0; hx < hsx; hx++) {
        int x = pos0[0] - dmax + hx, y = pos0[1] - dmax + hy, z = pos0[2] - dmax + hz;
        int inR = in_range(L, x, y, z);
        int32_t c = inR ? code_at(L, x, y, z) : OR_NULL;
        double v;
        if (inR && c != OR_NULL) {
          v = block_raw(L, c, vid_at(L, x, y, z));
        } else if (opt & OR_OPT_BC_NEUMANN) {
          int cx = x < pos0[0] ? pos0[0] : (x > pos0[0] + b - 1 ? pos0[0] + b - 1 : x);
          int cy = y < pos0[1] ? pos0[1] : (y > pos0[1] + b - 1 ? pos0[1] + b - 1 : y);
          int cz = z < pos0[2] ? pos0[2] : (z > pos0[2] + b - 1 ? pos0[2] + b - 1 : z);
          v = block_raw(L, own, vid_at(L, cx, cy, cz));
        } else {
          v = L->empty_raw;
        }
        if (L->storage == OR_U16) v = v * (1.0 / 65535.0);
        out[(hz * hsx + hy) * hsx + hx] = v;
      }
}

/* ------------------------------------------------------------------------ */
/* A12: MCF voxel update, msbg4.cpp:559-637 and 813-825                     */
/* ------------------------------------------------------------------------ */
/* Vec8f max(a,b) = a > b ? a : b and min(a,b) = a < b ? a : b
 * (vectorf256.h, _mm256_max_ps/_mm256_min_ps operand order). */
static inline float vmaxf(float a, float b) { return a > b ? a : b; }
static inline float vminf(float a, float b) { return a < b ? a : b; }

float or_mcf_voxel(const float *h, int hsx, int hid, float dt,
                   int doConstrZeroOne) {
  const int s1 = hsx, s2 = hsx * hsx;
#define F_(dx, dy, dz) h[hid + (dx) + (dy)*s1 + (dz)*s2]
  float f0 = F_(0, 0, 0), f1 = F_(-1, 0, 0), f2 = F_(1, 0, 0), f3 = F_(0, -1, 0),
        f4 = F_(0, 1, 0), f5 = F_(0, 0, -1), f6 = F_(0, 0, 1);
  float fx = 0.5f * (f2 - f1), fy = 0.5f * (f4 - f3), fz = 0.5f * (f6 - f5);
  float fxx = f1 + f2 - 2.0f * f0, fyy = f3 + f4 - 2.0f * f0, fzz = f5 + f6 - 2.0f * f0;
  float fy_right = (F_(1, 1, 0) - F_(1, -1, 0)) * 0.5f,
        fy_left = (F_(-1, 1, 0) - F_(-1, -1, 0)) * 0.5f,
        fz_right = (F_(1, 0, 1) - F_(1, 0, -1)) * 0.5f,
        fz_left = (F_(-1, 0, 1) - F_(-1, 0, -1)) * 0.5f,
        fz_up = (F_(0, 1, 1) - F_(0, 1, -1)) * 0.5f,
        fz_down = (F_(0, -1, 1) - F_(0, -1, -1)) * 0.5f;
#undef F_
  float fyx = (fy_right - fy_left) * 0.5f, fzx = (fz_right - fz_left) * 0.5f,
        fzy = (fz_up - fz_down) * 0.5f;
  float H = (fy * fy + fz * fz) * fxx + (fx * fx + fz * fz) * fyy + (fx * fx + fy * fy) * fzz -
            2.0f * (fx * fy * fyx + fx * fz * fzx + fy * fz * fzy);
  float g = fx * fx + fy * fy + fz * fz;
  H = g > 1e-7f ? H / g : 0.0f;
  float D = dt * H;
  D = vminf(vmaxf(D, -0.1f), 0.1f);
  float phi = f0 + D;
  if (doConstrZeroOne) phi = vminf(vmaxf(phi, 0.0f), 1.0f);
  return phi;
}

static inline double vmaxd(double a, double b) { return a > b ? a : b; }
static inline double vmind(double a, double b) { return a < b ? a : b; }

double or_mcf_voxel_f64(const double *h, int hsx, int hid, double dt,
                        int doConstrZeroOne) {
  const int s1 = hsx, s2 = hsx * hsx;
#define F_(dx, dy, dz) h[hid + (dx) + (dy)*s1 + (dz)*s2]
  double f0 = F_(0, 0, 0), f1 = F_(-1, 0, 0), f2 = F_(1, 0, 0), f3 = F_(0, -1, 0),
         f4 = F_(0, 1, 0), f5 = F_(0, 0, -1), f6 = F_(0, 0, 1);
  double fx = 0.5 * (f2 - f1), fy = 0.5 * (f4 - f3), fz = 0.5 * (f6 - f5);
  double fxx = f1 + f2 - 2.0 * f0, fyy = f3 + f4 - 2.0 * f0, fzz = f5 + f6 - 2.0 * f0;
  double fyx = ((F_(1, 1, 0) - F_(1, -1, 0)) * 0.5 - (F_(-1, 1, 0) - F_(-1, -1, 0)) * 0.5) * 0.5;
  double fzx = ((F_(1, 0, 1) - F_(1, 0, -1)) * 0.5 - (F_(-1, 0, 1) - F_(-1, 0, -1)) * 0.5) * 0.5;
  double fzy = ((F_(0, 1, 1) - F_(0, 1, -1)) * 0.5 - (F_(0, -1, 1) - F_(0, -1, -1)) * 0.5) * 0.5;
#undef F_
  double H = (fy * fy + fz * fz) * fxx + (fx * fx + fz * fz) * fyy + (fx * fx + fy * fy) * fzz -
             2.0 * (fx * fy * fyx + fx * fz * fzx + fy * fz * fzy);
  double g = fx * fx + fy * fy + fz * fz;
  /* the reference's threshold literal is the float 1e-7f (msbg4.cpp:604) */
  H = g > (double)1e-7f ? H / g : 0.0;
  double D = dt * H;
  /* limiter with the reference's float literals -0.1f / 0.1f (msbg4.cpp:625) */
  D = vmind(vmaxd(D, (double)-0.1f), (double)0.1f);
  double phi = f0 + D;
  if (doConstrZeroOne) phi = vmind(vmaxd(phi, 0.0), 1.0);
  return phi;
}

/* store of the new value: UNORM16 via round-half-even (sbg.h:343-351),
 * UNORM8 with stochastic rounding at sweep step `step` (msbg4.cpp:829-837);
 * the other storages round to nearest even in their own format. */
static inline void store_float(const or_level *L, int64_t page, int vid, float v, uint32_t step) {
  size_t nv = (size_t)L->bsx * L->bsx * L->bsx;
  size_t i = (size_t)page * nv + (size_t)vid;
  switch (L->storage) {
    case OR_U16: ((uint16_t *)L->pages)[i] = or_u16_from_float_rne(v); break;
    case OR_F32: ((float *)L->pages)[i] = v; break;
    case OR_F16: ((uint16_t *)L->pages)[i] = or_f16_from_float(v); break;
    case OR_BF16: ((uint16_t *)L->pages)[i] = or_bf16_from_float(v); break;
    case OR_F64: ((double *)L->pages)[i] = (double)v; break;
    case OR_U8:
      ((uint8_t *)L->pages)[i] =
          or_u8_from_float_sr(v, or_sr_uniform(step, (uint32_t)page, (uint32_t)vid));
      break;
  }
}

/* applyChannelPdeFast<T> with laplTyp = -(PDE_MEAN_CURVATURE +
 * OPT_8_COLOR_SCHEME) (msbg4.cpp:60-957): colour mode, in place, colours
 * 0..7 per sweep, level-0 blocks only, Neumann halo of width 1. */
int or_mcf_sweeps(const or_level *L, const int32_t *blocks, int n,
                  const uint16_t *levels, int numIter, float dt,
                  int doConstrZeroOne, double *residual, int nthreads) {
  return or_mcf_sweeps_level(L, NULL, blocks, n, levels, 0, OR_OPT_BC_NEUMANN, numIter, dt,
                             doConstrZeroOne, residual, nthreads);
}

int or_mcf_sweeps_level(const or_level *L, const or_level *Lo, const int32_t *blocks, int n,
                        const uint16_t *levels, int sweepLevel, unsigned opt, int numIter,
                        float dt, int doConstrZeroOne, double *residual, int nthreads) {
  if (L->storage == OR_F64 && opt != OR_OPT_BC_NEUMANN) return -1;
  if ((opt & OR_OPT_BC_COARSE_LEVEL) && !Lo) return -1;
  int nMaxIter = numIter;
  if (nMaxIter % 2 == 1) nMaxIter++;
  const int b = L->bsx, hsx = b + 2, nh = hsx * hsx * hsx;
  const int nvb = b * b * b;
  const int isF64 = (L->storage == OR_F64);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int it = 0; it < nMaxIter; it++) {
    double rsum = 0.0;
    int64_t ncnt = 0;
    for (int pass = 0; pass < 8; pass++) {
#ifdef _OPENMP
#pragma omp parallel reduction(+ : rsum, ncnt)
#endif
      {
        float *hf = isF64 ? NULL : (float *)malloc(sizeof(float) * (size_t)(nh + 16));
        double *hd = isF64 ? (double *)malloc(sizeof(double) * (size_t)(nh + 16)) : NULL;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int i = 0; i < n; i++) {
          int bid = blocks[i];
          int bx = bid % L->nbx, by = (bid / L->nbx) % L->nby, bz = bid / (L->nbx * L->nby);
          if (or_color8(bx, by, bz) != pass) continue;
          if (levels && levels[bid] != sweepLevel) continue;
          int32_t pg = L->page[bid];
          if (pg < 0) continue; /* UT_ASSERT0(isValueBlock) in the reference */
          if (isF64) {
            or_halo_fill_f64(L, bid, 1, OR_OPT_BC_NEUMANN, hd);
            double ddt = (double)dt;
            for (int vz = 0, vid = 0; vz < b; vz++)
              for (int vy = 0; vy < b; vy++)
                for (int vx = 0; vx < b; vx++, vid++) {
                  int hid = (vx + 1) + (vy + 1) * hsx + (vz + 1) * hsx * hsx;
                  double ph = or_mcf_voxel_f64(hd, hsx, hid, ddt, doConstrZeroOne);
                  double d = ph - hd[hid];
                  rsum += d * d;
                  ((double *)L->pages)[(size_t)pg * nvb + vid] = ph;
                }
          } else {
            or_halo_fill(L, Lo, bid, 1, opt, hf);
            for (int vz = 0, vid = 0; vz < b; vz++)
              for (int vy = 0; vy < b; vy++)
                for (int vx = 0; vx < b; vx++, vid++) {
                  int hid = (vx + 1) + (vy + 1) * hsx + (vz + 1) * hsx * hsx;
                  float ph = or_mcf_voxel(hf, hsx, hid, dt, doConstrZeroOne);
                  store_float(L, pg, vid, ph, L->sr_step + (uint32_t)it);
                  if (residual) {
                    double nw = raw_to_float(L->storage, page_raw(L, pg, vid));
                    double d = nw - (double)hf[hid];
                    rsum += d * d;
                  }
                }
          }
          ncnt += nvb;
        }
        free(hf);
        free(hd);
      }
    }
    if (residual) residual[it] = ncnt ? sqrt(rsum / (double)ncnt) : 0.0;
  }
  return nMaxIter;
}

/* ------------------------------------------------------------------------ */
/* Other stencils and driver modes of applyChannelPdeFast, msbg4.cpp:60-957 */
/* ------------------------------------------------------------------------ */
float or_fast_exp(float x0) {
  /* fastmath.h:583-628: range reduction by ln 2 in two parts, a polynomial
   * whose second coefficient is added twice, scaling by 2^n through the
   * exponent bits, and the range selects */
  const float MAXLOGF = 88.72283905206835f, MINLOGF = -88.f;
  const float C1F = 0.693359375f, C2F = -2.12194440e-4f;
  const float PX1 = 1.9875691500E-4f, PX2 = 1.3981999507E-3f, PX3 = 8.3334519073E-3f,
              PX4 = 4.1665795894E-2f, PX5 = 1.6666665459E-1f, PX6 = 5.0000001201E-1f;
  const float LOG2EF = 1.44269504088896341f;
  float x = x0;
  float z = floorf(LOG2EF * x + 0.5f);
  x -= z * C1F;
  x -= z * C2F;
  const int n = (int)z;
  const float x2 = x * x;
  z = x * PX1;
  z += PX2;
  z *= x;
  z += PX2;
  z *= x;
  z += PX3;
  z *= x;
  z += PX4;
  z *= x;
  z += PX5;
  z *= x;
  z += PX6;
  z *= x2;
  z += x + 1.0f;
  union {
    uint32_t u;
    float f;
  } sc;
  sc.u = (uint32_t)(n + 0x7f) << 23;
  z *= sc.f;
  if (x0 < MINLOGF) z = 0.0f;
  if (x0 > MAXLOGF) z = INFINITY;
  return z;
}

/* G_ of msbg4.cpp:761-764 */
static inline float pm_weight(float f, float alpha) {
  if (alpha < 0) {
    const float s = fabsf(f) * (1.0f / -alpha);
    return or_fast_exp(-(s * s));
  }
  const float s = fabsf(f) / alpha;
  return 1.0f / (1.0f + s * s);
}

float or_pde_voxel(int op, const float *h, int hsx, int hid, float dt, float taubin,
                   float alpha, int iter) {
  const int s1 = hsx, s2 = hsx * hsx;
#define F_(dx, dy, dz) h[hid + (dx) + (dy)*s1 + (dz)*s2]
  const float F0 = F_(0, 0, 0), F1 = F_(-1, 0, 0), F2 = F_(1, 0, 0), F3 = F_(0, -1, 0),
              F4 = F_(0, 1, 0), F5 = F_(0, 0, -1), F6 = F_(0, 0, 1);
  float phi = F0;
  switch (op) {
    case 1: { /* msbg4.cpp:682-712 */
      float L = F1 + F2 + F3 + F4 + F5 + F6 - 6.0f * F0;
      float GX = 0.5f * (F2 - F1), GY = 0.5f * (F4 - F3), GZ = 0.5f * (F6 - F5);
      float G = sqrtf(GX * GX + GY * GY + GZ * GZ);
      float D = dt * L * G;
      if (taubin != 0.0f)
        phi = (iter & 1) ? phi + D : phi - taubin * D;
      else
        phi += D;
      break;
    }
    case 2: { /* msbg4.cpp:637-680, halo 2 */
      float H = 42.f * F0 -
                12.f * (F_(1, 0, 0) + F_(0, 1, 0) + F_(0, 0, 1) + F_(-1, 0, 0) + F_(0, -1, 0) +
                        F_(0, 0, -1)) +
                1.f * (F_(2, 0, 0) + F_(0, 2, 0) + F_(0, 0, 2) + F_(-2, 0, 0) + F_(0, -2, 0) +
                       F_(0, 0, -2)) +
                2.f * (F_(1, 0, 1) + F_(0, 1, 1) + F_(-1, 0, 1) + F_(0, -1, 1) + F_(1, 0, -1) +
                       F_(0, 1, -1) + F_(-1, 0, -1) + F_(0, -1, -1) + F_(1, 1, 0) +
                       F_(-1, 1, 0) + F_(-1, -1, 0) + F_(1, -1, 0));
      float GX = 0.5f * (F2 - F1), GY = 0.5f * (F4 - F3), GZ = 0.5f * (F6 - F5);
      float G = sqrtf(GX * GX + GY * GY + GZ * GZ);
      float D = -dt * H * G;
      D = vminf(vmaxf(D, -0.1f), 0.1f);
      phi = F0 + D;
      break;
    }
    case 3: { /* msbg4.cpp:713-735 */
      float S = (1.0f / 6.0f) * (F1 + F2 + F3 + F4 + F5 + F6);
      float D = dt * (S - phi);
      phi += D;
      break;
    }
    case 14:
    case 15: { /* msbg4.cpp:736-790 */
      float g1 = F1 - F0, g2 = F2 - F0, g3 = F3 - F0, g4 = F4 - F0, g5 = F5 - F0, g6 = F6 - F0;
      float G1 = pm_weight(g1, alpha), G2 = pm_weight(g2, alpha), G3 = pm_weight(g3, alpha),
            G4 = pm_weight(g4, alpha), G5 = pm_weight(g5, alpha), G6 = pm_weight(g6, alpha);
      if (op == 15) {
        float sum = G1 * F1 + G2 * F2 + G3 * F3 + G4 * F4 + G5 * F5 + G6 * F6;
        float wsum = G1 + G2 + G3 + G4 + G5 + G6;
        phi = (phi + dt * sum) * 1.0f / (1.0f + dt * wsum);
      } else {
        float sum = G1 * g1 + G2 * g2 + G3 * g3 + G4 * g4 + G5 * g5 + G6 * g6;
        phi += dt * sum;
      }
      break;
    }
    case 17: /* msbg4.cpp:791-805 */
      phi = (phi + dt * (F1 + F2 + F3 + F4 + F5 + F6)) * 1.0f / (1.0f + dt * 6.0f);
      break;
  }
#undef F_
  return phi;
}

static int or_pde_halo(int op) { return op == 2 ? 2 : 1; }

int or_pde_sweeps(const or_level *L, const int32_t *blocks, int n, const uint16_t *levels,
                  int numIter, const or_pde_params *P, double *residual, int nthreads) {
  const int op = P->op;
  if (!(op == 1 || op == 2 || op == 3 || op == 4 || op == 14 || op == 15 || op == 17)) return -1;
  if (L->storage == OR_F64 && op != 4) return -1;
  const int nMaxIter0 = numIter;
  int nMaxIter = numIter;
  if (nMaxIter % 2 == 1) nMaxIter++; /* msbg4.cpp:150-152 */
  const int b = L->bsx, hw = or_pde_halo(op), hsx = b + 2 * hw, nh = hsx * hsx * hsx;
  const int64_t nvb = (int64_t)b * b * b, nb = (int64_t)L->nbx * L->nby * L->nbz;
  const int esz = (int)storage_size(L->storage);
  const int hard = P->densThr > 1e-12f || P->densThrTol1 > 1e-12f; /* msbg4.cpp:192-193 */
  int64_t npages = 0;
  for (int64_t i = 0; i < nb; i++)
    if (L->page[i] >= 0 && L->page[i] + 1 > npages) npages = L->page[i] + 1;
  /* swept blocks: listed, level 0, value blocks */
  uint8_t *swept = (uint8_t *)calloc((size_t)nb, 1);
  for (int i = 0; i < n; i++) {
    const int bid = blocks[i];
    if (L->page[bid] < 0) continue;
    if (levels && levels[bid] != 0) continue;
    swept[bid] = 1;
  }
  /* phi0: the field at entry (hard constraints), as floats per page */
  float *phi0 = NULL;
  if (hard) {
    phi0 = (float *)malloc(sizeof(float) * (size_t)(npages * nvb + 1));
    for (int64_t p = 0; p < npages; p++)
      for (int64_t v = 0; v < nvb; v++)
        phi0[p * nvb + v] = raw_to_float(L->storage, page_raw(L, p, (int)v));
  }
  /* Jacobi: a temporary level with the same page numbering; value blocks
   * that are not swept are NULL in it (msbg4.cpp:285-297) */
  or_level T = *L;
  int32_t *tpage = NULL;
  void *tpages = NULL;
  if (P->mode == OR_MODE_JACOBI) {
    tpage = (int32_t *)malloc(sizeof(int32_t) * (size_t)nb);
    for (int64_t i = 0; i < nb; i++)
      tpage[i] = L->page[i] >= 0 ? (swept[i] ? L->page[i] : OR_NULL) : L->page[i];
    tpages = calloc((size_t)(npages * nvb + 1), (size_t)esz);
    T.page = tpage;
    T.pages = tpages;
  }
  const int npass = P->mode == OR_MODE_COLOR8 ? 8 : (P->mode == OR_MODE_REDBLACK ? 2 : 1);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const or_level *src = L, *dst = L;
  for (int it = 0; it < nMaxIter; it++) {
    if (P->mode == OR_MODE_JACOBI) {
      src = (it & 1) ? &T : L;
      dst = (it & 1) ? L : &T;
    }
    const int copyOnly = P->mode == OR_MODE_JACOBI && it == nMaxIter - 1 && (nMaxIter0 % 2) == 1;
    double rsum = 0.0;
    int64_t ncnt = 0;
    for (int pass = 0; pass < npass; pass++) {
#ifdef _OPENMP
#pragma omp parallel reduction(+ : rsum, ncnt)
#endif
      {
        float *hf = (float *)malloc(sizeof(float) * (size_t)(nh + 16));
        double *hd = (double *)malloc(sizeof(double) * (size_t)(nh + 16));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int i = 0; i < n; i++) {
          const int bid = blocks[i];
          if (!swept[bid]) continue;
          const int bx = bid % L->nbx, by = (bid / L->nbx) % L->nby, bz = bid / (L->nbx * L->nby);
          if (P->mode == OR_MODE_COLOR8 && or_color8(bx, by, bz) != pass) continue;
          if (P->mode == OR_MODE_REDBLACK && ((bx + by + bz) & 1) == pass) continue;
          const int32_t pg = L->page[bid];
          if (copyOnly) { /* msbg4.cpp:481-487 */
            memcpy((char *)dst->pages + (size_t)pg * nvb * esz,
                   (const char *)src->pages + (size_t)pg * nvb * esz, (size_t)(nvb * esz));
            ncnt += nvb;
            continue;
          }
          if (L->storage == OR_F64) {
            or_halo_fill_f64(src, bid, 1, OR_OPT_BC_NEUMANN, hd);
            for (int vz = 0, vid = 0; vz < b; vz++)
              for (int vy = 0; vy < b; vy++)
                for (int vx = 0; vx < b; vx++, vid++) {
                  const int hid = (vx + 1) + (vy + 1) * hsx + (vz + 1) * hsx * hsx;
                  double ph = or_mcf_voxel_f64(hd, hsx, hid, (double)P->dt,
                                               hard ? 0 : P->doConstrZeroOne);
                  if (hard) {
                    const double p0 = phi0[(size_t)pg * nvb + vid];
                    const double lo = vmaxd(0.0, p0 - (double)P->densThrTol1),
                                 hi = vmind(1.0, p0 + (double)P->densThrTol2);
                    ph = vmaxd(vmind(ph, hi), lo);
                  }
                  const double d = ph - hd[hid];
                  rsum += d * d;
                  ((double *)dst->pages)[(size_t)pg * nvb + vid] = ph;
                }
          } else {
            or_halo_fill(src, NULL, bid, hw, OR_OPT_BC_NEUMANN, hf);
            for (int vz = 0, vid = 0; vz < b; vz++)
              for (int vy = 0; vy < b; vy++)
                for (int vx = 0; vx < b; vx++, vid++) {
                  const int hid = (vx + hw) + (vy + hw) * hsx + (vz + hw) * hsx * hsx;
                  float ph;
                  if (op == 4) {
                    ph = or_mcf_voxel(hf, hsx, hid, P->dt, hard ? 0 : P->doConstrZeroOne);
                  } else {
                    ph = or_pde_voxel(op, hf, hsx, hid, P->dt, P->taubin, P->alpha, it);
                    if (!hard && P->doConstrZeroOne) ph = vminf(vmaxf(ph, 0.0f), 1.0f);
                  }
                  if (hard) { /* msbg4.cpp:813-818 */
                    const float p0 = phi0[(size_t)pg * nvb + vid];
                    const float lo = vmaxf(0.0f, p0 - P->densThrTol1),
                                hi = vminf(1.0f, p0 + P->densThrTol2);
                    ph = vmaxf(vminf(ph, hi), lo);
                  }
                  store_float(dst, pg, vid, ph, L->sr_step + (uint32_t)it);
                  if (residual) {
                    const double nw = raw_to_float(L->storage, page_raw(dst, pg, vid));
                    const double d = nw - (double)hf[hid];
                    rsum += d * d;
                  }
                }
          }
          ncnt += nvb;
        }
        free(hf);
        free(hd);
      }
    }
    if (residual) residual[it] = ncnt ? sqrt(rsum / (double)ncnt) : 0.0;
  }
  free(swept);
  free(phi0);
  free(tpage);
  free(tpages);
  return nMaxIter;
}

/* ------------------------------------------------------------------------ */
/* A14: ghost downsample, msbg.cpp:298-343 (weights), 2902-2974             */
/* ------------------------------------------------------------------------ */
static void ghost_weights(float w[64]) {
  /* cell-centred kernel: goff = 0.5, dx in -1..2, weight
   * (1-0.5|2*0.5-(dx+0.5)|) per axis, normalised by the double sum. */
  double wd[64], wsum = 0;
  for (int dz = -1, ik = 0; dz <= 2; dz++)
    for (int dy = -1; dy <= 2; dy++)
      for (int dx = -1; dx <= 2; dx++, ik++) {
        double v = (1 - 0.5 * fabs(1.0 - (dx + 0.5))) * (1 - 0.5 * fabs(1.0 - (dy + 0.5))) *
                   (1 - 0.5 * fabs(1.0 - (dz + 0.5)));
        wd[ik] = v;
        wsum += v;
      }
  for (int i = 0; i < 64; i++) {
    float t = (float)wd[i];      /* weights[ik] = w (stored as float) */
    w[i] = (float)((double)t / wsum); /* weights[ik] /= wsum (double) */
  }
}

void or_ghost_downsample(const float *halo, int bsxHi, float *lo) {
  float w[64];
  ghost_weights(w);
  const int hsx = bsxHi + 2, hsx2 = hsx * hsx, bl = bsxHi / 2;
  for (int vz = 0; vz < bl; vz++)
    for (int vy = 0; vy < bl; vy++)
      for (int vx = 0; vx < bl; vx++) {
        float S[4][4] = {{0}};
        int hid = 2 * vx + 2 * vy * hsx + 2 * vz * hsx2;
        for (int k = 0; k < 4; k++, hid += hsx2)
          for (int j = 0; j < 4; j++)
            for (int l = 0; l < 4; l++) S[j][l] += w[16 * k + 4 * j + l] * halo[hid + j * hsx + l];
        float T[4];
        for (int l = 0; l < 4; l++) T[l] = ((S[0][l] + S[1][l]) + S[2][l]) + S[3][l];
        /* horizontal_add(Vec4f): (a0+a2)+(a1+a3) (vectorf128.h:870-883) */
        lo[(vz * bl + vy) * bl + vx] = (T[0] + T[2]) + (T[1] + T[3]);
      }
}

void or_prepare_ghost_blocks(or_level *lvl, int nLevels, const int32_t *fc,
                             const int32_t *fc_off, const uint16_t *flags) {
  for (int l = 0; l < nLevels - 1; l++) {
    or_level *L = &lvl[l], *Lo = &lvl[l + 1];
    int b = L->bsx, hsx = b + 2;
    float *halo = (float *)malloc(sizeof(float) * (size_t)(hsx * hsx * hsx + 16));
    int bl = b / 2;
    float *lo = (float *)malloc(sizeof(float) * (size_t)(bl * bl * bl));
    for (int i = fc_off[l]; i < fc_off[l + 1]; i++) {
      int bid = fc[i];
      if (!(flags[bid] & OR_BLK_EXISTS)) continue; /* setEmptyBlock at level+1 */
      or_halo_fill(L, Lo, bid, 1, OR_OPT_BC_COARSE_LEVEL, halo);
      or_ghost_downsample(halo, b, lo);
      int32_t pg = Lo->page[bid];
      if (pg < 0) continue;
      for (int v = 0; v < bl * bl * bl; v++) {
        size_t idx = (size_t)pg * bl * bl * bl + v;
        switch (Lo->storage) {
          case OR_F32: ((float *)Lo->pages)[idx] = lo[v]; break;
          case OR_F16: ((uint16_t *)Lo->pages)[idx] = or_f16_from_float(lo[v]); break;
          case OR_BF16: ((uint16_t *)Lo->pages)[idx] = or_bf16_from_float(lo[v]); break;
          case OR_U16: ((uint16_t *)Lo->pages)[idx] = or_u16_from_float_rne(lo[v]); break;
          case OR_F64: ((double *)Lo->pages)[idx] = lo[v]; break;
        }
      }
    }
    free(halo);
    free(lo);
  }

}

/* ------------------------------------------------------------------------ */
/* Field consumers: interpolation of a level and the demo's ray-march        */
/* renderer (msbg.h:469-484, sbg.h:3109-3245, 3903-4135, 4920-5031,          */
/* render.h:58-125, render.cpp:19-100).                                      */
/*                                                                          */
/* The reference compiles these with -O3 -march=native -ffast-math, so the  */
/* compiler contracts multiply-adds into FMAs where it sees fit. This       */
/* restatement fixes one choice, stated at each expression: vectorclass's    */
/* mul_add is an FMA (vectorf128.h:1133-1141 with __FMA__), "a*b + c" and    */
/* "a += b*c" contract into fma(a, b, c) / fma(b, c, a) with the first       */
/* product fused, everything else is rounded per operation.                 */
/* ------------------------------------------------------------------------ */

/* interpolant input at voxel (x,y,z): the raw UNORM16 integer for UNORM16
 * levels (SparseGrid<uint16_t> interpolates raw values and scales after,
 * msbg.h:478-484), the decoded value otherwise. NULL blocks read as the
 * empty value (getValuePtr returns NULL and the caller substitutes
 * _emptyValue, sbg.h:3211-3222; the fast paths would dereference NULL). */
static inline float ip_input(const or_level *L, int x, int y, int z) {
  int lb = ilog2i(L->bsx), m = L->bsx - 1;
  int32_t c = L->page[((int64_t)(z >> lb) * L->nby + (y >> lb)) * L->nbx + (x >> lb)];
  double raw;
  if (c >= 0)
    raw = page_raw(L, c, ((z & m) * L->bsx + (y & m)) * L->bsx + (x & m));
  else
    raw = c == OR_FULL ? L->full_raw : L->empty_raw;
  return L->storage == OR_U16 ? (float)raw : raw_to_float(L->storage, raw);
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* interpolateGen0<float,float,IP_LINEAR> border branch (sbg.h:3170-3245):
 * indices clamped to the domain; Interpolate1D (sbg.h:2934-2939)
 * p1*t + p0*(1-t) as fma(p1, t, p0*(1-t)). */
float or_interp_linear_border(const or_level *L, const float p[3]) {
  float q[3], t[3];
  int i0[3], s[3] = {L->sx, L->sy, L->sz};
  for (int k = 0; k < 3; k++) {
    q[k] = p[k] - 0.5f;
    i0[k] = (int)floorf(q[k]);
    t[k] = q[k] - (float)i0[k];
  }
  float bz[2];
  for (int k = 0; k < 2; k++) {
    int zz = clampi(i0[2] + k, 0, s[2] - 1);
    float by[2];
    for (int j = 0; j < 2; j++) {
      int yy = clampi(i0[1] + j, 0, s[1] - 1);
      float bx[2];
      for (int i = 0; i < 2; i++) bx[i] = ip_input(L, clampi(i0[0] + i, 0, s[0] - 1), yy, zz);
      by[j] = fmaf(bx[1], t[0], bx[0] * (1.0f - t[0]));
    }
    bz[k] = fmaf(by[1], t[1], by[0] * (1.0f - t[1]));
  }
  return fmaf(bz[1], t[2], bz[0] * (1.0f - t[2]));
}

/* interpolateUint16ToFloatFast / interpolateFloatFast (sbg.h:4920-5031,
 * 4623-4800): cell-centred trilinear interpolation; the lerps are
 * vlerp_ps = mul_add(b - a, t, a) in x, then y, then z; a stencil that
 * leaves [0, s-1] falls back to the border branch. Not scaled. */
float or_interp_linear(const or_level *L, const float p[3]) {
  float q[3], t[3];
  int i0[3], s[3] = {L->sx, L->sy, L->sz};
  for (int k = 0; k < 3; k++) {
    q[k] = p[k] - 0.5f;
    float f = floorf(q[k]);
    i0[k] = (int)f;
    t[k] = q[k] - f;
  }
  for (int k = 0; k < 3; k++)
    if (i0[k] < 0 || i0[k] >= s[k] - 1) return or_interp_linear_border(L, p);
  float c[8];
  for (int n = 0; n < 8; n++)
    c[n] = ip_input(L, i0[0] + (n & 1), i0[1] + ((n >> 1) & 1), i0[2] + (n >> 2));
  float x0 = fmaf(c[1] - c[0], t[0], c[0]), x1 = fmaf(c[3] - c[2], t[0], c[2]);
  float x2 = fmaf(c[5] - c[4], t[0], c[4]), x3 = fmaf(c[7] - c[6], t[0], c[6]);
  float y0 = fmaf(x1 - x0, t[1], x0), y1 = fmaf(x3 - x2, t[1], x2);
  return fmaf(y1 - y0, t[2], y0);
}

/* MtCubicBSplineWeights1D (mtool.h:344-363), products left to right */
void or_bspline3_weights(float t, float w[4], float d[4]) {
  w[0] = (1.f / 6.f) * (1.f - t) * (1.f - t) * (1.f - t);
  w[1] = (2.f / 3.f) - 0.5f * t * t * (2.f - t);
  w[2] = (2.f / 3.f) - 0.5f * (1.f - t) * (1.f - t) * (1.f + t);
  w[3] = (1.f / 6.f) * t * t * t;
  d[0] = -(1.f - t) * (1.f - t) / 2.f;
  d[1] = 0.5f * t * t - 1.0f * (2.f - t) * t;
  d[2] = 1.f * (1.f - t) * (t + 1.f) - 0.5f * (1.f - t) * (1.f - t);
  d[3] = t * t / 2.f;
}

/* interpolateWithDerivs0 with IP_BSPLINE_CUBIC, options 0
 * (sbg.h:3903-4135): 4^3 window from floor(p - 0.5) - 1, clamped to the
 * domain; value += w*f and grad += w'*f as FMAs in z, y, x order. */
void or_interp_bspline3(const or_level *L, const float p[3], float *val, float grad[3]) {
  float t[3], wv[3][4], wd[3][4];
  int i0[3], s[3] = {L->sx, L->sy, L->sz};
  for (int k = 0; k < 3; k++) {
    float q = p[k] - 0.5f, f = floorf(q);
    t[k] = q - f;
    i0[k] = (int)f - 1;
    or_bspline3_weights(t[k], wv[k], wd[k]);
  }
  float v = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
  for (int dz = 0; dz < 4; dz++) {
    int zz = clampi(i0[2] + dz, 0, s[2] - 1);
    for (int dy = 0; dy < 4; dy++) {
      int yy = clampi(i0[1] + dy, 0, s[1] - 1);
      float wgy = wv[2][dz] * wv[1][dy];
      float ax = wv[2][dz] * wv[1][dy], ay = wv[2][dz] * wd[1][dy], az = wd[2][dz] * wv[1][dy];
      for (int dx = 0; dx < 4; dx++) {
        float f = ip_input(L, clampi(i0[0] + dx, 0, s[0] - 1), yy, zz);
        float w = wgy * wv[0][dx];
        v = fmaf(w, f, v);
        gx = fmaf(ax * wd[0][dx], f, gx);
        gy = fmaf(ay * wv[0][dx], f, gy);
        gz = fmaf(az * wv[0][dx], f, gz);
      }
    }
  }
  *val = v;
  grad[0] = gx;
  grad[1] = gy;
  grad[2] = gz;
}

/* ---- RaymarchRenderer (render.h, render.cpp), in double ---- */
static inline double v3dot(const double a[3], const double b[3]) {
  /* (a.x*b.x + a.y*b.y) + a.z*b.z, both additions contracted */
  return fma(a[2], b[2], fma(a[0], b[0], a[1] * b[1]));
}
static inline void v3normalize(const double v[3], double o[3]) {
  double len = sqrt(v3dot(v, v));
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
static inline void v3cross(const double a[3], const double b[3], double o[3]) {
  o[0] = fma(a[1], b[2], -(a[2] * b[1]));
  o[1] = fma(a[2], b[0], -(a[0] * b[2]));
  o[2] = fma(a[0], b[1], -(a[1] * b[0]));
}

/* setCamera (render.h:66-77): forward, right, up */
void or_camera_basis(const double pos[3], const double look[3], double fwd[3], double right[3],
                     double up[3]) {
  const double worldUp[3] = {0.0, 1.0, 0.0};
  double d[3] = {look[0] - pos[0], look[1] - pos[1], look[2] - pos[2]}, r[3];
  v3normalize(d, fwd);
  v3cross(fwd, worldUp, r);
  v3normalize(r, right);
  v3cross(right, fwd, up);
}

/* sampleDensity (render.h:82-110) */
static double rd_sample(const or_level *L, int res, float scale, const double p[3]) {
  float pos[3];
  int ip[3], s[3] = {L->sx, L->sy, L->sz};
  for (int k = 0; k < 3; k++) {
    pos[k] = (float)res * (float)p[k];
    ip[k] = (int)floorf(pos[k]);
    if (ip[k] < 0 || ip[k] > s[k] - 1) return 0.0;
  }
  int lb = ilog2i(L->bsx);
  int32_t c = L->page[((int64_t)(ip[2] >> lb) * L->nby + (ip[1] >> lb)) * L->nbx + (ip[0] >> lb)];
  if (c == OR_EMPTY || c == OR_NULL) return 0.0; /* isEmptyBlock (sbg.h:1990-2000) */
  float f = or_interp_linear(L, pos);
  return (double)(f * scale);
}

/* sampleDensityGradient (render.h:112-125) */
static void rd_gradient(const or_level *L, int res, float scale, const float dmin[3],
                        const float dmax[3], const double p[3], double g[3]) {
  float pos[3], val, gr[3] = {0.0f, 0.0f, 0.0f};
  int in = 1;
  for (int k = 0; k < 3; k++) {
    pos[k] = (float)res * (float)p[k];
    if (pos[k] < dmin[k] || pos[k] > dmax[k]) in = 0;
  }
  if (in) {
    or_interp_bspline3(L, pos, &val, gr);
    for (int k = 0; k < 3; k++) gr[k] *= scale;
  }
  for (int k = 0; k < 3; k++) g[k] = -1.0 * (double)gr[k];
}

/* render one pixel (render.cpp:37-94) */
static uint32_t rd_pixel(const or_level *L, const or_camera *C, const double fwd[3],
                         const double right[3], const double up[3], const float dmin[3],
                         const float dmax[3], int W, int H, int x, int y) {
  const int res = L->sx > L->sy ? (L->sx > L->sz ? L->sx : L->sz) : (L->sy > L->sz ? L->sy : L->sz);
  const float scale = L->storage == OR_U16 ? 1.0f / 65535.0f : 1.0f;
  const double h = 1.0 / (double)res;
  const double ISO = 0.5;
  const int maxSteps = (int)(2.0 / h);
  const double invW = 1.0 / (double)W, invH = 1.0 / (double)H, aspect = (double)W / (double)H;
  /* ((x + 0.5f)*invW - 0.5f)*2*aspect and (0.5f - (y + 0.5f)*invH)*2, the
   * pixel centre in float, the rest in double */
  double ndcX = fma((double)((float)x + 0.5f), invW, -0.5) * 2.0 * aspect;
  double ndcY = fma(-(double)((float)y + 0.5f), invH, 0.5) * 2.0;
  double dir0[3], dir[3];
  for (int k = 0; k < 3; k++) dir0[k] = fma(up[k], ndcY, fma(fwd[k], C->focal, right[k] * ndcX));
  v3normalize(dir0, dir);
  double pos[3] = {C->pos[0], C->pos[1], C->pos[2]}, prevPos[3] = {pos[0], pos[1], pos[2]};
  double prevD = rd_sample(L, res, scale, pos), hitPos[3] = {0, 0, 0};
  int hit = 0;
  for (int step = 0; step < maxSteps; ++step) {
    for (int k = 0; k < 3; k++) pos[k] = fma(dir[k], h, pos[k]);
    double d = rd_sample(L, res, scale, pos);
    if (prevD < ISO && d >= ISO) {
      double t = (ISO - prevD) / (d - prevD);
      for (int k = 0; k < 3; k++) hitPos[k] = fma(pos[k] - prevPos[k], t, prevPos[k]);
      hit = 1;
      break;
    }
    prevD = d;
    for (int k = 0; k < 3; k++) prevPos[k] = pos[k];
  }
  uint32_t out = 0x02000000u;
  if (hit) {
    double g[3], n[3], l0[3], l[3];
    rd_gradient(L, res, scale, dmin, dmax, hitPos, g);
    v3normalize(g, n);
    for (int k = 0; k < 3; k++) l0[k] = C->sun[k] - hitPos[k];
    v3normalize(l0, l);
    double diff = v3dot(n, l);
    diff = diff < 0.0 ? 0.0 : diff; /* std::max(dot, 0.0) */
    uint32_t b8[3];
    for (int k = 0; k < 3; k++) {
      double c = C->color[k] * diff;
      c = c < 0.0 ? 0.0 : (1.0 < c ? 1.0 : c); /* std::clamp */
      b8[k] = (uint32_t)(uint8_t)fma(c, 255.0, 0.5);
    }
    out |= (b8[2] << 16) | (b8[1] << 8) | b8[0];
  }
  return out;
}

void or_render(const or_level *L, const or_camera *C, int W, int H, int y0, int y1,
               uint32_t *rgb, int nthreads) {
  double fwd[3], right[3], up[3];
  float dmin[3], dmax[3];
  or_camera_basis(C->pos, C->look, fwd, right, up);
  or_domain_bounds(L->sx, L->sy, L->sz, dmin, dmax);
  if (nthreads <= 0) nthreads = omp_get_max_threads(); /* OMP_NUM_THREADS */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int y = y0; y < y1; y++)
    for (int x = 0; x < W; x++)
      rgb[(size_t)(y - y0) * W + x] = rd_pixel(L, C, fwd, right, up, dmin, dmax, W, H, x, y);
}

/* batch form of the interpolators: kind 0 = trilinear (val only),
 * 1 = cubic B-spline (val and grad[3*i..]) */
void or_interp_points(const or_level *L, int kind, const float *pts, int64_t n, float *val,
                      float *grad) {
  for (int64_t i = 0; i < n; i++) {
    if (kind == 0) {
      val[i] = or_interp_linear(L, pts + 3 * i);
    } else {
      float g[3];
      or_interp_bspline3(L, pts + 3 * i, &val[i], g);
      if (grad)
        for (int k = 0; k < 3; k++) grad[3 * i + k] = g[k];
    }
  }
}
