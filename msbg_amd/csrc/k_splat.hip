// Particle splat and finalize kernels (msbg_demo.cpp:541-680).
//
// The reference splats colour by colour with one task per home block so
// that plain read-modify-writes never race (msbg_demo.cpp:541-636). Here the
// particles are binned into cells and each cell's footprints are min'ed into
// an LDS tile, merged into the pages colour by colour (k_splat_tile). Wider
// footprints take k_splat: one thread per particle with atomics. Both give
// the reference's result: min is order-independent, and the block value-ness
// that decides the early abandon does not change while splatting.
//
// Reference paths are relative to the reference's src/ directory.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <type_traits>

#include "internal.h"
#include "storage.h"

namespace msbg {

namespace {

constexpr int kThreads = 256;

// atomic min on a 16-bit element through a CAS on its aligned 32-bit word
__device__ inline void atomic_min_u16(uint16_t *addr, uint16_t v) {
  uintptr_t a = (uintptr_t)addr;
  uint32_t *w = (uint32_t *)(a & ~(uintptr_t)3);
  const int sh = (int)(a & 2) * 8;
  uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    uint32_t cur = (old >> sh) & 0xffffu;
    if (cur <= v) return;
    uint32_t nw = (old & ~(0xffffu << sh)) | ((uint32_t)v << sh);
    uint32_t prev = atomicCAS(w, old, nw);
    if (prev == old) return;
    old = prev;
  }
}

template <int ESZ>
__device__ inline void atomic_min_code(uint8_t *elem, uint16_t v) {
  if constexpr (ESZ == 2) {
    atomic_min_u16((uint16_t *)elem, v);
  } else {
    // FP32 / FP64 pages hold the code in the low 32-bit word of the element
    uint32_t *w = (uint32_t *)elem;
    if (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > v) atomicMin(w, (uint32_t)v);
  }
}

// msbg_demo.cpp:571-633 for one particle.
template <int ESZ>
__global__ void k_splat(const float *__restrict__ xyz, int64_t n, float rScan, float distSqMax,
                        float distSqMaxInv, int sx, int sy, int sz, int lg, int nbx, int nby,
                        const int32_t *__restrict__ pagemap, uint8_t *__restrict__ pages) {
#pragma clang fp contract(off)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  const int smax[3] = {sx - 1, sy - 1, sz - 1};
  int lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int a = (int)__builtin_ceilf(p[k] - rScan - 0.5f);
    int b = (int)__builtin_floorf(p[k] + rScan - 0.5f);
    lo[k] = a > 0 ? a : 0;
    hi[k] = b < smax[k] ? b : smax[k];
  }
  const float x0 = p[0] - 0.5f, y0 = p[1] - 0.5f, z0 = p[2] - 0.5f;
  const int msk = (1 << lg) - 1;
  const int64_t pageElems = (int64_t)1 << (3 * lg);
  for (int iz = lo[2]; iz <= hi[2]; iz++) {
    float dz = (float)iz - z0;
    for (int iy = lo[1]; iy <= hi[1]; iy++) {
      float dy = (float)iy - y0;
      float dzy = dz * dz + dy * dy;
      int64_t ibzy = ((int64_t)(iz >> lg) * nby + (iy >> lg)) * nbx;
      int ivzy = (((iz & msk) << lg) + (iy & msk)) << lg;
      for (int ix = lo[0]; ix <= hi[0]; ix++) {
        float dx = (float)ix - x0;
        float d2 = dzy + dx * dx;
        if (d2 > distSqMax) continue;
        int32_t c = pagemap[ibzy + (ix >> lg)];
        if (c < 0) return;  // !isValueBlock -> goto loop_exit
        float f = d2 * distSqMaxInv;
        uint16_t v = (uint16_t)(int)__builtin_roundf(f * 65535.0f);  // renderDensFromFloat_
        atomic_min_code<ESZ>(pages + (c * pageElems + ivzy + (ix & msk)) * ESZ, v);
      }
    }
  }
}

// ---- binned splat: LDS tiles, scattered colour by colour ------------------
// Particles are binned into cells of kCell^3 voxels (counting sort). A
// workgroup takes one cell: every footprint of its particles lies in the
// cell grown by kHalo voxels per side (a kTile^3 tile of u32 codes in LDS),
// so the particles min their codes into LDS with ds_min_u32, one wave per
// particle and one lane per footprint voxel. The tile is then merged into
// the pages with plain loads and stores: the cells of one colour (cell
// coordinate parities) are 2 kCell apart and kCell >= 2 kHalo, so their
// tiles are disjoint, and the 8 colours run as 8 launches in stream order.
// The particle's walk in z,y,x order with its early abandon
// (msbg_demo.cpp:615-632) becomes a ballot over the footprint in the same
// order: lanes before the first in-range voxel of a non-value block apply,
// and the particle ends there.
constexpr int kCell = 16, kHalo = 5, kTile = kCell + 2 * kHalo, kTileVox = kTile * kTile * kTile;
constexpr int kTileBlk = 125;  // blocks a tile can touch: (kTile/8 + 1)^3 at 8^3 blocks
constexpr uint32_t kNoCode = 0xffffu;

// Cells per axis (ncx..) and the cell keys: cells are numbered in 4x4x4
// groups (nsx.. groups per axis, x fastest within and across groups), so
// consecutive keys, and the particles sorted by them, stay close in space.
// Keys of the partial groups at the domain's far faces are never used.
struct CellGrid {
  int ncx, ncy, ncz, sx, sy, sz, nsx, nsy, nsz;
  int64_t nkeys;
};

__host__ __device__ inline int cell_key(const CellGrid &G, int cx, int cy, int cz) {
  return ((((cz >> 2) * G.nsy + (cy >> 2)) * G.nsx + (cx >> 2)) << 6) | ((cz & 3) << 4) |
         ((cy & 3) << 2) | (cx & 3);
}

__device__ inline void key_cell(const CellGrid &G, int key, int &cx, int &cy, int &cz) {
  const int s = key >> 6, l = key & 63;
  cx = (s % G.nsx) * 4 + (l & 3);
  cy = ((s / G.nsx) % G.nsy) * 4 + ((l >> 2) & 3);
  cz = (s / (G.nsx * G.nsy)) * 4 + (l >> 4);
}

__device__ inline int cell_of(const CellGrid &G, float x, float y, float z) {
  auto ax = [](float p, int s, int n) {
    float f = __builtin_floorf(p);
    f = __builtin_fminf(__builtin_fmaxf(f, 0.0f), (float)(s - 1));  // NaN -> 0
    int c = (int)f / kCell;
    return c < n ? c : n - 1;
  };
  return cell_key(G, ax(x, G.sx, G.ncx), ax(y, G.sy, G.ncy), ax(z, G.sz, G.ncz));
}

__global__ void k_cell_count(const float *__restrict__ xyz, int64_t n, CellGrid G,
                             uint32_t *__restrict__ cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicAdd(&cnt[cell_of(G, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2])], 1u);
}

__global__ void k_cell_scatter(const float *__restrict__ xyz, int64_t n, CellGrid G,
                               const uint32_t *__restrict__ off, uint32_t *__restrict__ cursor,
                               float *__restrict__ sorted) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  const int c = cell_of(G, x, y, z);
  const uint32_t at = off[c] + atomicAdd(&cursor[c], 1u);
  sorted[3 * (int64_t)at] = x;
  sorted[3 * (int64_t)at + 1] = y;
  sorted[3 * (int64_t)at + 2] = z;
}

// selects the non-empty cells of one colour (cell coordinate parities)
struct CellOfColour {
  const uint32_t *cnt;
  CellGrid G;
  int col;
  __device__ bool operator()(int c) const {
    int cx, cy, cz;
    key_cell(G, c, cx, cy, cz);
    return ((cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2)) == col && cnt[c] != 0;
  }
};

constexpr int kSplatThreads = 1024, kMergeBatch = 9;
// footprint boxes of the tile path have at most 9 voxels a side
// (2 rScan + 1 with rScan + 0.5 <= kHalo - 0.25): 81 <= 2 * 64 per face
constexpr int kFaceSets = 2;
static_assert((2 * kHalo - 1) * (2 * kHalo - 1) <= kFaceSets * 64, "face sets");

template <int ESZ>
__global__ __launch_bounds__(kSplatThreads) void k_splat_tile(
    const float *__restrict__ sorted, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ cnt, const int32_t *__restrict__ list, CellGrid G, float rScan,
    float distSqMax, float distSqMaxInv, int lg, int nbx, int nby,
    const int32_t *__restrict__ pagemap, uint8_t *__restrict__ pages) {
#pragma clang fp contract(off)
  __shared__ uint32_t tile[kTileVox];
  __shared__ int32_t bcode[kTileBlk];
  const int cell = list[blockIdx.x];
  int cx, cy, cz;
  key_cell(G, cell, cx, cy, cz);
  const int tx0 = cx * kCell - kHalo, ty0 = cy * kCell - kHalo, tz0 = cz * kCell - kHalo;
  // blocks under the in-domain part of the tile
  const int bxa = (tx0 > 0 ? tx0 : 0) >> lg, bya = (ty0 > 0 ? ty0 : 0) >> lg,
            bza = (tz0 > 0 ? tz0 : 0) >> lg;
  const int bxb = (tx0 + kTile - 1 < G.sx - 1 ? tx0 + kTile - 1 : G.sx - 1) >> lg;
  const int byb = (ty0 + kTile - 1 < G.sy - 1 ? ty0 + kTile - 1 : G.sy - 1) >> lg;
  const int bzb = (tz0 + kTile - 1 < G.sz - 1 ? tz0 + kTile - 1 : G.sz - 1) >> lg;
  const int nbtx = bxb - bxa + 1, nbty = byb - bya + 1, nbtz = bzb - bza + 1;
  for (int e = threadIdx.x; e < kTileVox; e += blockDim.x) tile[e] = kNoCode;
  for (int e = threadIdx.x; e < nbtx * nbty * nbtz; e += blockDim.x) {
    const int ex = e % nbtx, ey = (e / nbtx) % nbty, ez = e / (nbtx * nbty);
    bcode[e] = pagemap[((int64_t)(bza + ez) * nby + (bya + ey)) * nbx + (bxa + ex)];
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const uint32_t p0 = off[cell], np = cnt[cell];
  const int smax[3] = {G.sx - 1, G.sy - 1, G.sz - 1};
  // the wave's particles are wave, wave + nwaves, ...: each lane loads one
  // of the next 64, then the wave takes them one by one from its lanes
  auto lane_f = [](float v, int k) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), k));
  };
  for (uint32_t qb = wave; qb < np; qb += 64u * nwaves) {
    const uint32_t ql = qb + (uint32_t)lane * nwaves;
    float lx = 0.0f, ly = 0.0f, lz = 0.0f;
    if (ql < np) {
      const float *pp = sorted + 3 * (int64_t)(p0 + ql);
      lx = pp[0];
      ly = pp[1];
      lz = pp[2];
    }
    const uint32_t left = (np - qb + nwaves - 1) / nwaves;
    const int nq = left < 64u ? (int)left : 64;
    for (int kq = 0; kq < nq; kq++) {
      const float p[3] = {lane_f(lx, kq), lane_f(ly, kq), lane_f(lz, kq)};
      int lo[3], hi[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        int a = (int)__builtin_ceilf(p[k] - rScan - 0.5f);
        int b = (int)__builtin_floorf(p[k] + rScan - 0.5f);
        lo[k] = a > 0 ? a : 0;
        hi[k] = b < smax[k] ? b : smax[k];
      }
      const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
      if (nx <= 0 || ny <= 0 || nz <= 0) continue;
      const int nxy = nx * ny;
      const float x0 = p[0] - 0.5f, y0 = p[1] - 0.5f, z0 = p[2] - 0.5f;
      // lanes over the box's (y, x) face in y,x order (at most kFaceSets
      // sets of 64), the planes in z order: z,y,x order is lane order
      // within a plane. Per lane and set, what does not depend on z.
      // t -> (t / nx, t % nx) through (t + 0.5) / nx in float: exact for
      // t < 2^15 and nx <= kTile.
      const float rx = 1.0f / (float)nx;
      bool fv[kFaceSets];
      float fdy2[kFaceSets], fdx2[kFaceSets];
      int fblk[kFaceSets], ftile[kFaceSets];
#pragma unroll
      for (int u = 0; u < kFaceSets; u++) {
        const int t = u * 64 + lane;
        const int qy = (int)(((float)t + 0.5f) * rx);
        const int iy = lo[1] + qy, ix = lo[0] + (t - qy * nx);
        fv[u] = t < nxy;
        const float dy = (float)iy - y0, dx = (float)ix - x0;
        fdy2[u] = dy * dy;
        fdx2[u] = dx * dx;
        fblk[u] = ((iy >> lg) - bya) * nbtx + ((ix >> lg) - bxa);
        ftile[u] = (iy - ty0) * kTile + (ix - tx0);
      }
      const int nsets = (nxy + 63) >> 6;
      // the box spans at most two blocks in z: both codes per lane, read once
      const int bzlo = lo[2] >> lg;
      const int zb0 = (bzlo - bza) * (nbty * nbtx), zb1 = ((hi[2] >> lg) - bza) * (nbty * nbtx);
      int fc0[kFaceSets], fc1[kFaceSets];
#pragma unroll
      for (int u = 0; u < kFaceSets; u++) {
        fc0[u] = fv[u] ? bcode[zb0 + fblk[u]] : 0;
        fc1[u] = fv[u] ? bcode[zb1 + fblk[u]] : 0;
      }
      bool done = false;
      for (int iz = lo[2]; iz <= hi[2] && !done; iz++) {
        const float dz = (float)iz - z0, dz2 = dz * dz;
        const bool zlo = (iz >> lg) == bzlo;
        const int tz = (iz - tz0) * (kTile * kTile);
#pragma unroll
        for (int u = 0; u < kFaceSets; u++) {
          if (u >= nsets || done) break;
          const float dzy = dz2 + fdy2[u];  // dz*dz + dy*dy, then + dx*dx
          const float d2 = dzy + fdx2[u];
          const bool inR = fv[u] && !(d2 > distSqMax);
          const int c = zlo ? fc0[u] : fc1[u];
          const uint64_t bad = __ballot(inR && c < 0);
          const int lim = bad ? __builtin_ctzll(bad) : 64;
          if (inR && lane < lim) {
            const float f = d2 * distSqMaxInv;
            const uint32_t v = (uint32_t)(uint16_t)(int)__builtin_roundf(f * 65535.0f);
            atomicMin(&tile[tz + ftile[u]], v);
          }
          done = bad != 0;  // !isValueBlock -> goto loop_exit
        }
      }
    }
  }
  __syncthreads();

  // merge the touched voxels (only value blocks were touched), the loads of
  // kMergeBatch voxels in flight together
  const int64_t pageElems = (int64_t)1 << (3 * lg);
  const int msk = (1 << lg) - 1;
  for (int e0 = threadIdx.x; e0 < kTileVox; e0 += kMergeBatch * kSplatThreads) {
    uint32_t v[kMergeBatch], cur[kMergeBatch];
    int64_t el[kMergeBatch];
#pragma unroll
    for (int u = 0; u < kMergeBatch; u++) {
      const int e = e0 + u * kSplatThreads;
      v[u] = e < kTileVox ? tile[e] : kNoCode;
      el[u] = 0;
      if (v[u] != kNoCode) {
        const int tz = e / (kTile * kTile), r = e - tz * (kTile * kTile), ty = r / kTile;
        const int ix = tx0 + (r - ty * kTile), iy = ty0 + ty, iz = tz0 + tz;
        const int c =
            bcode[(((iz >> lg) - bza) * nbty + ((iy >> lg) - bya)) * nbtx + ((ix >> lg) - bxa)];
        el[u] = c * pageElems + ((((iz & msk) << lg) + (iy & msk)) << lg) + (ix & msk);
      }
    }
#pragma unroll
    for (int u = 0; u < kMergeBatch; u++) {
      cur[u] = kNoCode;
      if (v[u] != kNoCode) {
        if constexpr (ESZ == 2)
          cur[u] = ((const uint16_t *)pages)[el[u]];
        else  // FP32 / FP64 pages hold the code in the low 32-bit word of the element
          cur[u] = *(const uint32_t *)(pages + el[u] * ESZ);
      }
    }
#pragma unroll
    for (int u = 0; u < kMergeBatch; u++) {
      if (v[u] == kNoCode || v[u] >= cur[u]) continue;
      if constexpr (ESZ == 2)
        ((uint16_t *)pages)[el[u]] = (uint16_t)v[u];
      else
        *(uint32_t *)(pages + el[u] * ESZ) = v[u];
    }
  }
}

// distance codes start at 65535 (UNORM 1.0, msbg_demo.cpp:494-496) in the
// pages of the non-UNORM16 storages
template <typename W>
__global__ void k_init_codes(W *__restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = (W)65535u;
}

// msbg_demo.cpp:664-678, MT_LINSTEPF (mtool.h:400-409)
__device__ inline float finalize_value(uint32_t u, float rParticle, float nbDist, float distSqMax,
                                       bool *isZero) {
#pragma clang fp contract(off)
  float distSq = (float)u * (1.0f / 65535.0f);
  *isZero = !(distSq < 0.999f);
  if (*isZero) return 0.0f;
  distSq *= distSqMax;
  float dist = __builtin_sqrtf(distSq) - rParticle;
  float a = -rParticle, b = rParticle + nbDist, ls;
  if (dist < a)
    ls = 0.0f;
  else if (dist >= b)
    ls = 1.0f;
  else
    ls = (dist - a) / (b - a);
  return 1.0f - ls;
}

// one element: the code in its storage word -> the density in the same word
template <int ST, typename Wd>
__device__ inline Wd finalize_word(Wd w, float rParticle, float nbDist, float distSqMax) {
#pragma clang fp contract(off)
  bool z;
  if constexpr (ST == MSBG_STORAGE_U16) {
    float f = finalize_value(w, rParticle, nbDist, distSqMax, &z);
    return z ? (Wd)0 : (Wd)(uint16_t)(int)__builtin_roundf(f * 65535.0f);
  } else if constexpr (ST == MSBG_STORAGE_F16 || ST == MSBG_STORAGE_BF16) {
    float f = finalize_value(w, rParticle, nbDist, distSqMax, &z);
    return Stor<ST>::enc(z ? 0.0f : f);
  } else if constexpr (ST == MSBG_STORAGE_F32) {
    float f = finalize_value(w, rParticle, nbDist, distSqMax, &z);
    return __builtin_bit_cast(uint32_t, z ? 0.0f : f);
  } else {
    float f = finalize_value((uint32_t)w, rParticle, nbDist, distSqMax, &z);
    return __builtin_bit_cast(uint64_t, z ? 0.0 : (double)f);
  }
}

// 16 bytes of elements per thread and step; n is a whole number of pages
template <int ST>
__global__ void k_finalize(uint8_t *__restrict__ pages, int64_t n, float rParticle, float nbDist,
                           float distSqMax) {
  using Wd = typename std::conditional<
      ST == MSBG_STORAGE_F32, uint32_t,
      typename std::conditional<ST == MSBG_STORAGE_F64, uint64_t, uint16_t>::type>::type;
  constexpr int K = 16 / (int)sizeof(Wd);
  const int64_t nv = n / K;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = i; v < nv; v += stride) {
    const uint4 q = ((const uint4 *)pages)[v];
    Wd e[K];
    __builtin_memcpy(e, &q, 16);
#pragma unroll
    for (int k = 0; k < K; k++) e[k] = finalize_word<ST>(e[k], rParticle, nbDist, distSqMax);
    uint4 o;
    __builtin_memcpy(&o, e, 16);
    ((uint4 *)pages)[v] = o;
  }
  for (int64_t t = nv * K + i; t < n; t += stride)
    ((Wd *)pages)[t] = finalize_word<ST>(((Wd *)pages)[t], rParticle, nbDist, distSqMax);
}

inline unsigned grid_cap(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

CellGrid cell_grid(const msbg_grid *g) {
  CellGrid G{};
  G.ncx = (g->sx + kCell - 1) / kCell;
  G.ncy = (g->sy + kCell - 1) / kCell;
  G.ncz = (g->sz + kCell - 1) / kCell;
  G.sx = g->sx;
  G.sy = g->sy;
  G.sz = g->sz;
  G.nsx = (G.ncx + 3) / 4;
  G.nsy = (G.ncy + 3) / 4;
  G.nsz = (G.ncz + 3) / 4;
  G.nkeys = (int64_t)G.nsx * G.nsy * G.nsz * 64;
  return G;
}

int splat_binned(msbg_grid *g, ChanLevel &cl, int esz, float rScan, float distSqMax,
                 float distSqMaxInv) {
  hipStream_t s = g->stream;
  MSBG_TRY(ensure_binned(g));
  const CellGrid G = cell_grid(g);
  const int64_t nc = G.nkeys;
  // the even cells (colour 0) are the most numerous colour
  const int64_t cap = (int64_t)((G.ncx + 1) / 2) * ((G.ncy + 1) / 2) * ((G.ncz + 1) / 2);
  DevBuf<uint32_t> ncol;
  DevBuf<int32_t> lists;
  MSBG_TRY(ncol.alloc(8));
  MSBG_TRY(lists.alloc(8 * cap));
  MSBG_HIP(hipMemsetAsync(ncol.p, 0, sizeof(uint32_t) * 8, s));
  const uint32_t *cnt = g->bin_cnt.p, *off = g->bin_off.p;
  const float *sorted = g->bin_xyz.p;
  // one list of non-empty cells per colour
  hipcub::CountingInputIterator<int> cells(0);
  size_t sbytes = 0;
  MSBG_HIP(hipcub::DeviceSelect::If(nullptr, sbytes, cells, lists.p, ncol.p, (int)nc,
                                    CellOfColour{cnt, G, 0}, s));
  DevBuf<uint8_t> swork;
  MSBG_TRY(swork.alloc(sbytes > 0 ? sbytes : 1));
  for (int col = 0; col < 8; col++)
    MSBG_HIP(hipcub::DeviceSelect::If(swork.p, sbytes, cells, lists.p + col * cap, ncol.p + col,
                                      (int)nc, CellOfColour{cnt, G, col}, s));
  uint32_t h[8];
  MSBG_HIP(hipMemcpyAsync(h, ncol.p, sizeof(h), hipMemcpyDeviceToHost, s));
  MSBG_HIP(hipStreamSynchronize(s));
  for (int col = 0; col < 8; col++) {
    if (!h[col]) continue;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(h[col]), dim3(kSplatThreads), 0, s, sorted, off, cnt,
                         lists.p + col * cap, G, rScan, distSqMax, distSqMaxInv, g->lg0, g->nbx,
                         g->nby, cl.d_page.p, cl.d_pages.p);
    };
    if (esz == 2)
      go(k_splat_tile<2>);
    else if (esz == 4)
      go(k_splat_tile<4>);
    else
      go(k_splat_tile<8>);
    MSBG_HIP(hipGetLastError());
  }
  // the work buffers are freed on return
  MSBG_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace

int ensure_binned(msbg_grid *g) {
  if (g->bin_valid) return 0;
  hipStream_t s = g->stream;
  const int64_t n = g->npoints;
  const CellGrid G = cell_grid(g);
  const int64_t nc = G.nkeys;
  if (n >= ((int64_t)1 << 31) || nc >= ((int64_t)1 << 31)) {
    set_error("binning: %lld particles / %lld cells exceed the 32-bit offsets", (long long)n,
              (long long)nc);
    return MSBG_EUNSUPP;
  }
  DevBuf<uint32_t> cursor;
  MSBG_TRY(g->bin_cnt.ensure(nc));
  MSBG_TRY(g->bin_off.ensure(nc));
  MSBG_TRY(cursor.alloc(nc));
  MSBG_TRY(g->bin_xyz.ensure(3 * std::max<int64_t>(n, 1)));
  MSBG_HIP(hipMemsetAsync(g->bin_cnt.p, 0, sizeof(uint32_t) * nc, s));
  MSBG_HIP(hipMemsetAsync(cursor.p, 0, sizeof(uint32_t) * nc, s));
  if (n) {
    const unsigned gp = (unsigned)((n + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_cell_count, dim3(gp), dim3(kThreads), 0, s, g->d_xyz.p, n, G,
                       g->bin_cnt.p);
    MSBG_HIP(hipGetLastError());
  }
  size_t bytes = 0;
  MSBG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, g->bin_cnt.p, g->bin_off.p, (int)nc,
                                            s));
  DevBuf<uint8_t> work;
  MSBG_TRY(work.alloc(bytes > 0 ? bytes : 1));
  MSBG_HIP(hipcub::DeviceScan::ExclusiveSum(work.p, bytes, g->bin_cnt.p, g->bin_off.p, (int)nc,
                                            s));
  if (n) {
    const unsigned gp = (unsigned)((n + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_cell_scatter, dim3(gp), dim3(kThreads), 0, s, g->d_xyz.p, n, G,
                       g->bin_off.p, cursor.p, g->bin_xyz.p);
    MSBG_HIP(hipGetLastError());
  }
  // the work buffers are freed on return
  MSBG_HIP(hipStreamSynchronize(s));
  g->bin_valid = true;
  return 0;
}

int launch_splat(msbg_grid *g, Channel &ch, float rScan) {
  ChanLevel &cl = ch.lv[0];
  int64_t nv = (int64_t)cl.bsx * cl.bsx * cl.bsx;
  int64_t nel = cl.npages * nv;
  if (ch.storage != MSBG_STORAGE_U16 && nel) {
    switch (ch.storage) {
      case MSBG_STORAGE_F16:
      case MSBG_STORAGE_BF16:
        hipLaunchKernelGGL(k_init_codes<uint16_t>, dim3(grid_cap(nel)), dim3(kThreads), 0,
                           g->stream, (uint16_t *)cl.d_pages.p, nel);
        break;
      case MSBG_STORAGE_F32:
        hipLaunchKernelGGL(k_init_codes<uint32_t>, dim3(grid_cap(nel)), dim3(kThreads), 0,
                           g->stream, (uint32_t *)cl.d_pages.p, nel);
        break;
      case MSBG_STORAGE_F64:
        hipLaunchKernelGGL(k_init_codes<uint64_t>, dim3(grid_cap(nel)), dim3(kThreads), 0,
                           g->stream, (uint64_t *)cl.d_pages.p, nel);
        break;
    }
    MSBG_HIP(hipGetLastError());
  }
  if (!g->npoints) return 0;
  const float distSqMax = rScan * rScan, distSqMaxInv = 1.0f / distSqMax;
  int esz = storage_bytes(ch.storage);
  // the binned splat needs every footprint inside its cell's tile, with a
  // quarter voxel to spare for the float rounding of the footprint bounds
  if (rScan + 0.5f <= (float)kHalo - 0.25f && g->npoints < ((int64_t)1 << 31) && esz != 1)
    return splat_binned(g, cl, esz, rScan, distSqMax, distSqMaxInv);
  int64_t nb = (g->npoints + kThreads - 1) / kThreads;
  auto args = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(kThreads), 0, g->stream, g->d_xyz.p,
                       g->npoints, rScan, distSqMax, distSqMaxInv, g->sx, g->sy, g->sz, g->lg0,
                       g->nbx, g->nby, cl.d_page.p, cl.d_pages.p);
  };
  if (esz == 2)
    args(k_splat<2>);
  else if (esz == 4)
    args(k_splat<4>);
  else
    args(k_splat<8>);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_finalize(msbg_grid *g, Channel &ch, float rParticle, float nbDist) {
  ChanLevel &cl = ch.lv[0];
  int64_t n = cl.npages * (int64_t)cl.bsx * cl.bsx * cl.bsx;
  if (!n) return 0;
  float rScan = rParticle + nbDist;
  float distSqMax = rScan * rScan;  // MT::sqf(rParticle + nbDist)
  switch (ch.storage) {
    case MSBG_STORAGE_U16:
      hipLaunchKernelGGL(k_finalize<MSBG_STORAGE_U16>, dim3(grid_cap(n)), dim3(kThreads), 0,
                         g->stream, cl.d_pages.p, n, rParticle, nbDist, distSqMax);
      break;
    case MSBG_STORAGE_F16:
      hipLaunchKernelGGL(k_finalize<MSBG_STORAGE_F16>, dim3(grid_cap(n)), dim3(kThreads), 0,
                         g->stream, cl.d_pages.p, n, rParticle, nbDist, distSqMax);
      break;
    case MSBG_STORAGE_BF16:
      hipLaunchKernelGGL(k_finalize<MSBG_STORAGE_BF16>, dim3(grid_cap(n)), dim3(kThreads), 0,
                         g->stream, cl.d_pages.p, n, rParticle, nbDist, distSqMax);
      break;
    case MSBG_STORAGE_F32:
      hipLaunchKernelGGL(k_finalize<MSBG_STORAGE_F32>, dim3(grid_cap(n)), dim3(kThreads), 0,
                         g->stream, cl.d_pages.p, n, rParticle, nbDist, distSqMax);
      break;
    case MSBG_STORAGE_F64:
      hipLaunchKernelGGL(k_finalize<MSBG_STORAGE_F64>, dim3(grid_cap(n)), dim3(kThreads), 0,
                         g->stream, cl.d_pages.p, n, rParticle, nbDist, distSqMax);
      break;
  }
  MSBG_HIP(hipGetLastError());
  return 0;
}

}  // namespace msbg
