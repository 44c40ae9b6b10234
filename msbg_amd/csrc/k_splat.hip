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
#include <cmath>
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
// workgroup takes a cell: every footprint of its particles lies in the cell
// grown by H voxels per side (H from rScan, a tile of u32 codes in LDS), so
// the particles min their codes into the tile with ds_min_u32, one wave per
// particle and one lane per (y, x) position of its footprint box. The tile
// is then merged into the pages with plain loads and stores: the cells of
// one colour (cell coordinate parities) are 2 kCell apart and the tile is at
// most 2 kCell wide, so the tiles of one colour are disjoint, and the 8
// colours run as 8 launches in stream order. Workgroups stay resident and
// walk their colour's cells in turn, fetching the next cell's block codes
// and particles while they merge the current one.
// The particle's walk in z,y,x order with its early abandon
// (msbg_demo.cpp:615-632): a footprint whose blocks all hold a page never
// abandons; any other becomes a ballot over the footprint in the same
// order: lanes before the first in-range voxel of a block without a page
// apply, and the particle ends there.
constexpr int kCell = 16, kMaxHalo = 4;
constexpr int kTileBlk = 125;  // blocks a tile can touch: (32/8 + 1)^3 at 8^3 blocks
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

constexpr int kSplatThreads = 1024;
// particles set up per round of the walk, and their set-up words in LDS:
// info, magic, tile origin, x and y origin (ints), x0, y0, dz^2 of up to 9
// planes, the position (floats)
constexpr uint32_t kRound = 64;
constexpr int kParWords = 5;
// the wave that sets rounds up (it holds no merge item)
constexpr int kSetupWave = kSplatThreads / 64 - 1;
// the merge takes a pair of voxels in x over kMergeZ planes per thread
constexpr int kMergeZ = 8;
// footprint faces of at most (2 kMaxHalo + 1)^2 voxels: kFaceSets sets of 64 lanes
constexpr int kFaceSets = 2;
static_assert((2 * kMaxHalo + 1) * (2 * kMaxHalo + 1) <= kFaceSets * 64, "face sets");

// The tile of halo H: x grown by HX (H rounded up to even, so that voxel
// pairs in x stay aligned to pairs of page elements), y and z by H. Footprint
// boxes have at most 2H + 1 voxels a side.
template <int H>
struct TileDims {
  static constexpr int HX = (H + 1) & ~1;
  static constexpr int TX = kCell + 2 * HX, TY = kCell + 2 * H, TZ = kCell + 2 * H;
  static constexpr int PLANE = TX * TY, NV = PLANE * TZ, MAXPL = 2 * H + 1;
  static constexpr int PARTS = (TZ + kMergeZ - 1) / kMergeZ, ITEMS = (TX / 2) * TY * PARTS;
  static_assert(TX <= 2 * kCell && TY <= 2 * kCell, "tiles of one colour overlap");
  static_assert((TX / 8 + 1) * (TY / 8 + 1) * (TZ / 8 + 1) <= kTileBlk, "tile blocks");
  static_assert(ITEMS <= kSplatThreads - 64, "one merge item per thread, none in the set-up wave");
  static_assert((MAXPL - 1) * PLANE * 4 < 65536, "plane offsets");
  static_assert(MAXPL <= 9, "dz^2 words of a particle's set-up");
  static_assert(NV % 4 == 0 && TX % 2 == 0, "tile fill, voxel pairs");
};

// one cell of a colour: its key, first particle and particle count
struct CellMeta {
  int key;
  uint32_t off, cnt;
  int pad;
};

__global__ void k_cell_meta(const int32_t *__restrict__ list, int n, const uint32_t *__restrict__ off,
                            const uint32_t *__restrict__ cnt, CellMeta *__restrict__ meta) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const int k = list[i];
  meta[i] = CellMeta{k, off[k], cnt[k], 0};
}

typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));

__device__ inline int lane_i(int v, int k) { return __builtin_amdgcn_readlane(v, k); }

// The UNORM16 distance code of an in-range voxel, (uint16_t)roundf(f * 65535)
// with f = d2 * distSqMaxInv (renderDensFromFloat_, sbg.h:272-315). For
// x = f * 65535 in [0, 65536], trunc(x + 0.49999997f) in float is exactly
// roundf(x); out-of-range voxels (d2 > distSqMax, or +inf) give 65535 or more
// and are clamped to 65535, which is the tile's empty code and changes
// nothing. codeScale = 65535 * distSqMaxInv when distSqMaxInv is a power of
// two: then d2 * distSqMaxInv is exact and one multiply rounds as two do.
template <bool ONE_MUL>
__device__ inline uint32_t splat_code(float d2, float distSqMaxInv, float codeScale) {
#pragma clang fp contract(off)
  const float x = ONE_MUL ? d2 * codeScale : (d2 * distSqMaxInv) * 65535.0f;
  return (uint32_t)__builtin_fminf(x + 0.49999997f, 65535.5f);
}

// where a cell's tile lies and which blocks lie under it
struct TileGeo {
  int tx0, ty0, tz0, bxa, bya, bza, nbtx, nbty, nbt;
};

template <int H>
__device__ inline TileGeo tile_geo(const CellGrid &G, int key, int lg) {
  using T = TileDims<H>;
  int cx, cy, cz;
  key_cell(G, key, cx, cy, cz);
  TileGeo t;
  t.tx0 = cx * kCell - T::HX;
  t.ty0 = cy * kCell - H;
  t.tz0 = cz * kCell - H;
  t.bxa = (t.tx0 > 0 ? t.tx0 : 0) >> lg;
  t.bya = (t.ty0 > 0 ? t.ty0 : 0) >> lg;
  t.bza = (t.tz0 > 0 ? t.tz0 : 0) >> lg;
  const int bxb = (t.tx0 + T::TX - 1 < G.sx - 1 ? t.tx0 + T::TX - 1 : G.sx - 1) >> lg;
  const int byb = (t.ty0 + T::TY - 1 < G.sy - 1 ? t.ty0 + T::TY - 1 : G.sy - 1) >> lg;
  const int bzb = (t.tz0 + T::TZ - 1 < G.sz - 1 ? t.tz0 + T::TZ - 1 : G.sz - 1) >> lg;
  t.nbtx = bxb - t.bxa + 1;
  t.nbty = byb - t.bya + 1;
  t.nbt = t.nbtx * t.nbty * (bzb - t.bza + 1);
  return t;
}

// what a thread fetches for a cell ahead of its walk: the page code of one
// block under the tile and, in the set-up wave, one particle of the first round
struct CellFetch {
  int bval;
  float x, y, z;
};

template <int ESZ, int H>
__global__ __launch_bounds__(kSplatThreads, 8) void k_splat_tile(
    const float *__restrict__ sorted, const CellMeta *__restrict__ meta, int ncells, CellGrid G,
    float rScan, float distSqMax, float distSqMaxInv, float codeScale, int lg, int nbx, int nby,
    const int32_t *__restrict__ pagemap, uint8_t *__restrict__ pages) {
#pragma clang fp contract(off)
  using T = TileDims<H>;
  __shared__ __align__(16) uint32_t tile[T::NV];
  __shared__ int32_t bcode[2][kTileBlk];
  __shared__ uint4 par[kRound][kParWords];  // a round's particle set-up
  const int stride = (int)gridDim.x;
  int ci = (int)blockIdx.x;
  if (ci >= ncells) return;
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr uint32_t kWaves = kSplatThreads / 64;
  for (int e = tid; e < T::NV / 4; e += (int)blockDim.x)
    ((uint4 *)tile)[e] = make_uint4(kNoCode, kNoCode, kNoCode, kNoCode);

  auto fetch = [&](const CellMeta &m, const TileGeo &t) {
    CellFetch f{0, 0.0f, 0.0f, 0.0f};
    if (tid < t.nbt) {
      const int ex = tid % t.nbtx, ey = (tid / t.nbtx) % t.nbty, ez = tid / (t.nbtx * t.nbty);
      f.bval = pagemap[((int64_t)(t.bza + ez) * nby + (t.bya + ey)) * nbx + (t.bxa + ex)];
    }
    if (wave == kSetupWave && (uint32_t)lane < m.cnt) {
      const float *pp = sorted + 3 * (int64_t)(m.off + lane);
      f.x = pp[0];
      f.y = pp[1];
      f.z = pp[2];
    }
    return f;
  };
  // the meta of cell i: a vector load (odd lanes read cell i + stride), so
  // that it does not hold up the LDS waits of the merge, taken from lane 0
  // an iteration later
  auto meta_raw = [&](int i) {
    const int j = (lane & 1) && i + stride < ncells ? i + stride : i;
    return meta[j];
  };
  auto uni = [&](const CellMeta &v) {
    return CellMeta{lane_i(v.key, 0), (uint32_t)lane_i((int)v.off, 0),
                    (uint32_t)lane_i((int)v.cnt, 0), 0};
  };

  const int smax[3] = {G.sx - 1, G.sy - 1, G.sz - 1};
  // A round's set-up, one particle per lane of the set-up wave: its
  // footprint box (msbg_demo.cpp:571-614), whether its blocks (at most two
  // per axis, codes from the page map) all hold a page, so that the walk can
  // skip the early abandon, the tile origin and dz^2 of each plane.
  auto setup = [&](bool mine, float lx, float ly, float lz, const TileGeo &t) {
    const float lp[3] = {lx, ly, lz};
    int lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      int a = (int)__builtin_ceilf(lp[k] - rScan - 0.5f);
      int b = (int)__builtin_floorf(lp[k] + rScan - 0.5f);
      lo[k] = a > 0 ? a : 0;
      hi[k] = b < smax[k] ? b : smax[k];
    }
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    const bool live = mine && nx > 0 && ny > 0 && nz > 0;
    bool fast = live && __mul24(nx, ny) <= 64 && nz <= T::MAXPL &&
                (hi[0] >> lg) - (lo[0] >> lg) <= 1 && (hi[1] >> lg) - (lo[1] >> lg) <= 1 &&
                (hi[2] >> lg) - (lo[2] >> lg) <= 1;
    if (fast) {
      const int64_t xl = lo[0] >> lg, xh = hi[0] >> lg;
      const int64_t r[4] = {((int64_t)(lo[2] >> lg) * nby + (lo[1] >> lg)) * nbx,
                            ((int64_t)(lo[2] >> lg) * nby + (hi[1] >> lg)) * nbx,
                            ((int64_t)(hi[2] >> lg) * nby + (lo[1] >> lg)) * nbx,
                            ((int64_t)(hi[2] >> lg) * nby + (hi[1] >> lg)) * nbx};
      int c[8];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        c[2 * e] = pagemap[r[e] + xl];
        c[2 * e + 1] = pagemap[r[e] + xh];
      }
#pragma unroll
      for (int e = 0; e < 8; e++) fast = fast && c[e] >= 0;
    }
    uint32_t w[4 * kParWords];
#pragma unroll
    for (int k = 0; k < 4 * kParWords; k++) w[k] = 0u;
    // info, t -> t / nx as (t * magic) >> 16 (exact for t < 64 and nx <= 64;
    // 65536 / nx through the reciprocal is within 0.008 of the quotient,
    // whose fraction is 0 or at least 1/64, so the floor is exact), the tile
    // index of the box origin, the box's x and y origin
    w[0] = (uint32_t)((live ? 1 : 0) | (fast ? 2 : 0) | ((nx & 63) << 2) | ((ny & 63) << 8) |
                      ((nz & 63) << 14));
    w[1] = fast ? (uint32_t)(65536.0f * __builtin_amdgcn_rcpf((float)nx)) + 1u : 0u;
    w[2] = (uint32_t)(((lo[2] - t.tz0) * T::TY + (lo[1] - t.ty0)) * T::TX + (lo[0] - t.tx0));
    w[3] = (uint32_t)lo[0];
    w[4] = (uint32_t)lo[1];
    w[5] = __builtin_bit_cast(uint32_t, lx - 0.5f);
    w[6] = __builtin_bit_cast(uint32_t, ly - 0.5f);
    const float z0l = lz - 0.5f;
#pragma unroll
    for (int k = 0; k < T::MAXPL; k++) {
      const float dz = (float)(lo[2] + k) - z0l;
      w[7 + k] = __builtin_bit_cast(uint32_t, dz * dz);
    }
    w[16] = __builtin_bit_cast(uint32_t, lx);
    w[17] = __builtin_bit_cast(uint32_t, ly);
    w[18] = __builtin_bit_cast(uint32_t, lz);
#pragma unroll
    for (int k = 0; k < kParWords; k++)
      par[lane][k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
  };

  CellMeta mc = uni(meta_raw(ci));
  CellMeta mnRaw = ci + stride < ncells ? meta_raw(ci + stride) : CellMeta{0, 0, 0, 0};
  TileGeo tg = tile_geo<H>(G, mc.key, lg);
  CellFetch fc = fetch(mc, tg);
  // the first cell's first round; later cells' first rounds are set up
  // while the cell before them merges
  if (wave == kSetupWave) setup((uint32_t)lane < mc.cnt, fc.x, fc.y, fc.z, tg);
  const bool oneMul = codeScale != 0.0f;
  int buf = 0;
  for (;;) {
    int32_t *const bc = bcode[buf];
    if (tid < tg.nbt) bc[tid] = fc.bval;
    // the next cell's meta, loaded an iteration ago
    const CellMeta mn = ci + stride < ncells ? uni(mnRaw) : CellMeta{0, 0, 0, 0};
    __syncthreads();  // block codes in, the previous merge has emptied the tile

    // ---- walk, in rounds of kRound particles set up in LDS (the first
    // round before this cell's walk began), each wave walking particles
    // wave, wave + kWaves, ... of the round
    for (uint32_t rb = 0; rb < mc.cnt; rb += kRound) {
      if (rb) {
        __syncthreads();  // the previous round is walked
        if (wave == kSetupWave) {
          const uint32_t ql = rb + (uint32_t)lane;
          const bool mine = ql < mc.cnt;
          float lx = 0.0f, ly = 0.0f, lz = 0.0f;
          if (mine) {
            const float *pp = sorted + 3 * (int64_t)(mc.off + ql);
            lx = pp[0];
            ly = pp[1];
            lz = pp[2];
          }
          setup(mine, lx, ly, lz, tg);
        }
        __syncthreads();  // the round is set up
      }

      const uint32_t left = mc.cnt - rb;
      const int nr = left < kRound ? (int)left : (int)kRound;
      for (int j = wave; j < nr; j += (int)kWaves) {
        int lane = tid & 63;
        __asm__ volatile("" : "+v"(lane));
        const uint4 q0 = par[j][0];
        const int inf = __builtin_amdgcn_readfirstlane((int)q0.x);
        if (!(inf & 1)) continue;
        if (inf & 2) {
          // lanes over the box's (y, x) face, the planes in z order; lanes
          // past the face take +inf distances, whose codes change nothing
          const uint4 q1 = par[j][1], q2 = par[j][2], q3 = par[j][3];
          const uint32_t dzw[12] = {q1.w, q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w, 0, 0, 0};
          const int fnx = (inf >> 2) & 63, fny = (inf >> 8) & 63, fnz = (inf >> 14) & 63;
          const int qy = (int)(__umul24((uint32_t)lane, q0.y) >> 16);
          const int rx = lane - __mul24(qy, fnx);
          const bool on = lane < __mul24(fnx, fny);
          const float dx = (float)((int)q0.w + rx) - __builtin_bit_cast(float, q1.y);
          const float dy = (float)((int)q1.x + qy) - __builtin_bit_cast(float, q1.z);
          const float dx2 = dx * dx, dy2 = on ? dy * dy : __builtin_inff();
          uint32_t *const t = tile + (on ? (int)q0.z + qy * T::TX + rx : 0);
          auto planes = [&](auto one) {
#pragma unroll
            for (int k = 0; k < T::MAXPL; k++) {
              if (k >= fnz) break;
              const float dzy = __builtin_bit_cast(float, dzw[k]) + dy2;  // dz*dz + dy*dy, then + dx*dx
              atomicMin(t + k * T::PLANE,
                        splat_code<decltype(one)::value>(dzy + dx2, distSqMaxInv, codeScale));
            }
          };
          if (oneMul)
            planes(std::true_type{});
          else
            planes(std::false_type{});
          continue;
        }
        // a box that may touch a block without a page: lanes over the face
        // in y,x order (at most kFaceSets sets of 64), a ballot per plane
        const uint4 q4 = par[j][4];
        const float p[3] = {__builtin_bit_cast(float, q4.x), __builtin_bit_cast(float, q4.y),
                            __builtin_bit_cast(float, q4.z)};
        int blo[3], bhi[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          int a = (int)__builtin_ceilf(p[k] - rScan - 0.5f);
          int b = (int)__builtin_floorf(p[k] + rScan - 0.5f);
          blo[k] = a > 0 ? a : 0;
          bhi[k] = b < smax[k] ? b : smax[k];
        }
        const int bnx = bhi[0] - blo[0] + 1, bny = bhi[1] - blo[1] + 1;
        const int nxy = bnx * bny;
        const float x0 = p[0] - 0.5f, y0 = p[1] - 0.5f, z0 = p[2] - 0.5f;
        // t -> (t / nx, t % nx) through (t + 0.5) / nx in float: exact for
        // t < 2^15 and nx <= 64
        const float rx = 1.0f / (float)bnx;
        bool fv[kFaceSets];
        float fdy2[kFaceSets], fdx2[kFaceSets];
        int fblk[kFaceSets], ftile[kFaceSets];
#pragma unroll
        for (int u = 0; u < kFaceSets; u++) {
          const int t = u * 64 + lane;
          const int qy = (int)(((float)t + 0.5f) * rx);
          const int iy = blo[1] + qy, ix = blo[0] + (t - qy * bnx);
          fv[u] = t < nxy;
          const float dy = (float)iy - y0, dx = (float)ix - x0;
          fdy2[u] = dy * dy;
          fdx2[u] = dx * dx;
          fblk[u] = ((iy >> lg) - tg.bya) * tg.nbtx + ((ix >> lg) - tg.bxa);
          ftile[u] = (iy - tg.ty0) * T::TX + (ix - tg.tx0);
        }
        const int nsets = (nxy + 63) >> 6;
        // the box spans at most two blocks in z: both codes per lane, read once
        const int bzlo = blo[2] >> lg;
        const int zb0 = (bzlo - tg.bza) * (tg.nbty * tg.nbtx);
        const int zb1 = ((bhi[2] >> lg) - tg.bza) * (tg.nbty * tg.nbtx);
        int fc0[kFaceSets], fc1[kFaceSets];
#pragma unroll
        for (int u = 0; u < kFaceSets; u++) {
          fc0[u] = fv[u] ? bc[zb0 + fblk[u]] : 0;
          fc1[u] = fv[u] ? bc[zb1 + fblk[u]] : 0;
        }
        bool done = false;
        for (int iz = blo[2]; iz <= bhi[2] && !done; iz++) {
          const float dz = (float)iz - z0, dz2s = dz * dz;
          const bool zlo = (iz >> lg) == bzlo;
          const int tz = (iz - tg.tz0) * T::PLANE;
#pragma unroll
          for (int u = 0; u < kFaceSets; u++) {
            if (u >= nsets || done) break;
            const float dzy = dz2s + fdy2[u];  // dz*dz + dy*dy, then + dx*dx
            const float d2 = dzy + fdx2[u];
            const bool inR = fv[u] && !(d2 > distSqMax);
            const int c = zlo ? fc0[u] : fc1[u];
            const uint64_t bad = __ballot(inR && c < 0);
            const int lim = bad ? __builtin_ctzll(bad) : 64;
            if (inR && lane < lim)
              atomicMin(&tile[tz + ftile[u]], splat_code<false>(d2, distSqMaxInv, 0.0f));
            done = bad != 0;  // !isValueBlock -> goto loop_exit
          }
        }
      }
    }

    // fetch the next cell while this one merges
    const int cn = ci + stride;
    const bool more = cn < ncells;
    CellMeta mnnRaw{0, 0, 0, 0};
    TileGeo tn = tg;
    CellFetch fn{0, 0.0f, 0.0f, 0.0f};
    if (more) {
      if (cn + stride < ncells) mnnRaw = meta_raw(cn + stride);
      tn = tile_geo<H>(G, mn.key, lg);
      fn = fetch(mn, tn);
    }
    __syncthreads();  // every particle is in the tile

    // ---- merge: a thread per pair of voxels in x over kMergeZ planes of
    // the tile (only value blocks were touched). The tile's values are read
    // and the tile emptied first, then the pages' values of the touched pairs
    // are loaded together, then the smaller codes stored.
    // the set-up wave has no merge item: it sets up the next cell's first round
    if (wave == kSetupWave && more) setup((uint32_t)lane < mn.cnt, fn.x, fn.y, fn.z, tn);
    // (the thread's item is derived anew each cell, not kept live across the walk)
    int mt = tid;
    __asm__ volatile("" : "+v"(mt));
    if (mt < T::ITEMS) {
      constexpr int NP = T::TX / 2;
      const int part = mt / (NP * T::TY), pc = mt - part * (NP * T::TY);
      const int ty = pc / NP, jx = pc - ty * NP;
      const int ix = tg.tx0 + 2 * jx, iy = tg.ty0 + ty, tzb = part * kMergeZ;
      const int tbase = ty * T::TX + 2 * jx;
      // the pair's two codes packed in one word (codes are at most 0xffff)
      // planes past the tile (the last run only) read the tile's last plane,
      // which is this thread's own, and count as untouched
      uint32_t v[kMergeZ];
#pragma unroll
      for (int u = 0; u < kMergeZ; u++) {
        const int tz = tzb + u < T::TZ ? tzb + u : T::TZ - 1;
        const uint2 w = *(const uint2 *)&tile[tz * T::PLANE + tbase];
        v[u] = tzb + u < T::TZ ? w.x | (w.y << 16) : 0xffffffffu;
      }
#pragma unroll
      for (int u = 0; u < kMergeZ; u++) {
        const int tz = tzb + u < T::TZ ? tzb + u : T::TZ - 1;
        *(uint2 *)&tile[tz * T::PLANE + tbase] = make_uint2(kNoCode, kNoCode);
      }
      // voxels outside the domain are never touched; ix is even and the
      // domain's x size a multiple of the block size, so ix + 1 is inside too
      if (ix >= 0 && iy >= 0 && ix < G.sx && iy < G.sy) {
        const int msk = (1 << lg) - 1;
        const int colBlk = ((iy >> lg) - tg.bya) * tg.nbtx + ((ix >> lg) - tg.bxa);
        const int64_t colEl = ((int64_t)(iy & msk) << lg) + (ix & msk);
        // the run's planes in the domain lie in at most two blocks in z
        // (kMergeZ <= the block edge)
        int za = tg.tz0 + tzb, zb = za + kMergeZ - 1;
        za = za > 0 ? za : 0;
        zb = zb < G.sz - 1 ? zb : G.sz - 1;
        const int zbl = za >> lg;
        uint8_t *pb[2] = {pages, pages};
        if (za <= zb) {
          const int c0 = bc[(zbl - tg.bza) * (tg.nbty * tg.nbtx) + colBlk];
          const int c1 = bc[((zb >> lg) - tg.bza) * (tg.nbty * tg.nbtx) + colBlk];
          pb[0] = pages + ((int64_t)c0 * ((int64_t)1 << (3 * lg)) + colEl) * ESZ;
          pb[1] = pages + ((int64_t)c1 * ((int64_t)1 << (3 * lg)) + colEl) * ESZ;
        }
        // the pages' codes of the touched pairs, packed the same way
        uint8_t *qa[kMergeZ];
#pragma unroll
        for (int u = 0; u < kMergeZ; u++) {
          const int iz = tg.tz0 + tzb + u;
          qa[u] = pb[(iz >> lg) != zbl] + ((int64_t)(iz & msk) << (2 * lg)) * ESZ;
        }
        uint32_t cur[kMergeZ];
#pragma unroll
        for (int u = 0; u < kMergeZ; u++) {
          cur[u] = 0xffffffffu;
          if (v[u] != 0xffffffffu) {
            const uint8_t *q = qa[u];
            if constexpr (ESZ == 2) {
              cur[u] = *(const uint32_t *)q;
            } else if constexpr (ESZ == 4) {
              const uint2 w = *(const uint2 *)q;  // codes: the splat starts them at 65535
              cur[u] = w.x | (w.y << 16);
            } else {  // FP64 pages hold the code in the low 32-bit word of the element
              cur[u] = *(const uint32_t *)q | (*(const uint32_t *)(q + 8) << 16);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kMergeZ; u++) {
          // both codes' minima in one packed 16-bit min
          const uint32_t w = __builtin_bit_cast(
              uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, v[u]),
                                                  __builtin_bit_cast(u16x2, cur[u])));
          if (w == cur[u]) continue;
          uint8_t *q = qa[u];
          if constexpr (ESZ == 2) {
            *(uint32_t *)q = w;
          } else if constexpr (ESZ == 4) {
            *(uint2 *)q = make_uint2(w & 0xffffu, w >> 16);
          } else {
            *(uint32_t *)q = w & 0xffffu;
            *(uint32_t *)(q + 8) = w >> 16;
          }
        }
      }
    }
    if (!more) break;
    ci = cn;
    mc = mn;
    mnRaw = mnnRaw;
    tg = tn;
    fc = fn;
    buf ^= 1;
  }
}

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

// The tile halo for a footprint radius: every box lies within H voxels of
// its particle's cell when H > rScan - 0.5 plus the float rounding of the
// box bounds (two roundings at coordinates below 2^e: 2^(e-23)).
int splat_halo(const msbg_grid *g, float rScan) {
  const double m = (double)std::max(std::max(g->sx, g->sy), g->sz) + rScan + 1.0;
  const double slack = std::ldexp(1.0, (int)std::ceil(std::log2(m)) - 23);
  const int h = (int)std::floor((double)rScan - 0.5 + slack) + 1;
  return h > 1 ? h : 1;
}

int device_cu_count() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  return cus;
}

template <int ESZ, int H>
int launch_splat_tile(hipStream_t s, int ncells, const float *sorted, const CellMeta *meta,
                      const CellGrid &G, float rScan, float distSqMax, float distSqMaxInv,
                      float codeScale, const msbg_grid *g, const ChanLevel &cl) {
  auto kern = k_splat_tile<ESZ, H>;
  int per = 0;
  MSBG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, kSplatThreads, 0));
  const int nwg = std::max(1, std::min(ncells, (per > 0 ? per : 1) * device_cu_count()));
  hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(kSplatThreads), 0, s, sorted, meta, ncells, G,
                     rScan, distSqMax, distSqMaxInv, codeScale, g->lg0, g->nbx, g->nby,
                     cl.d_page.p, cl.d_pages.p);
  MSBG_HIP(hipGetLastError());
  return 0;
}

template <int ESZ>
int launch_splat_tile_h(int h, hipStream_t s, int ncells, const float *sorted,
                        const CellMeta *meta, const CellGrid &G, float rScan, float distSqMax,
                        float distSqMaxInv, float codeScale, const msbg_grid *g,
                        const ChanLevel &cl) {
  switch (h) {
    case 1: return launch_splat_tile<ESZ, 1>(s, ncells, sorted, meta, G, rScan, distSqMax, distSqMaxInv, codeScale, g, cl);
    case 2: return launch_splat_tile<ESZ, 2>(s, ncells, sorted, meta, G, rScan, distSqMax, distSqMaxInv, codeScale, g, cl);
    case 3: return launch_splat_tile<ESZ, 3>(s, ncells, sorted, meta, G, rScan, distSqMax, distSqMaxInv, codeScale, g, cl);
    case 4: return launch_splat_tile<ESZ, 4>(s, ncells, sorted, meta, G, rScan, distSqMax, distSqMaxInv, codeScale, g, cl);
  }
  set_error("splat: tile halo %d", h);
  return MSBG_EINVAL;
}

int splat_binned(msbg_grid *g, ChanLevel &cl, int esz, int halo, float rScan, float distSqMax,
                 float distSqMaxInv) {
  hipStream_t s = g->stream;
  MSBG_TRY(ensure_binned(g));
  const CellGrid G = cell_grid(g);
  // one multiply for the code when distSqMaxInv is a power of two: the
  // squared distances are 0 or at least 2^-50 (differences of floats from
  // whole voxel positions), so d2 * distSqMaxInv stays exact and normal
  int ex = 0;
  const float codeScale = std::frexp(distSqMaxInv, &ex) == 0.5f && ex > -60 && ex < 60
                              ? 65535.0f * distSqMaxInv
                              : 0.0f;
  const int64_t nc = G.nkeys;
  // the even cells (colour 0) are the most numerous colour
  const int64_t cap = (int64_t)((G.ncx + 1) / 2) * ((G.ncy + 1) / 2) * ((G.ncz + 1) / 2);
  DevBuf<uint32_t> ncol;
  DevBuf<int32_t> lists;
  DevBuf<CellMeta> meta;
  MSBG_TRY(ncol.alloc(8));
  MSBG_TRY(lists.alloc(8 * cap));
  MSBG_TRY(meta.alloc(cap));
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
    const int n = (int)h[col];
    if (!n) continue;
    hipLaunchKernelGGL(k_cell_meta, dim3((unsigned)((n + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, s, lists.p + col * cap, n, off, cnt, meta.p);
    MSBG_HIP(hipGetLastError());
    if (esz == 2)
      MSBG_TRY(launch_splat_tile_h<2>(halo, s, n, sorted, meta.p, G, rScan, distSqMax,
                                      distSqMaxInv, codeScale, g, cl));
    else if (esz == 4)
      MSBG_TRY(launch_splat_tile_h<4>(halo, s, n, sorted, meta.p, G, rScan, distSqMax,
                                      distSqMaxInv, codeScale, g, cl));
    else
      MSBG_TRY(launch_splat_tile_h<8>(halo, s, n, sorted, meta.p, G, rScan, distSqMax,
                                      distSqMaxInv, codeScale, g, cl));
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
  const int esz = storage_bytes(ch.storage);
  if (esz != 2 && esz != 4 && esz != 8) {  // codes need 16 bits (UNORM8 channels hold none)
    set_error("splat: storage %d cannot hold the distance codes", ch.storage);
    return MSBG_EUNSUPP;
  }
  // the binned splat needs every footprint inside its cell's tile
  const int halo = splat_halo(g, rScan);
  if (halo <= kMaxHalo && g->npoints < ((int64_t)1 << 31))
    return splat_binned(g, cl, esz, halo, rScan, distSqMax, distSqMaxInv);
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
