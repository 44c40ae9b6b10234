// Generic halo fill (test hook) and the cross-level ghost downsample.
//
// halo_value() is the per-voxel rule of HaloBlockSet::fillHaloBlock_ /
// getOutOfBlockValue (halo.cpp:267-479, halo.h:267-355) on device-side block
// maps of one channel level and its coarser neighbour level.
//
// Reference paths are relative to the reference's src/ directory.

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "internal.h"
#include "storage.h"

namespace msbg {

namespace {

struct LevelView {
  int sx, sy, sz, bsx, lg, nbx, nby, nbz;
  const int32_t *page;
  const uint8_t *pages;
  int storage;
  double empty_raw, full_raw;
};

struct GhostWeights {
  float w[64];
};

__device__ inline int code_at(const LevelView &L, int x, int y, int z) {
  return L.page[((int64_t)(z >> L.lg) * L.nby + (y >> L.lg)) * L.nbx + (x >> L.lg)];
}

__device__ inline int vid_at(const LevelView &L, int x, int y, int z) {
  int m = L.bsx - 1;
  return ((((z & m) << L.lg) + (y & m)) << L.lg) + (x & m);
}

__device__ inline bool in_range(const LevelView &L, int x, int y, int z) {
  return x >= 0 && x < L.sx && y >= 0 && y < L.sy && z >= 0 && z < L.sz;
}

__device__ inline void clip3(const LevelView &L, int &x, int &y, int &z) {
  x = x < 0 ? 0 : (x > L.sx - 1 ? L.sx - 1 : x);
  y = y < 0 ? 0 : (y > L.sy - 1 ? L.sy - 1 : y);
  z = z < 0 ? 0 : (z > L.sz - 1 ? L.sz - 1 : z);
}

__device__ inline double page_raw(const LevelView &L, int64_t page, int vid) {
  int64_t i = page * ((int64_t)1 << (3 * L.lg)) + vid;
  switch (L.storage) {
    case MSBG_STORAGE_U16: return ((const uint16_t *)L.pages)[i];
    case MSBG_STORAGE_F32: return ((const float *)L.pages)[i];
    case MSBG_STORAGE_F16: return f16_bits_to_float(((const uint16_t *)L.pages)[i]);
    case MSBG_STORAGE_BF16: return bf16_bits_to_float(((const uint16_t *)L.pages)[i]);
    case MSBG_STORAGE_U8: return ((const uint8_t *)L.pages)[i];
    default: return ((const double *)L.pages)[i];
  }
}

__device__ inline double block_raw(const LevelView &L, int c, int vid) {
  if (c >= 0) return page_raw(L, c, vid);
  return c == MSBG_BLOCK_EMPTY ? L.empty_raw : L.full_raw;
}

// transferValue: UNORM16 -> u*(1.f/65535) (halo.h:484-501), floats as is
__device__ inline float raw_to_float(int storage, double raw) {
  if (storage == MSBG_STORAGE_U16) return (float)(uint16_t)raw * (1.0f / 65535.0f);
  if (storage == MSBG_STORAGE_U8) return (float)(uint8_t)raw * (1.0f / 255.0f);
  return (float)raw;
}

// SparseGrid::getValue(int) (sbg.h:2622-2648)
__device__ inline double get_value_raw(const LevelView &L, int x, int y, int z) {
  int c = code_at(L, x, y, z);
  if (c == MSBG_BLOCK_NULL) return L.empty_raw;
  return block_raw(L, c, vid_at(L, x, y, z));
}

__device__ inline double to_raw_type(int storage, float v) {
  if (storage == MSBG_STORAGE_U16) return (double)(uint16_t)v;  // Data_T = uint16_t truncates
  return (double)v;
}

// SparseGrid<Data_T>::interpolateGhost (sbg.h:3002-3095), options 0: the
// arithmetic, with the corners supplied by corner(X, Y, Z, &val0, &have0),
// which returns the raw value the reference reads for that coarse voxel.
template <typename Corner>
__device__ inline double interpolate_ghost_with(int storage, float px, float py, float pz,
                                                Corner corner) {
#pragma clang fp contract(off)
  const float po = 0.5f;
  float x = px - po, y = py - po, z = pz - po;
  int ix = (int)floorf(x), iy = (int)floorf(y), iz = (int)floorf(z);
  float u = x - (float)ix, v = y - (float)iy, w = z - (float)iz;
  double bx[2], by[2], bzv[2];
  for (int k = 0; k < 2; k++) {
    for (int j = 0; j < 2; j++) {
      for (int i = 0; i < 2; i++) bx[i] = corner(ix + i, iy + j, iz + k);
      by[j] = to_raw_type(storage, (float)bx[1] * u + (float)bx[0] * (1.0f - u));
    }
    bzv[k] = to_raw_type(storage, (float)by[1] * v + (float)by[0] * (1.0f - v));
  }
  return to_raw_type(storage, (float)bzv[1] * w + (float)bzv[0] * (1.0f - w));
}

// The corners from the level's block map: out of range, the clipped
// voxel's value; in a NULL block, the value of the voxel holding the sample
// point (clipped); else the voxel itself.
__device__ double interpolate_ghost_raw(const LevelView &Lo, float px, float py, float pz) {
  double val0 = Lo.empty_raw;
  bool have0 = false;
  auto corner = [&](int X, int Y, int Z) -> double {
    if (!in_range(Lo, X, Y, Z)) {
      clip3(Lo, X, Y, Z);
      return get_value_raw(Lo, X, Y, Z);
    }
    const int c = code_at(Lo, X, Y, Z);
    if (c == MSBG_BLOCK_NULL) {
      if (!have0) {
        int qx = (int)floorf(px), qy = (int)floorf(py), qz = (int)floorf(pz);
        clip3(Lo, qx, qy, qz);
        val0 = get_value_raw(Lo, qx, qy, qz);
        have0 = true;
      }
      return val0;
    }
    return block_raw(Lo, c, vid_at(Lo, X, Y, Z));
  };
  return interpolate_ghost_with(Lo.storage, px, py, pz, corner);
}

// fillHaloBlock_ per-voxel rule + getOutOfBlockValue.
__device__ float halo_value(const LevelView &L, const LevelView &Lo, int x, int y, int z,
                            int px0, int py0, int pz0, int own, unsigned opt) {
#pragma clang fp contract(off)
  bool inR = in_range(L, x, y, z);
  if (inR) {
    int c = code_at(L, x, y, z);
    if (c != MSBG_BLOCK_NULL) return raw_to_float(L.storage, block_raw(L, c, vid_at(L, x, y, z)));
  }
  if (opt & MSBG_OPT_BC_NEUMANN) {
    int b = L.bsx;
    int cx = x < px0 ? px0 : (x > px0 + b - 1 ? px0 + b - 1 : x);
    int cy = y < py0 ? py0 : (y > py0 + b - 1 ? py0 + b - 1 : y);
    int cz = z < pz0 ? pz0 : (z > pz0 + b - 1 ? pz0 + b - 1 : z);
    return raw_to_float(L.storage, block_raw(L, own, vid_at(L, cx, cy, cz)));
  }
  if (opt & MSBG_OPT_BC_COARSE_LEVEL) {
    if (!inR) {
      clip3(L, x, y, z);
      int c = code_at(L, x, y, z);
      if (c != MSBG_BLOCK_NULL) return raw_to_float(L.storage, block_raw(L, c, vid_at(L, x, y, z)));
    }
    if (opt & MSBG_OPT_INTERPOLATE) {
      float qx = ((float)x + 0.5f) / 2, qy = ((float)y + 0.5f) / 2, qz = ((float)z + 0.5f) / 2;
      return raw_to_float(Lo.storage, interpolate_ghost_raw(Lo, qx, qy, qz));
    }
    int xl = x / 2, yl = y / 2, zl = z / 2;
    clip3(Lo, xl, yl, zl);
    // a NULL coarse block is an assertion failure in the reference
    // (UT_ASSERT2(pVal), halo.h:347); its empty value is read here
    return raw_to_float(Lo.storage, get_value_raw(Lo, xl, yl, zl));
  }
  return raw_to_float(L.storage, L.empty_raw);
}

__global__ void k_halo_fill(LevelView L, LevelView Lo, int bid, int dmax, unsigned opt,
                            float *__restrict__ out) {
  int hsx = L.bsx + 2 * dmax;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= hsx * hsx * hsx) return;
  int hx = t % hsx, hy = (t / hsx) % hsx, hz = t / (hsx * hsx);
  int bx = bid % L.nbx, by = (bid / L.nbx) % L.nby, bz = bid / (L.nbx * L.nby);
  int px0 = bx * L.bsx, py0 = by * L.bsx, pz0 = bz * L.bsx;
  int own = L.page[bid];
  out[t] = halo_value(L, Lo, px0 - dmax + hx, py0 - dmax + hy, pz0 - dmax + hz, px0, py0, pz0, own,
                      opt);
}

// downsampleGhostBlockFloat (msbg.cpp:2902-2974): coarse voxel (vx,vy,vz)
// sums the 4x4x4 halo window at halo index (2vx,2vy,2vz) with four row
// accumulators S0..S3 over k, then horizontal_add((S0+S1)+S2)+S3) =
// (a0+a2)+(a1+a3) (vectorf128.h:870-883).
template <int ST>
__global__ void k_ghost(LevelView L, LevelView Lo, GhostWeights W, const int32_t *__restrict__ bids,
                        uint8_t *__restrict__ loPages) {
#pragma clang fp contract(off)
  int bid = bids[blockIdx.y];
  int bl = L.bsx / 2;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bl * bl * bl) return;
  int vx = t % bl, vy = (t / bl) % bl, vz = t / (bl * bl);
  int bx = bid % L.nbx, by = (bid / L.nbx) % L.nby, bz = bid / (L.nbx * L.nby);
  int px0 = bx * L.bsx, py0 = by * L.bsx, pz0 = bz * L.bsx;
  int own = L.page[bid];
  float S[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int l = 0; l < 4; l++) S[j][l] = 0.0f;
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int l = 0; l < 4; l++) {
        // halo index (2v + l) -> grid coordinate pos0 - 1 + 2v + l
        float h = halo_value(L, Lo, px0 - 1 + 2 * vx + l, py0 - 1 + 2 * vy + j,
                             pz0 - 1 + 2 * vz + k, px0, py0, pz0, own, MSBG_OPT_BC_COARSE_LEVEL);
        S[j][l] = S[j][l] + W.w[16 * k + 4 * j + l] * h;
      }
  float T[4];
#pragma unroll
  for (int l = 0; l < 4; l++) T[l] = ((S[0][l] + S[1][l]) + S[2][l]) + S[3][l];
  float val = (T[0] + T[2]) + (T[1] + T[3]);
  int pg = Lo.page[bid];
  int64_t idx = (int64_t)pg * bl * bl * bl + t;
  if constexpr (ST == MSBG_STORAGE_F32)
    ((float *)loPages)[idx] = val;
  else
    ((uint16_t *)loPages)[idx] = Stor<ST>::enc(val);
}

// The same downsample with the halo staged once in LDS: one workgroup per
// fine-coarse block. Each fine voxel feeds eight coarse voxels, so the
// per-tap version above evaluates the halo rule 8x per voxel; here the
// (bsx+2)^3 halo is evaluated once, in z-chunks of 2*CZ+2 planes (the
// windows of CZ coarse planes), as float. The taps and their summation order
// are those of k_ghost, so the results are identical.
//
// Staging: the block's own voxels come from its page in 16-byte loads; the
// one-voxel shell resolves halo_value(..., MSBG_OPT_BC_COARSE_LEVEL) from two
// 27-entry tables of the block's neighbour codes at this level and the
// coarser one (the levels share one block grid), so every shell voxel becomes
// a constant or one page load, kGhostBatch of them issued together.
constexpr int kGhostThreads = 512, kGhostBatch = 8, kGhostVec = 4;

template <int ST, int BSX>
__global__ __launch_bounds__(kGhostThreads) void k_ghost_tile(LevelView L, LevelView Lo,
                                                              GhostWeights W,
                                                              const int32_t *__restrict__ bids,
                                                              uint8_t *__restrict__ loPages) {
#pragma clang fp contract(off)
  using R = typename Stor<ST>::raw;
  constexpr int H = BSX + 2, PL = H * H, BL = BSX / 2, CZ = BL < 4 ? BL : 4, NP = 2 * CZ + 2;
  constexpr int M = BSX - 1, LG = BSX == 4 ? 2 : (BSX == 8 ? 3 : (BSX == 16 ? 4 : 5));
  constexpr int ML = BL - 1, LGL = LG - 1;
  __shared__ float tile[NP * PL];
  __shared__ int32_t nbL[27], nbLo[27];
  const int bid = bids[blockIdx.x];
  const int bx = bid % L.nbx, by = (bid / L.nbx) % L.nby, bz = bid / (L.nbx * L.nby);
  if (threadIdx.x < 27) {
    const int k = threadIdx.x, x2 = bx + k % 3 - 1, y2 = by + (k / 3) % 3 - 1, z2 = bz + k / 9 - 1;
    const bool in = x2 >= 0 && x2 < L.nbx && y2 >= 0 && y2 < L.nby && z2 >= 0 && z2 < L.nbz;
    const int64_t b2 = ((int64_t)z2 * L.nby + y2) * L.nbx + x2;
    nbL[k] = in ? L.page[b2] : MSBG_BLOCK_OUT;
    nbLo[k] = in ? Lo.page[b2] : MSBG_BLOCK_OUT;
  }
  const int pg = Lo.page[bid];
  const R *Lp = (const R *)L.pages, *Lop = (const R *)Lo.pages;
  const float lEmpty = (float)L.empty_raw, lFull = (float)L.full_raw;
  const float loEmpty = (float)Lo.empty_raw, loFull = (float)Lo.full_raw;
  // the block-local voxel (x, y, z), each in [-1, BSX]: a page element or a
  // constant, as halo_value resolves it
  auto source = [&](int x, int y, int z, float &cv) -> const R * {
    int dx = x < 0 ? -1 : (x > M ? 1 : 0), dy = y < 0 ? -1 : (y > M ? 1 : 0),
        dz = z < 0 ? -1 : (z > M ? 1 : 0);
    int c = nbL[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1];
    if (c == MSBG_BLOCK_OUT) {
      // outside the domain: clip to it (the axes that leave it land on the
      // block's own face)
      if (dx && nbL[13 + dx] == MSBG_BLOCK_OUT) { x = x < 0 ? 0 : M; dx = 0; }
      if (dy && nbL[13 + 3 * dy] == MSBG_BLOCK_OUT) { y = y < 0 ? 0 : M; dy = 0; }
      if (dz && nbL[13 + 9 * dz] == MSBG_BLOCK_OUT) { z = z < 0 ? 0 : M; dz = 0; }
      c = nbL[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1];
    }
    const int k = (dz + 1) * 9 + (dy + 1) * 3 + dx + 1;
    if (c >= 0)
      return Lp + (int64_t)c * (BSX * BSX * BSX) + ((((z & M) << LG) + (y & M)) << LG) + (x & M);
    if (c != MSBG_BLOCK_NULL) {
      cv = c == MSBG_BLOCK_EMPTY ? lEmpty : lFull;
      return nullptr;
    }
    // NULL block: the coarser level's voxel (x/2, y/2, z/2) of the same
    // block (getValue: a NULL coarse block reads its empty value)
    const int c2 = nbLo[k];
    const int gx = bx * BSX + x, gy = by * BSX + y, gz = bz * BSX + z;
    const int xl = gx / 2, yl = gy / 2, zl = gz / 2;
    if (c2 >= 0)
      return Lop + (int64_t)c2 * (BL * BL * BL) + ((((zl & ML) << LGL) + (yl & ML)) << LGL) +
             (xl & ML);
    cv = c2 == MSBG_BLOCK_FULL ? loFull : loEmpty;
    return nullptr;
  };
  __syncthreads();
  // interior rows come from the block's own page in vector loads of VB bytes
  constexpr int VB = BSX * (int)sizeof(R) < 16 ? BSX * (int)sizeof(R) : 16;
  constexpr int VEC = VB / (int)sizeof(R), SEGS = BSX / VEC, PER = 4 * H - 4;
  using V = typename std::conditional<VB == 16, uint4, uint2>::type;
  static_assert(VB == 16 || VB == 8, "vector width");
  const R *ownp = nbL[13] >= 0 ? Lp + (int64_t)nbL[13] * (BSX * BSX * BSX) : nullptr;
  for (int vz0 = 0; vz0 < BL; vz0 += CZ) {
    const int zlo = 2 * vz0 - 1;  // block-local z of tile plane 0
    // tile planes [pa, pb) lie inside the block
    const int pa = zlo < 0 ? 1 : 0, pb = zlo + NP - 1 > M ? NP - 1 : NP;
    int nsrc = NP * PL;  // elements resolved through source()
    if (ownp) {
      const int nseg = (pb - pa) * BSX * SEGS;
      for (int s0 = threadIdx.x; s0 < nseg; s0 += kGhostVec * kGhostThreads) {
        V raw[kGhostVec];
        int dst[kGhostVec];
#pragma unroll
        for (int u = 0; u < kGhostVec; u++) {
          const int sg = s0 + u * kGhostThreads;
          dst[u] = -1;
          if (sg < nseg) {
            const int hzl = pa + sg / (BSX * SEGS), r = sg % (BSX * SEGS), y = r / SEGS;
            const int xs = (r % SEGS) * VEC;
            raw[u] = *(const V *)(ownp + ((zlo + hzl) * BSX + y) * BSX + xs);
            dst[u] = hzl * PL + (y + 1) * H + xs + 1;
          }
        }
#pragma unroll
        for (int u = 0; u < kGhostVec; u++) {
          if (dst[u] < 0) continue;
          R v[VEC];
          __builtin_memcpy(v, &raw[u], sizeof(V));
#pragma unroll
          for (int i = 0; i < VEC; i++) tile[dst[u] + i] = (float)Stor<ST>::dec(v[i]);
        }
      }
      // the rest: whole planes outside the block, the perimeter of the others
      nsrc = pa * PL + (pb - pa) * PER + (NP - pb) * PL;
    }
    for (int e0 = threadIdx.x; e0 < nsrc; e0 += kGhostBatch * kGhostThreads) {
      const R *src[kGhostBatch];
      float val[kGhostBatch];
      int dst[kGhostBatch];
#pragma unroll
      for (int u = 0; u < kGhostBatch; u++) {
        int e = e0 + u * kGhostThreads;
        src[u] = nullptr;
        val[u] = 0.0f;
        dst[u] = -1;
        if (e < nsrc) {
          int hzl, hy, hx;
          if (!ownp) {
            hzl = e / PL;
            hy = (e - hzl * PL) / H;
            hx = e - hzl * PL - hy * H;
          } else if (e < pa * PL) {
            hzl = 0;
            hy = e / H;
            hx = e - hy * H;
          } else if ((e -= pa * PL) < (pb - pa) * PER) {
            hzl = pa + e / PER;
            int q = e % PER;
            if (q < 2 * H) {
              hy = q < H ? 0 : H - 1;
              hx = q < H ? q : q - H;
            } else {
              q -= 2 * H;
              hy = 1 + q / 2;
              hx = (q & 1) ? H - 1 : 0;
            }
          } else {
            e -= (pb - pa) * PER;
            hzl = NP - 1;
            hy = e / H;
            hx = e - hy * H;
          }
          dst[u] = hzl * PL + hy * H + hx;
          src[u] = source(hx - 1, hy - 1, zlo + hzl, val[u]);
        }
      }
      R raw[kGhostBatch];
#pragma unroll
      for (int u = 0; u < kGhostBatch; u++)
        if (src[u]) raw[u] = *src[u];
#pragma unroll
      for (int u = 0; u < kGhostBatch; u++)
        if (dst[u] >= 0) tile[dst[u]] = src[u] ? (float)Stor<ST>::dec(raw[u]) : val[u];
    }
    __syncthreads();
    for (int o = threadIdx.x; o < CZ * BL * BL; o += kGhostThreads) {
      const int vx = o % BL, vy = (o / BL) % BL, vzl = o / (BL * BL);
      const float *t0 = tile + (2 * vzl) * PL + (2 * vy) * H + 2 * vx;
      float S[4][4];
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int l = 0; l < 4; l++) S[j][l] = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; k++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
          for (int l = 0; l < 4; l++)
            S[j][l] = S[j][l] + W.w[16 * k + 4 * j + l] * t0[k * PL + j * H + l];
      float T[4];
#pragma unroll
      for (int l = 0; l < 4; l++) T[l] = ((S[0][l] + S[1][l]) + S[2][l]) + S[3][l];
      const float v = (T[0] + T[2]) + (T[1] + T[3]);
      const int64_t idx = (int64_t)pg * (BL * BL * BL) + ((vz0 + vzl) * BL + vy) * BL + vx;
      if constexpr (ST == MSBG_STORAGE_F32)
        ((float *)loPages)[idx] = v;
      else
        ((uint16_t *)loPages)[idx] = Stor<ST>::enc(v);
    }
    __syncthreads();
  }
}

LevelView make_view(msbg_grid *g, Channel &ch, int level) {
  LevelView v{};
  if (level >= g->nlevels || !ch.lv[level].defined) return v;
  ChanLevel &cl = ch.lv[level];
  v.sx = g->sx >> level;
  v.sy = g->sy >> level;
  v.sz = g->sz >> level;
  v.bsx = cl.bsx;
  v.lg = g->lg0 - level;
  v.nbx = g->nbx;
  v.nby = g->nby;
  v.nbz = g->nbz;
  v.page = cl.d_page.p;
  v.pages = cl.d_pages.p;
  v.storage = ch.storage;
  v.empty_raw = cl.empty_raw;
  v.full_raw = cl.full_raw;
  return v;
}

// cell-centred 4x4x4 kernel weights (msbg.cpp:298-343): per axis
// 1 - 0.5|1 - (d + 0.5)| for d = -1..2, product in double stored as float,
// then divided by the double sum.
GhostWeights ghost_weights() {
  GhostWeights W;
  double wd[64], wsum = 0;
  for (int dz = -1, ik = 0; dz <= 2; dz++)
    for (int dy = -1; dy <= 2; dy++)
      for (int dx = -1; dx <= 2; dx++, ik++) {
        double v = (1 - 0.5 * std::fabs(1.0 - (dx + 0.5))) * (1 - 0.5 * std::fabs(1.0 - (dy + 0.5))) *
                   (1 - 0.5 * std::fabs(1.0 - (dz + 0.5)));
        wd[ik] = v;
        wsum += v;
      }
  for (int i = 0; i < 64; i++) W.w[i] = (float)((double)(float)wd[i] / wsum);
  return W;
}

// Ghost pages of a coarse-level sweep: page gp (block bids[i]) of level L
// gets, in the voxels a swept neighbour reads, the value getOutOfBlockValue
// gives an in-domain voxel of a NULL block with OPT_BC_COARSE_LEVEL
// [| OPT_INTERPOLATE] (halo.h:288-352): interpolateGhost of level Lo at
// ((x+.5)/2, ...) or its voxel (x/2, y/2, z/2). Stored raw in the channel's
// storage, so the sweep reads back what fillHaloBlock_ would have produced.
//
// A swept block at +d (d in {-1,0,1}^3) reads the ghost's one-voxel layer
// facing it: voxels with v_k = bsx-1 where d_k = 1 and v_k = 0 where
// d_k = -1 (rows and planes, the side-car faces and edges built from them,
// and out-of-domain cells clipped next to the swept block all lie there).
// masks[i] has bit (d_z+1)*9 + (d_y+1)*3 + d_x+1 set for each swept
// neighbour; only those layers are filled (the whole shell when masks is
// null). Threads walk the shell: the planes z = 0 and z = bsx-1, then the
// ring of every plane between them.
template <int ST>
__global__ void k_fill_ghost_pages(LevelView L, LevelView Lo, unsigned opt,
                                   const int32_t *__restrict__ bids,
                                   const uint32_t *__restrict__ masks, int64_t firstPage,
                                   uint8_t *__restrict__ pages) {
  using R = typename Stor<ST>::raw;
  const int B = L.bsx, m = B - 1, per = 4 * B - 4;
  const int nshell = 2 * B * B + (B - 2) * per;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nshell) return;
  int vx, vy, vz;
  if (s < 2 * B * B) {
    vz = s < B * B ? 0 : m;
    const int r = s - (vz ? B * B : 0);
    vy = r / B;
    vx = r - vy * B;
  } else {
    const int t = s - 2 * B * B, q = t % per;
    vz = 1 + t / per;
    if (q < 2 * B) {
      vy = q < B ? 0 : m;
      vx = q < B ? q : q - B;
    } else {
      vy = 1 + (q - 2 * B) / 2;
      vx = ((q - 2 * B) & 1) ? m : 0;
    }
  }
  const int64_t i = blockIdx.y;
  if (masks) {
    const uint32_t mk = masks[i];
    const int cx = vx == 0 ? -1 : (vx == m ? 1 : 0), cy = vy == 0 ? -1 : (vy == m ? 1 : 0),
              cz = vz == 0 ? -1 : (vz == m ? 1 : 0);
    bool need = false;
    for (int k = 1; k < 8 && !need; k++) {  // the directions whose layer holds the voxel
      if (((k & 1) && !cx) || ((k & 2) && !cy) || ((k & 4) && !cz)) continue;
      const int dx = (k & 1) ? cx : 0, dy = (k & 2) ? cy : 0, dz = (k & 4) ? cz : 0;
      need = (mk >> ((dz + 1) * 9 + (dy + 1) * 3 + dx + 1)) & 1u;
    }
    if (!need) return;
  }
  const int v = (((vz << L.lg) + vy) << L.lg) + vx;
  const int bid = bids[i];
  const int bx = bid % L.nbx, by = (bid / L.nbx) % L.nby, bz = bid / (L.nbx * L.nby);
  const int x = bx * L.bsx + vx, y = by * L.bsx + vy, z = bz * L.bsx + vz;
  double raw;
  if (opt & MSBG_OPT_INTERPOLATE) {
    raw = interpolate_ghost_raw(Lo, ((float)x + 0.5f) / 2, ((float)y + 0.5f) / 2,
                                ((float)z + 0.5f) / 2);
  } else {
    int xl = x / 2, yl = y / 2, zl = z / 2;
    clip3(Lo, xl, yl, zl);
    raw = get_value_raw(Lo, xl, yl, zl);
  }
  R out;
  if constexpr (ST == MSBG_STORAGE_U16) out = (R)(uint16_t)raw;
  else out = Stor<ST>::enc((float)raw);
  ((R *)pages)[(firstPage + i) * ((int64_t)B * B * B) + v] = out;
}

// The INTERPOLATE fill with the coarse neighbourhood staged in LDS: one
// workgroup per ghost page. Its voxels interpolate level Lo at
// ((x+.5)/2, ...), whose corners lie in the page's coarse voxels grown by
// one, a (B/2+2)^3 region. Each region voxel is staged once as the value
// interpolateGhost reads there (out of range: the clipped voxel's value;
// else getValue's), with a flag for voxels of NULL blocks, whose corners
// take the value at the sample point's voxel. The arithmetic is
// interpolate_ghost_with's, so the pages equal k_fill_ghost_pages'.
template <int ST, int B>
__global__ __launch_bounds__(256) void k_fill_ghost_interp(LevelView L, LevelView Lo,
                                                           const int32_t *__restrict__ bids,
                                                           const uint32_t *__restrict__ masks,
                                                           int64_t firstPage,
                                                           uint8_t *__restrict__ pages,
                                                           uint8_t *__restrict__ side,
                                                           int64_t nTotal) {
  using R = typename Stor<ST>::raw;
  constexpr int HB = B / 2 + 2, HV = HB * HB * HB, M = B - 1;
  constexpr int PS = B * B, FS = 2 * PS, ES = 4 * B;
  __shared__ float sval[HV];
  __shared__ uint8_t snull[HV];
  const int64_t i = blockIdx.x;
  const int bid = bids[i];
  const int bx = bid % L.nbx, by = (bid / L.nbx) % L.nby, bz = bid / (L.nbx * L.nby);
  const int cx0 = bx * (B / 2) - 1, cy0 = by * (B / 2) - 1, cz0 = bz * (B / 2) - 1;
  for (int e = threadIdx.x; e < HV; e += blockDim.x) {
    const int hx = e % HB, hy = (e / HB) % HB, hz = e / (HB * HB);
    int X = cx0 + hx, Y = cy0 + hy, Z = cz0 + hz;
    bool isNull = false;
    if (in_range(Lo, X, Y, Z))
      isNull = code_at(Lo, X, Y, Z) == MSBG_BLOCK_NULL;
    else
      clip3(Lo, X, Y, Z);
    sval[e] = (float)get_value_raw(Lo, X, Y, Z);
    snull[e] = isNull ? 1 : 0;
  }
  __syncthreads();
  const uint32_t mk = masks ? masks[i] : 0xffffffffu;
  // The needed voxels are the union of the layers facing the swept
  // neighbours: for direction d, the voxels with v = 0 on the axes where
  // d < 0 and v = M where d > 0, free on the others. The workgroup walks the
  // set directions in turn (a uniform loop) and each layer's voxels with its
  // threads; a voxel shared by two layers is written twice, with one value.
  for (int bit = 0; bit < 27; bit++) {
    if (bit == 13 || !((mk >> bit) & 1u)) continue;
    const int dx = bit % 3 - 1, dy = (bit / 3) % 3 - 1, dz = bit / 9 - 1;
    const int size = (dx ? 1 : B) * (dy ? 1 : B) * (dz ? 1 : B);
    for (int t = threadIdx.x; t < size; t += blockDim.x) {
      int r = t, vx, vy, vz;
      if (dx) vx = dx < 0 ? 0 : M;
      else { vx = r % B; r /= B; }
      if (dy) vy = dy < 0 ? 0 : M;
      else { vy = r % B; r /= B; }
      if (dz) vz = dz < 0 ? 0 : M;
      else vz = r;
      const int x = bx * B + vx, y = by * B + vy, z = bz * B + vz;
      const float px = ((float)x + 0.5f) / 2, py = ((float)y + 0.5f) / 2, pz = ((float)z + 0.5f) / 2;
      auto corner = [&](int X, int Y, int Z) -> double {
        const int e = ((Z - cz0) * HB + (Y - cy0)) * HB + (X - cx0);
        if (!snull[e]) return sval[e];
        // the sample point's voxel lies in the domain and in the region
        return sval[(((int)floorf(pz) - cz0) * HB + ((int)floorf(py) - cy0)) * HB +
                    ((int)floorf(px) - cx0)];
      };
      const double raw = interpolate_ghost_with(Lo.storage, px, py, pz, corner);
      R out;
      if constexpr (ST == MSBG_STORAGE_U16) out = (R)(uint16_t)raw;
      else out = Stor<ST>::enc((float)raw);
      ((R *)pages)[(firstPage + i) * ((int64_t)B * B * B) + (((vz * B) + vy) * B + vx)] = out;
      // side-cars of the pool's nTotal pages (when given): the voxel's entry
      // in its page's x-faces [2][z][y] and z-edges [4][z], as the refresh
      // kernel lays them out
      if (side && (vx == 0 || vx == M)) {
        R *sd = (R *)side;
        const int64_t p = firstPage + i;
        sd[p * FS + (vx ? PS : 0) + vz * B + vy] = out;
        if (vy == 0 || vy == M)
          sd[nTotal * FS + p * ES + ((vx == M) + 2 * (vy == M)) * B + vz] = out;
      }
    }
  }
}

}  // namespace

int launch_halo_fill(msbg_grid *g, Channel &ch, int level, int bid, int dmax, unsigned opt,
                     float *d_out) {
  LevelView L = make_view(g, ch, level);
  LevelView Lo = make_view(g, ch, level + 1);
  int hsx = L.bsx + 2 * dmax;
  int n = hsx * hsx * hsx;
  hipLaunchKernelGGL(k_halo_fill, dim3((n + 255) / 256), dim3(256), 0, g->stream, L, Lo, bid, dmax,
                     opt, d_out);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_fill_ghost_pages(msbg_grid *g, Channel &ch, int level, unsigned opt,
                            const int32_t *d_bids, const uint32_t *d_mask, int64_t n,
                            uint8_t *d_pages, int64_t firstPage, uint8_t *d_side,
                            int64_t nTotal, bool *sideDone) {
  *sideDone = false;
  LevelView L = make_view(g, ch, level);
  LevelView Lo = make_view(g, ch, level + 1);
  if ((opt & MSBG_OPT_INTERPOLATE) && n > 0 && (L.bsx == 8 || L.bsx == 16 || L.bsx == 32) &&
      ch.storage != MSBG_STORAGE_F64) {
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(256), 0, g->stream, L, Lo, d_bids, d_mask,
                         firstPage, d_pages, d_side, nTotal);
    };
    auto by_size = [&](auto st) {
      constexpr int ST = decltype(st)::value;
      if (L.bsx == 8) go(k_fill_ghost_interp<ST, 8>);
      else if (L.bsx == 16) go(k_fill_ghost_interp<ST, 16>);
      else go(k_fill_ghost_interp<ST, 32>);
    };
    switch (ch.storage) {
      case MSBG_STORAGE_U16: by_size(std::integral_constant<int, MSBG_STORAGE_U16>{}); break;
      case MSBG_STORAGE_F32: by_size(std::integral_constant<int, MSBG_STORAGE_F32>{}); break;
      case MSBG_STORAGE_F16: by_size(std::integral_constant<int, MSBG_STORAGE_F16>{}); break;
      default: by_size(std::integral_constant<int, MSBG_STORAGE_BF16>{}); break;
    }
    MSBG_HIP(hipGetLastError());
    *sideDone = d_side != nullptr;
    return 0;
  }
  const int B = L.bsx, nshell = 2 * B * B + (B - 2) * (4 * B - 4);
  for (int64_t off = 0; off < n; off += 65535) {
    const unsigned cnt = (unsigned)(n - off < 65535 ? n - off : 65535);
    const dim3 grid((unsigned)((nshell + 255) / 256), cnt);
    const uint32_t *mk = d_mask ? d_mask + off : nullptr;
    switch (ch.storage) {
      case MSBG_STORAGE_U16:
        hipLaunchKernelGGL(k_fill_ghost_pages<MSBG_STORAGE_U16>, grid, dim3(256), 0, g->stream, L,
                           Lo, opt, d_bids + off, mk, firstPage + off, d_pages);
        break;
      case MSBG_STORAGE_F32:
        hipLaunchKernelGGL(k_fill_ghost_pages<MSBG_STORAGE_F32>, grid, dim3(256), 0, g->stream, L,
                           Lo, opt, d_bids + off, mk, firstPage + off, d_pages);
        break;
      case MSBG_STORAGE_F16:
        hipLaunchKernelGGL(k_fill_ghost_pages<MSBG_STORAGE_F16>, grid, dim3(256), 0, g->stream, L,
                           Lo, opt, d_bids + off, mk, firstPage + off, d_pages);
        break;
      case MSBG_STORAGE_BF16:
        hipLaunchKernelGGL(k_fill_ghost_pages<MSBG_STORAGE_BF16>, grid, dim3(256), 0, g->stream,
                           L, Lo, opt, d_bids + off, mk, firstPage + off, d_pages);
        break;
      default: set_error("ghost pages: unsupported storage %d", ch.storage); return MSBG_EUNSUPP;
    }
    MSBG_HIP(hipGetLastError());
  }
  return 0;
}

int launch_ghost(msbg_grid *g, Channel &ch, int level, const int32_t *d_bids, int64_t n) {
  LevelView L = make_view(g, ch, level);
  LevelView Lo = make_view(g, ch, level + 1);
  GhostWeights W = ghost_weights();
  if (n <= 0) return 0;
  if (L.bsx >= 4 && L.bsx <= 32) {
    uint8_t *lp = ch.lv[level + 1].d_pages.p;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(kGhostThreads), 0, g->stream, L, Lo, W, d_bids,
                         lp);
    };
    auto by_size = [&](auto st) {
      constexpr int ST = decltype(st)::value;
      switch (L.bsx) {
        case 4: go(k_ghost_tile<ST, 4>); break;
        case 8: go(k_ghost_tile<ST, 8>); break;
        case 16: go(k_ghost_tile<ST, 16>); break;
        default: go(k_ghost_tile<ST, 32>); break;
      }
    };
    switch (ch.storage) {
      case MSBG_STORAGE_F32: by_size(std::integral_constant<int, MSBG_STORAGE_F32>{}); break;
      case MSBG_STORAGE_F16: by_size(std::integral_constant<int, MSBG_STORAGE_F16>{}); break;
      case MSBG_STORAGE_BF16: by_size(std::integral_constant<int, MSBG_STORAGE_BF16>{}); break;
      default: set_error("ghost: unsupported storage %d", ch.storage); return MSBG_EUNSUPP;
    }
    MSBG_HIP(hipGetLastError());
    return 0;
  }
  // 2^3 and 64^3 blocks: the per-tap kernel
  int bl = L.bsx / 2;
  int nv = bl * bl * bl;
  for (int64_t off = 0; off < n; off += 65535) {
    unsigned cnt = (unsigned)(n - off < 65535 ? n - off : 65535);
    dim3 grid((nv + 255) / 256, cnt);
    switch (ch.storage) {
      case MSBG_STORAGE_F32:
        hipLaunchKernelGGL(k_ghost<MSBG_STORAGE_F32>, grid, dim3(256), 0, g->stream, L, Lo, W,
                           d_bids + off, ch.lv[level + 1].d_pages.p);
        break;
      case MSBG_STORAGE_F16:
        hipLaunchKernelGGL(k_ghost<MSBG_STORAGE_F16>, grid, dim3(256), 0, g->stream, L, Lo, W,
                           d_bids + off, ch.lv[level + 1].d_pages.p);
        break;
      case MSBG_STORAGE_BF16:
        hipLaunchKernelGGL(k_ghost<MSBG_STORAGE_BF16>, grid, dim3(256), 0, g->stream, L, Lo, W,
                           d_bids + off, ch.lv[level + 1].d_pages.p);
        break;
      default: set_error("ghost: unsupported storage %d", ch.storage); return MSBG_EUNSUPP;
    }
    MSBG_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace msbg
