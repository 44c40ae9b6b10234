// Block-table kernels: activation, level bands, regularization, block flags,
// halo index tables, page fill and the residual reduction.
//
// Reference paths are relative to the reference's src/ directory.

#include <hip/hip_runtime.h>

#include "internal.h"
#include "storage.h"

namespace msbg {

namespace {

constexpr int kThreads = 256;

inline unsigned grid_for(int64_t n, int threads = kThreads) {
  int64_t b = (n + threads - 1) / threads;
  return (unsigned)(b < 1 ? 1 : b);
}

// msbg_demo.cpp:368-382. One thread per particle; the box bounds use the
// reference's FP32 expressions without contraction.
__global__ void k_activate(const float *__restrict__ xyz, int64_t n, float scale, float rScanBsx,
                           int nbx, int nby, int nbz, uint32_t *__restrict__ count) {
#pragma clang fp contract(off)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo[3], hi[3];
  const int bmax[3] = {nbx - 1, nby - 1, nbz - 1};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float bp = xyz[3 * i + k] * scale;
    int a = (int)(bp - rScanBsx), b = (int)(bp + rScanBsx);  // truncate_to_int
    lo[k] = a > 0 ? a : 0;
    hi[k] = b < bmax[k] ? b : bmax[k];
  }
  for (int bz = lo[2]; bz <= hi[2]; bz++)
    for (int by = lo[1]; by <= hi[1]; by++)
      for (int bx = lo[0]; bx <= hi[0]; bx++)
        atomicAdd(&count[((int64_t)bz * nby + by) * nbx + bx], 1u);
}

// The same counts over the particles in cell order (ensure_binned):
// neighbouring lanes share most of their blocks. Each wave ballots every
// block of the union of its lanes' boxes; when the workgroup's union holds
// at most kAggBlocks blocks, the counts gather in LDS and the workgroup adds
// each block once. Wider unions add per wave (up to 64 blocks) or per lane.
constexpr int kAggBlocks = 512;

__device__ inline int wave_min(int v) {
#pragma unroll
  for (int o = 32; o; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ inline int wave_max(int v) {
#pragma unroll
  for (int o = 32; o; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

__global__ __launch_bounds__(256) void k_activate_binned(const float *__restrict__ xyz, int64_t n,
                                                         float scale, float rScanBsx, int nbx,
                                                         int nby, int nbz,
                                                         uint32_t *__restrict__ count) {
#pragma clang fp contract(off)
  __shared__ int glo[3], ghi[3];
  __shared__ uint32_t agg[kAggBlocks];
  if (threadIdx.x < 3) {
    glo[threadIdx.x] = 0x7fffffff;
    ghi[threadIdx.x] = -1;
  }
  for (int t = threadIdx.x; t < kAggBlocks; t += blockDim.x) agg[t] = 0;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
  if (i < n) {
    const int bmax[3] = {nbx - 1, nby - 1, nbz - 1};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      float bp = xyz[3 * i + k] * scale;
      int a = (int)(bp - rScanBsx), b = (int)(bp + rScanBsx);  // truncate_to_int
      lo[k] = a > 0 ? a : 0;
      hi[k] = b < bmax[k] ? b : bmax[k];
    }
  }
  const bool has = lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2];
  int ulo[3], uhi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ulo[k] = wave_min(has ? lo[k] : 0x7fffffff);
    uhi[k] = wave_max(has ? hi[k] : -1);
  }
  const int ex = uhi[0] - ulo[0] + 1, ey = uhi[1] - ulo[1] + 1, ez = uhi[2] - ulo[2] + 1;
  const bool waveHas = ex > 0 && ey > 0 && ez > 0;
  const int lane = threadIdx.x & 63;
  __syncthreads();
  if (waveHas && lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      atomicMin(&glo[k], ulo[k]);
      atomicMax(&ghi[k], uhi[k]);
    }
  }
  __syncthreads();
  const int gx = ghi[0] - glo[0] + 1, gy = ghi[1] - glo[1] + 1, gz = ghi[2] - glo[2] + 1;
  if (gx <= 0 || gy <= 0 || gz <= 0) return;  // no lane of the workgroup has a block
  if (gx * gy * gz <= kAggBlocks) {
    if (waveHas)
      for (int bz = ulo[2]; bz <= uhi[2]; bz++)
        for (int by = ulo[1]; by <= uhi[1]; by++)
          for (int bx = ulo[0]; bx <= uhi[0]; bx++) {
            const bool in = has && bx >= lo[0] && bx <= hi[0] && by >= lo[1] && by <= hi[1] &&
                            bz >= lo[2] && bz <= hi[2];
            const uint64_t m = __ballot(in);
            if (m && lane == __builtin_ctzll(m))
              atomicAdd(&agg[((bz - glo[2]) * gy + (by - glo[1])) * gx + (bx - glo[0])],
                        (uint32_t)__builtin_popcountll(m));
          }
    __syncthreads();
    for (int t = threadIdx.x; t < gx * gy * gz; t += blockDim.x) {
      if (!agg[t]) continue;
      const int bx = glo[0] + t % gx, by = glo[1] + (t / gx) % gy, bz = glo[2] + t / (gx * gy);
      atomicAdd(&count[((int64_t)bz * nby + by) * nbx + bx], agg[t]);
    }
    return;
  }
  if (!waveHas) return;
  if (ex * ey * ez <= 64) {
    for (int bz = ulo[2]; bz <= uhi[2]; bz++)
      for (int by = ulo[1]; by <= uhi[1]; by++)
        for (int bx = ulo[0]; bx <= uhi[0]; bx++) {
          const bool in = has && bx >= lo[0] && bx <= hi[0] && by >= lo[1] && by <= hi[1] &&
                          bz >= lo[2] && bz <= hi[2];
          const uint64_t m = __ballot(in);
          if (m && lane == __builtin_ctzll(m))
            atomicAdd(&count[((int64_t)bz * nby + by) * nbx + bx],
                      (uint32_t)__builtin_popcountll(m));
        }
    return;
  }
  if (!has) return;
  for (int bz = lo[2]; bz <= hi[2]; bz++)
    for (int by = lo[1]; by <= hi[1]; by++)
      for (int bx = lo[0]; bx <= hi[0]; bx++)
        atomicAdd(&count[((int64_t)bz * nby + by) * nbx + bx], 1u);
}

// level bands: the widest band initialises (count ? band : nLevels-1),
// narrower bands overwrite where their count is nonzero.
__global__ void k_level_band(const uint32_t *__restrict__ count, int64_t nb, int band, int init,
                             int lastLevel, int32_t *__restrict__ levels) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  if (init)
    levels[i] = count[i] ? band : lastLevel;
  else if (count[i])
    levels[i] = band;
}

// One pass ll of regularizeRefinementMap (msbg.cpp:1319-1388): a block of
// level >= ll whose finest in-grid 26-neighbour has level ll-1 gets level ll.
__global__ void k_regularize(const int32_t *__restrict__ src, int32_t *__restrict__ dst, int nbx,
                             int nby, int nbz, int nLevels, int ll) {
  int64_t bid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t nb = (int64_t)nbx * nby * nbz;
  if (bid >= nb) return;
  int bx = (int)(bid % nbx), by = (int)((bid / nbx) % nby), bz = (int)(bid / ((int64_t)nbx * nby));
  int level = (int8_t)src[bid];  // blockmapTmp2 is int8_t
  if (level >= ll) {
    int level2 = nLevels - 1;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if (!dx && !dy && !dz) continue;
          int x = bx + dx, y = by + dy, z = bz + dz;
          if (x < 0 || y < 0 || z < 0 || x >= nbx || y >= nby || z >= nbz) continue;
          int l2 = (int8_t)src[((int64_t)z * nby + y) * nbx + x];
          level2 = l2 < level2 ? l2 : level2;
        }
    if (level2 == ll - 1) level = ll;
  }
  dst[bid] = level;
}

// setRefinementMap per-block part at levelMg 0 (msbg.cpp:1649-1745).
__global__ void k_flags(const int32_t *__restrict__ levels, const uint32_t *__restrict__ inFlags,
                        int nbx, int nby, int nbz, int nLevels, uint16_t *__restrict__ outLevel,
                        uint16_t *__restrict__ outFlags) {
  int64_t bid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t nb = (int64_t)nbx * nby * nbz;
  if (bid >= nb) return;
  int bx = (int)(bid % nbx), by = (int)((bid / nbx) % nby), bz = (int)(bid / ((int64_t)nbx * nby));
  int level = levels[bid];  // MAX(blockLevels[bid], levelMg=0)
  unsigned fl = inFlags ? (uint16_t)inFlags[bid] : 0u;
  bool border = bx == 0 || bx == nbx - 1 || by == 0 || by == nby - 1 || bz == 0 || bz == nbz - 1;
  fl = border ? (fl | MSBG_BLK_DOM_BORDER) : (fl & ~MSBG_BLK_DOM_BORDER);  // FLAG_COPY
  int l2max = 0, l2min = nLevels - 1;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        if (!dx && !dy && !dz) continue;
        int x = bx + dx, y = by + dy, z = bz + dz;
        if (x < 0 || y < 0 || z < 0 || x >= nbx || y >= nby || z >= nbz) continue;
        int l2 = levels[((int64_t)z * nby + y) * nbx + x];
        l2max = l2 > l2max ? l2 : l2max;
        l2min = l2 < l2min ? l2 : l2min;
      }
  if (l2max > level) fl |= MSBG_BLK_FINE_COARSE;
  if (l2min < level) fl |= MSBG_BLK_COARSE_FINE;
  fl &= ~MSBG_BLK_UNIFORM_EMPTY;
  if (!(fl & MSBG_BLK_EXISTS) && !(fl & (MSBG_BLK_COARSE_FINE | MSBG_BLK_FINE_COARSE)))
    fl |= MSBG_BLK_UNIFORM_EMPTY;
  outLevel[bid] = (uint16_t)level;
  outFlags[bid] = (uint16_t)fl;
}

// Halo index table: 27 neighbour codes per value page.
__global__ void k_build_nbr(const int32_t *__restrict__ pageBid, int64_t npages,
                            const int32_t *__restrict__ pagemap, int nbx, int nby, int nbz,
                            int32_t *__restrict__ nbr) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npages * 27) return;
  int64_t p = t / 27;
  int k = (int)(t % 27);
  int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
  int64_t bid = pageBid[p];
  int bx = (int)(bid % nbx) + dx, by = (int)((bid / nbx) % nby) + dy,
      bz = (int)(bid / ((int64_t)nbx * nby)) + dz;
  int32_t c = MSBG_BLOCK_OUT;
  if (bx >= 0 && by >= 0 && bz >= 0 && bx < nbx && by < nby && bz < nbz)
    c = pagemap[((int64_t)bz * nby + by) * nbx + bx];
  nbr[t] = c;
}

template <typename T>
__global__ void k_fill(T *__restrict__ p, int64_t n, T v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Deterministic FP64 sum of n partials (fixed order for a fixed n).
__global__ void k_reduce_sum(const double *__restrict__ a, int64_t n, double *__restrict__ out) {
  __shared__ double s[1024];
  double acc = 0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += a[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

}  // namespace

int launch_activate(msbg_grid *g, float rScan, uint32_t *d_count) {
  if (g->npoints == 0) return 0;
  const float scale = 1.0f / (float)g->bsx0;  // scale2DestBlockGrid
  const float rScanBsx = rScan / (float)g->bsx0;
  if (ensure_binned(g) == 0) {
    hipLaunchKernelGGL(k_activate_binned, dim3((unsigned)((g->npoints + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, g->stream, g->bin_xyz.p, g->npoints, scale, rScanBsx,
                       g->nbx, g->nby, g->nbz, d_count);
  } else {  // too many particles to bin with 32-bit offsets
    hipLaunchKernelGGL(k_activate, dim3(grid_for(g->npoints)), dim3(kThreads), 0, g->stream,
                       g->d_xyz.p, g->npoints, scale, rScanBsx, g->nbx, g->nby, g->nbz, d_count);
  }
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_level_band(msbg_grid *g, const uint32_t *d_count, int band, int nbands,
                      int32_t *d_levels) {
  hipLaunchKernelGGL(k_level_band, dim3(grid_for(g->nblocks)), dim3(kThreads), 0, g->stream,
                     d_count, g->nblocks, band, band == nbands - 1 ? 1 : 0, g->nlevels - 1,
                     d_levels);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_regularize(msbg_grid *g, const int32_t *d_in, int32_t *d_out, int ll) {
  hipLaunchKernelGGL(k_regularize, dim3(grid_for(g->nblocks)), dim3(kThreads), 0, g->stream, d_in,
                     d_out, g->nbx, g->nby, g->nbz, g->nlevels, ll);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_flags(msbg_grid *g, const int32_t *d_levels, const uint32_t *d_in_flags) {
  hipLaunchKernelGGL(k_flags, dim3(grid_for(g->nblocks)), dim3(kThreads), 0, g->stream, d_levels,
                     d_in_flags, g->nbx, g->nby, g->nbz, g->nlevels, g->d_level.p, g->d_flags.p);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_build_nbr(msbg_grid *g, ChanLevel &cl, int level) {
  (void)level;
  if (!cl.npages) return 0;
  hipLaunchKernelGGL(k_build_nbr, dim3(grid_for(cl.npages * 27)), dim3(kThreads), 0, g->stream,
                     cl.d_page_bid.p, cl.npages, cl.d_page.p, g->nbx, g->nby, g->nbz, cl.d_nbr.p);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_fill_pages(msbg_grid *g, ChanLevel &cl, int storage, double fill_raw,
                      int64_t first_page, int64_t npages) {
  int64_t nv = (int64_t)cl.bsx * cl.bsx * cl.bsx;
  int64_t n = npages * nv;
  if (!n) return 0;
  unsigned grid = grid_for(n);
  if (grid > 65536u) grid = 65536u;
  uint8_t *base = cl.d_pages.p + (size_t)first_page * nv * storage_bytes(storage);
  switch (storage) {
    case MSBG_STORAGE_U16:
      hipLaunchKernelGGL(k_fill<uint16_t>, dim3(grid), dim3(kThreads), 0, g->stream,
                         (uint16_t *)base, n, (uint16_t)fill_raw);
      break;
    case MSBG_STORAGE_F16:
      hipLaunchKernelGGL(k_fill<uint16_t>, dim3(grid), dim3(kThreads), 0, g->stream,
                         (uint16_t *)base, n, f16_bits_from_float((float)fill_raw));
      break;
    case MSBG_STORAGE_BF16:
      hipLaunchKernelGGL(k_fill<uint16_t>, dim3(grid), dim3(kThreads), 0, g->stream,
                         (uint16_t *)base, n, bf16_bits_from_float((float)fill_raw));
      break;
    case MSBG_STORAGE_F32:
      hipLaunchKernelGGL(k_fill<float>, dim3(grid), dim3(kThreads), 0, g->stream, (float *)base, n,
                         (float)fill_raw);
      break;
    case MSBG_STORAGE_F64:
      hipLaunchKernelGGL(k_fill<double>, dim3(grid), dim3(kThreads), 0, g->stream, (double *)base,
                         n, fill_raw);
      break;
    case MSBG_STORAGE_U8:
      hipLaunchKernelGGL(k_fill<uint8_t>, dim3(grid), dim3(kThreads), 0, g->stream, (uint8_t *)base,
                         n, (uint8_t)fill_raw);
      break;
    default: set_error("fill_pages: bad storage %d", storage); return MSBG_EINVAL;
  }
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_reduce_residual(msbg_grid *g, const double *d_partial, int64_t n, double *d_out) {
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(1024), 0, g->stream, d_partial, n, d_out);
  MSBG_HIP(hipGetLastError());
  return 0;
}

}  // namespace msbg
