// Particle splat and finalize kernels (msbg_demo.cpp:541-680).
//
// The reference splats colour by colour with one task per home block so
// that plain read-modify-writes never race (msbg_demo.cpp:541-636). Here one
// thread owns one particle and the min goes through atomics, which makes the
// colour order unnecessary: min is order-independent, and the block
// value-ness that decides the early abandon does not change while splatting.
//
// Reference paths are relative to the reference's src/ directory.

#include <hip/hip_runtime.h>

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

template <int ST>
__global__ void k_finalize(uint8_t *__restrict__ pages, int64_t n, float rParticle, float nbDist,
                           float distSqMax) {
#pragma clang fp contract(off)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    bool z;
    if constexpr (ST == MSBG_STORAGE_U16) {
      uint16_t *p = (uint16_t *)pages;
      float f = finalize_value(p[i], rParticle, nbDist, distSqMax, &z);
      p[i] = z ? (uint16_t)0 : (uint16_t)(int)__builtin_roundf(f * 65535.0f);
    } else if constexpr (ST == MSBG_STORAGE_F16 || ST == MSBG_STORAGE_BF16) {
      uint16_t *p = (uint16_t *)pages;
      float f = finalize_value(p[i], rParticle, nbDist, distSqMax, &z);
      p[i] = Stor<ST>::enc(z ? 0.0f : f);
    } else if constexpr (ST == MSBG_STORAGE_F32) {
      uint32_t *p = (uint32_t *)pages;
      float f = finalize_value(p[i], rParticle, nbDist, distSqMax, &z);
      ((float *)pages)[i] = z ? 0.0f : f;
    } else {
      uint64_t *p = (uint64_t *)pages;
      float f = finalize_value((uint32_t)p[i], rParticle, nbDist, distSqMax, &z);
      ((double *)pages)[i] = z ? 0.0 : (double)f;
    }
  }
}

inline unsigned grid_cap(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

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
  int64_t nb = (g->npoints + kThreads - 1) / kThreads;
  int esz = storage_bytes(ch.storage);
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
