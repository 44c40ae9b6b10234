// Channel-to-channel copy with storage conversion (SparseGrid::copyDataTo,
// sbg.h:550-557, extended to change the storage format) and the FP64
// difference statistics of two channels with the same block map.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.h"
#include "storage.h"

namespace msbg {
namespace {

__device__ inline double load_as_double(const uint8_t *p, int64_t i, int st) {
  switch (st) {
    case MSBG_STORAGE_U16: return (double)((const uint16_t *)p)[i] * (double)(1.0f / 65535.0f);
    case MSBG_STORAGE_F32: return ((const float *)p)[i];
    case MSBG_STORAGE_F16: return f16_bits_to_float(((const uint16_t *)p)[i]);
    case MSBG_STORAGE_BF16: return bf16_bits_to_float(((const uint16_t *)p)[i]);
    case MSBG_STORAGE_U8: return (double)((const uint8_t *)p)[i] * (double)(1.0f / 255.0f);
    default: return ((const double *)p)[i];
  }
}

// UNORM16 and UNORM8 targets round half to even (the reference's
// round_to_int stores, sbg.h:319-323, 343-351)
__device__ inline void store_from_double(uint8_t *p, int64_t i, int st, double v) {
  switch (st) {
    case MSBG_STORAGE_U16:
      ((uint16_t *)p)[i] = (uint16_t)(int)__builtin_rintf((float)v * 65535.0f);
      break;
    case MSBG_STORAGE_F32: ((float *)p)[i] = (float)v; break;
    case MSBG_STORAGE_F16: ((uint16_t *)p)[i] = f16_bits_from_float((float)v); break;
    case MSBG_STORAGE_BF16: ((uint16_t *)p)[i] = bf16_bits_from_float((float)v); break;
    case MSBG_STORAGE_U8: ((uint8_t *)p)[i] = (uint8_t)(int)__builtin_rintf((float)v * 255.0f); break;
    default: ((double *)p)[i] = v; break;
  }
}

__global__ void k_convert(const uint8_t *__restrict__ src, int sst, uint8_t *__restrict__ dst,
                          int dst_st, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) store_from_double(dst, i, dst_st, load_as_double(src, i, sst));
}

__global__ void k_diff(const uint8_t *__restrict__ a, int ast, const uint8_t *__restrict__ b,
                       int bst, int64_t n, double *__restrict__ pmax, double *__restrict__ psum) {
  __shared__ double smax[256], ssum[256];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double mx = 0, sm = 0;
  for (; i < n; i += stride) {
    double d = load_as_double(a, i, ast) - load_as_double(b, i, bst);
    mx = fmax(mx, fabs(d));
    sm += d * d;
  }
  smax[threadIdx.x] = mx;
  ssum[threadIdx.x] = sm;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      ssum[threadIdx.x] += ssum[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmax[blockIdx.x] = smax[0];
    psum[blockIdx.x] = ssum[0];
  }
}

}  // namespace

int launch_convert(msbg_grid *g, const ChanLevel &src, int sst, ChanLevel &dst, int dst_st) {
  int64_t n = src.npages * (int64_t)src.bsx * src.bsx * src.bsx;
  if (!n) return 0;
  int64_t nb = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_convert, dim3((unsigned)nb), dim3(256), 0, g->stream, src.d_pages.p, sst,
                     dst.d_pages.p, dst_st, n);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_diff(msbg_grid *g, const ChanLevel &a, int ast, const ChanLevel &b, int bst,
                double *maxAbs, double *sumSq) {
  int64_t n = a.npages * (int64_t)a.bsx * a.bsx * a.bsx;
  *maxAbs = 0;
  *sumSq = 0;
  if (!n) return 0;
  const int nb = 4096;
  DevBuf<double> pm, ps;
  MSBG_TRY(pm.alloc(nb));
  MSBG_TRY(ps.alloc(nb));
  hipLaunchKernelGGL(k_diff, dim3(nb), dim3(256), 0, g->stream, a.d_pages.p, ast, b.d_pages.p, bst,
                     n, pm.p, ps.p);
  MSBG_HIP(hipGetLastError());
  std::vector<double> hm(nb), hs(nb);
  MSBG_HIP(hipMemcpyAsync(hm.data(), pm.p, sizeof(double) * nb, hipMemcpyDeviceToHost, g->stream));
  MSBG_HIP(hipMemcpyAsync(hs.data(), ps.p, sizeof(double) * nb, hipMemcpyDeviceToHost, g->stream));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  for (int i = 0; i < nb; i++) {
    *maxAbs = std::max(*maxAbs, hm[i]);
    *sumSq += hs[i];
  }
  return 0;
}

}  // namespace msbg
