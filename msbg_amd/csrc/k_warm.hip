// Block windows in and predictions out of the warm-start network.
//
// The window of a 32^3 block is its 36^3 neighbourhood (a 2-voxel halo)
// under the halo rule of HaloBlockSet::fillHaloBlock_ with OPT_BC_NEUMANN
// (halo.cpp:267-479, halo.h:267-287): a voxel inside the domain whose block
// is non-NULL reads that block (const EMPTY/FULL blocks read their value);
// NULL blocks and the outside of the domain read the block's own voxel at
// the clamped position.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cnn_kernels.h"
#include "internal.h"
#include "storage.h"

namespace msbg {
namespace cnn {

namespace {

constexpr int B = 32, W = 36, H = 2;

template <int ST>
__global__ void k_gather36(const int32_t *__restrict__ page_map, const uint8_t *__restrict__ pages,
                           float empty_v, float full_v, int nbx, int nby, int nbz,
                           const int32_t *__restrict__ bids, int n, uint16_t *__restrict__ out) {
  using S = Stor<ST>;
  using R = typename S::raw;
  constexpr int WV = W * W * W;
  constexpr int64_t PE = (int64_t)B * B * B;
  // window k = blockIdx.y, voxel w of it (no 64-bit division per voxel)
  const int w = (int)blockIdx.x * (int)blockDim.x + (int)threadIdx.x;
  const int k = (int)blockIdx.y;
  if (w >= WV || k >= n) return;
  const int64_t i = (int64_t)k * WV + w;
  const int wx = w % W, wy = (w / W) % W, wz = w / (W * W);
  const int bid = bids[k];
  const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
  int gx = bx * B + wx - H, gy = by * B + wy - H, gz = bz * B + wz - H;
  const R *P = (const R *)pages;
  float v;
  int code = MSBG_BLOCK_NULL;
  if (gx >= 0 && gy >= 0 && gz >= 0 && gx < nbx * B && gy < nby * B && gz < nbz * B)
    code = page_map[((int64_t)(gz / B) * nby + gy / B) * nbx + gx / B];
  if (code == MSBG_BLOCK_EMPTY) {
    v = empty_v;
  } else if (code == MSBG_BLOCK_FULL) {
    v = full_v;
  } else {
    if (code < 0) {  // Neumann: the own block's voxel at the clamped position
      code = page_map[bid];
      gx = gx < bx * B ? bx * B : (gx > bx * B + B - 1 ? bx * B + B - 1 : gx);
      gy = gy < by * B ? by * B : (gy > by * B + B - 1 ? by * B + B - 1 : gy);
      gz = gz < bz * B ? bz * B : (gz > bz * B + B - 1 ? bz * B + B - 1 : gz);
    }
    v = (float)S::dec(P[(int64_t)code * PE + ((gz & (B - 1)) * B + (gy & (B - 1))) * B +
                         (gx & (B - 1))]);
  }
  // round to nearest even in hardware (v_cvt_pk_bf16_f32): the same bits as
  // bf16_bits_from_float for every non-NaN value with FP32 denormals kept
  out[i] = __builtin_bit_cast(uint16_t, (__bf16)v);
}

template <int ST>
__global__ void k_scatter32(const int32_t *__restrict__ page_map, uint8_t *__restrict__ pages,
                            const int32_t *__restrict__ bids, int n, const float *__restrict__ y) {
#pragma clang fp contract(off)
  using S = Stor<ST>;
  using R = typename S::raw;
  constexpr int64_t PE = (int64_t)B * B * B;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * PE) return;
  const int64_t k = i / PE, v = i % PE;
  R *dst = (R *)pages + (int64_t)page_map[bids[k]] * PE + v;
  float f = (float)S::dec(*dst) + y[i];
  f = f > 0.f ? f : 0.f;  // the sweep's [0,1] constraint (msbg4.cpp:627-630)
  f = f < 1.f ? f : 1.f;
  *dst = S::enc(f);
}

template <int ST>
__global__ void k_extract32(const int32_t *__restrict__ page_map, const uint8_t *__restrict__ pages,
                            const int32_t *__restrict__ bids, int n, float *__restrict__ y) {
  using S = Stor<ST>;
  using R = typename S::raw;
  constexpr int64_t PE = (int64_t)B * B * B;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * PE) return;
  const int64_t k = i / PE, v = i % PE;
  y[i] = (float)S::dec(((const R *)pages)[(int64_t)page_map[bids[k]] * PE + v]);
}

}  // namespace

int launch_extract32(hipStream_t s, int storage, const int32_t *page_map, const uint8_t *pages,
                     const int32_t *bids, int n, float *y) {
  const int64_t total = (int64_t)n * B * B * B;
  const dim3 grid((unsigned)((total + 255) / 256)), blk(256);
  switch (storage) {
    case MSBG_STORAGE_U16:
      hipLaunchKernelGGL(k_extract32<MSBG_STORAGE_U16>, grid, blk, 0, s, page_map, pages, bids, n, y);
      break;
    case MSBG_STORAGE_F32:
      hipLaunchKernelGGL(k_extract32<MSBG_STORAGE_F32>, grid, blk, 0, s, page_map, pages, bids, n, y);
      break;
    case MSBG_STORAGE_F16:
      hipLaunchKernelGGL(k_extract32<MSBG_STORAGE_F16>, grid, blk, 0, s, page_map, pages, bids, n, y);
      break;
    case MSBG_STORAGE_BF16:
      hipLaunchKernelGGL(k_extract32<MSBG_STORAGE_BF16>, grid, blk, 0, s, page_map, pages, bids, n, y);
      break;
    default:
      set_error("extract: storage %d not supported", storage);
      return MSBG_EUNSUPP;
  }
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_gather36(hipStream_t s, int storage, const int32_t *page_map, const uint8_t *pages,
                    double empty_raw, double full_raw, int nbx, int nby, int nbz,
                    const int32_t *bids, int n, uint16_t *out) {
  MSBG_CHECK(n >= 0, MSBG_EINVAL, "gather: %d windows", n);
  const float ev = raw_to_float_host(storage, empty_raw), fv = raw_to_float_host(storage, full_raw);
  // one grid row per window, at most 65535 windows per launch
  for (int off = 0; off < n; off += 65535) {
    const int cnt = n - off < 65535 ? n - off : 65535;
    const dim3 grid((unsigned)((W * W * W + 255) / 256), (unsigned)cnt), blk(256);
    const int32_t *b = bids + off;
    uint16_t *o = out + (int64_t)off * W * W * W;
    switch (storage) {
      case MSBG_STORAGE_U16:
        hipLaunchKernelGGL(k_gather36<MSBG_STORAGE_U16>, grid, blk, 0, s, page_map, pages, ev, fv,
                           nbx, nby, nbz, b, cnt, o);
        break;
      case MSBG_STORAGE_F32:
        hipLaunchKernelGGL(k_gather36<MSBG_STORAGE_F32>, grid, blk, 0, s, page_map, pages, ev, fv,
                           nbx, nby, nbz, b, cnt, o);
        break;
      case MSBG_STORAGE_F16:
        hipLaunchKernelGGL(k_gather36<MSBG_STORAGE_F16>, grid, blk, 0, s, page_map, pages, ev, fv,
                           nbx, nby, nbz, b, cnt, o);
        break;
      case MSBG_STORAGE_BF16:
        hipLaunchKernelGGL(k_gather36<MSBG_STORAGE_BF16>, grid, blk, 0, s, page_map, pages, ev, fv,
                           nbx, nby, nbz, b, cnt, o);
        break;
      default:
        set_error("gather: storage %d not supported by the network input", storage);
        return MSBG_EUNSUPP;
    }
    MSBG_HIP(hipGetLastError());
  }
  return 0;
}

int launch_scatter32(hipStream_t s, int storage, const int32_t *page_map, uint8_t *pages,
                     const int32_t *bids, int n, const float *y) {
  const int64_t total = (int64_t)n * B * B * B;
  const dim3 grid((unsigned)((total + 255) / 256)), blk(256);
  switch (storage) {
    case MSBG_STORAGE_U16:
      hipLaunchKernelGGL(k_scatter32<MSBG_STORAGE_U16>, grid, blk, 0, s, page_map, pages, bids, n, y);
      break;
    case MSBG_STORAGE_F32:
      hipLaunchKernelGGL(k_scatter32<MSBG_STORAGE_F32>, grid, blk, 0, s, page_map, pages, bids, n, y);
      break;
    case MSBG_STORAGE_F16:
      hipLaunchKernelGGL(k_scatter32<MSBG_STORAGE_F16>, grid, blk, 0, s, page_map, pages, bids, n, y);
      break;
    case MSBG_STORAGE_BF16:
      hipLaunchKernelGGL(k_scatter32<MSBG_STORAGE_BF16>, grid, blk, 0, s, page_map, pages, bids, n, y);
      break;
    default:
      set_error("scatter: storage %d not supported by the network output", storage);
      return MSBG_EUNSUPP;
  }
  MSBG_HIP(hipGetLastError());
  return 0;
}

}  // namespace cnn
}  // namespace msbg
