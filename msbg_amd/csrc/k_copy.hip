// Plain streaming copy between two device buffers: the bandwidth reference
// (and the profiler byte-counter calibration) for the stencil kernels.

#include <hip/hip_runtime.h>

#include "internal.h"

namespace msbg {
namespace {

// one 16-byte element per thread: each workgroup streams one contiguous
// 4 KiB chunk, which measured faster than grid-stride or multi-element
// variants on MI355X
__global__ void k_copy16(const uint4 *__restrict__ a, uint4 *__restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

}  // namespace
}  // namespace msbg

using namespace msbg;

extern "C" int msbg_copy_bandwidth(int64_t bytes, int reps, double *gbps) {
  MSBG_CHECK(bytes >= 16 && reps >= 1 && gbps, MSBG_EINVAL, "msbg_copy_bandwidth: bad args");
  int64_t n = bytes / 16;
  DevBuf<uint4> a, b;
  MSBG_TRY(a.alloc((size_t)n));
  MSBG_TRY(b.alloc((size_t)n));
  MSBG_HIP(hipMemset(a.p, 1, (size_t)n * 16));
  hipEvent_t e0, e1;
  MSBG_HIP(hipEventCreate(&e0));
  MSBG_HIP(hipEventCreate(&e1));
  MSBG_CHECK(n / 256 < 0x7fffffffLL, MSBG_EINVAL, "msbg_copy_bandwidth: too many bytes");
  unsigned grid = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, nullptr, a.p, b.p, n);  // warm-up
  MSBG_HIP(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, nullptr, a.p, b.p, n);
  MSBG_HIP(hipEventRecord(e1, nullptr));
  MSBG_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MSBG_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbps = 2.0 * (double)n * 16 * reps / (ms * 1e-3) / 1e9;
  return 0;
}
