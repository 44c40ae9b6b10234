// Plain streaming copy between two device buffers: the bandwidth reference
// (and the profiler byte-counter calibration) for the stencil kernels.

#include <hip/hip_runtime.h>

#include "internal.h"

namespace msbg {
namespace {

// four independent 16-byte loads in flight per thread before the stores
__global__ void k_copy16(const uint4 *__restrict__ a, uint4 *__restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    uint4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
    b[i] = v0;
    b[i + stride] = v1;
    b[i + 2 * stride] = v2;
    b[i + 3 * stride] = v3;
  }
  for (; i < n; i += stride) b[i] = a[i];
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
  int dev = 0, cus = 0;
  MSBG_HIP(hipGetDevice(&dev));
  MSBG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  unsigned grid = (unsigned)(cus > 0 ? cus * 8 : 2048);
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
