// Persistent batched sweep: one launch per batch group runs every sweep and
// colour pass of the group's grids from a work queue (k_sweep_queue in
// sweep_core.h), each grid keeping the reference's colour order
// (msbg4.cpp:384-900) through per-grid completion counts instead of a launch
// boundary per pass.
//
// Reference paths are relative to the reference's src/ directory.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "storage.h"
#include "sweep_core.h"

namespace msbg {

namespace {

using namespace sweep;

// zero n words (the queue state, before each run; a kernel, so that a
// captured graph holds no memset node)
__global__ void k_queue_reset(unsigned *__restrict__ p, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = 0u;
}

int device_cus() {
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

template <int BSX, int ST, bool EXACT, int PX, int PY, bool C01>
int launch_q(hipStream_t s, const SweepGrid *d_grids, const BatchQueue &q, int nwg) {
  constexpr int NT = Tile<BSX, PX, PY>::NT;
  auto kern = k_sweep_queue<BSX, ST, EXACT, PX, PY, C01>;
  if (nwg <= 0) {
    int per = 0;
    MSBG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, NT, 0));
    nwg = (per > 0 ? per : 1) * device_cus();
  }
  const int64_t total = (int64_t)q.off[8] * q.sweeps;
  if (total <= 0) return 0;
  int widest = 0;  // the largest colour pass
  for (int c = 0; c < 8; c++) widest = std::max(widest, q.off[c + 1] - q.off[c]);
  nwg = std::max(1, std::min(nwg, widest));
  hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(NT), 0, s, d_grids, q);
  MSBG_HIP(hipGetLastError());
  return 0;
}

// one wave per workgroup: 4x4 columns at 32^3, 2x2 at 16^3, in both
// arithmetic modes (FP64 channels are swept one launch per pass)
template <int BSX, int ST>
int launch_q_st(hipStream_t s, bool exact, bool c01, const SweepGrid *d_grids,
                const BatchQueue &q, int nwg) {
  constexpr int P = BSX >= 32 ? 4 : 2;
  if constexpr (ST == MSBG_STORAGE_F64) {
    set_error("persistent sweep: FP64 channels run one launch per colour pass");
    return MSBG_EUNSUPP;
  } else {
    if (exact) return launch_q<BSX, ST, true, P, P, false>(s, d_grids, q, nwg);
    return c01 ? launch_q<BSX, ST, false, P, P, true>(s, d_grids, q, nwg)
               : launch_q<BSX, ST, false, P, P, false>(s, d_grids, q, nwg);
  }
}

template <int BSX>
int launch_q_bsx(hipStream_t s, int storage, bool exact, bool c01, const SweepGrid *d_grids,
                 const BatchQueue &q, int nwg) {
  switch (storage) {
    case MSBG_STORAGE_U16: return launch_q_st<BSX, MSBG_STORAGE_U16>(s, exact, c01, d_grids, q, nwg);
    case MSBG_STORAGE_F32: return launch_q_st<BSX, MSBG_STORAGE_F32>(s, exact, c01, d_grids, q, nwg);
    case MSBG_STORAGE_F16: return launch_q_st<BSX, MSBG_STORAGE_F16>(s, exact, c01, d_grids, q, nwg);
    case MSBG_STORAGE_BF16: return launch_q_st<BSX, MSBG_STORAGE_BF16>(s, exact, c01, d_grids, q, nwg);
    case MSBG_STORAGE_F64: return launch_q_st<BSX, MSBG_STORAGE_F64>(s, exact, c01, d_grids, q, nwg);
    case MSBG_STORAGE_U8: return launch_q_st<BSX, MSBG_STORAGE_U8>(s, exact, c01, d_grids, q, nwg);
  }
  set_error("persistent sweep: bad storage %d", storage);
  return MSBG_EINVAL;
}

}  // namespace

int launch_sweep_queue(hipStream_t s, int bsx, int storage, bool exact, bool c01,
                       const SweepGrid *d_grids, const BatchQueue &q, int nwg) {
  switch (bsx) {
    case 16: return launch_q_bsx<16>(s, storage, exact, c01, d_grids, q, nwg);
    case 32: return launch_q_bsx<32>(s, storage, exact, c01, d_grids, q, nwg);
  }
  set_error("persistent sweep: block size %d (16 and 32 are supported)", bsx);
  return MSBG_EUNSUPP;
}

uint64_t wall_ticks(uint64_t us) {
  static int khz = 0;  // s_memrealtime rate
  if (!khz) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        khz <= 0)
      khz = 100000;
  }
  return us * (uint64_t)khz / 1000u;
}

int launch_queue_reset(hipStream_t s, unsigned *d_state, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_queue_reset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_state,
                     n);
  MSBG_HIP(hipGetLastError());
  return 0;
}

}  // namespace msbg
