// The other stencils and driver modes of applyChannelPdeFast on the sweep
// kernel: laplacian flow (1), diffusion (3), Perona-Malik (14, 15), Jacobi
// diffusion (17), and mean curvature (4) with hard constraints
// (msbg4.cpp:60-957). The kernel is the z-marching pass of k_sweep.hip
// (sweep_core.h) with the operator and the constraint chosen at compile
// time; the Jacobi mode writes into another channel with the same page
// numbering.
//
// Ops 1, 3 and 17 read their neighbours at fixed halo offsets 18 and 18*18
// in the reference (msbg4.cpp:688-693, 719-724, 795-800), the strides of a
// 16^3 block's halo, so they are built for 16^3 blocks only.

#include <hip/hip_runtime.h>

#include "internal.h"
#include "storage.h"
#include "sweep_core.h"

namespace msbg {

namespace {

using namespace sweep;

template <int BSX, int ST, int OP, bool HARD>
int launch_op(hipStream_t s, const SweepGrid *d_grids, const int2 *d_items, int64_t n,
              double *d_partial, int iter) {
  constexpr int PX = BSX >= 32 ? 4 : 2, PY = 2;
  constexpr int NT = Tile<BSX, PX, PY>::NT;
  for (int64_t off = 0; off < n; off += 0x7fffffff) {
    const unsigned cnt = (unsigned)(n - off < 0x7fffffff ? n - off : 0x7fffffff);
    hipLaunchKernelGGL((k_sweep<BSX, ST, true, PX, PY, OP, HARD>), dim3(cnt), dim3(NT), 0, s,
                       d_grids, d_items + off, d_partial ? d_partial + off : nullptr, iter);
    MSBG_HIP(hipGetLastError());
  }
  return 0;
}

template <int BSX, int ST, bool HARD>
int launch_ops(hipStream_t s, int op, const SweepGrid *d_grids, const int2 *d_items, int64_t n,
               double *d_partial, int iter) {
  if (op == OP_MCF) return launch_op<BSX, ST, OP_MCF, HARD>(s, d_grids, d_items, n, d_partial, iter);
  if constexpr (ST != MSBG_STORAGE_F64) {
    if (op == OP_PM) return launch_op<BSX, ST, OP_PM, HARD>(s, d_grids, d_items, n, d_partial, iter);
    if (op == OP_PMJ) return launch_op<BSX, ST, OP_PMJ, HARD>(s, d_grids, d_items, n, d_partial, iter);
    if constexpr (BSX == 16) {
      if (op == OP_LAPL) return launch_op<BSX, ST, OP_LAPL, HARD>(s, d_grids, d_items, n, d_partial, iter);
      if (op == OP_DIFF) return launch_op<BSX, ST, OP_DIFF, HARD>(s, d_grids, d_items, n, d_partial, iter);
      if (op == OP_DIFFJ) return launch_op<BSX, ST, OP_DIFFJ, HARD>(s, d_grids, d_items, n, d_partial, iter);
    }
  }
  set_error("applyChannelPdeFast: laplTyp %d with block size %d / storage %d is not supported", op,
            BSX, ST);
  return MSBG_EUNSUPP;
}

template <int BSX, int ST>
int launch_st(hipStream_t s, int op, bool hard, const SweepGrid *d_grids, const int2 *d_items,
              int64_t n, double *d_partial, int iter) {
  return hard ? launch_ops<BSX, ST, true>(s, op, d_grids, d_items, n, d_partial, iter)
              : launch_ops<BSX, ST, false>(s, op, d_grids, d_items, n, d_partial, iter);
}

template <int BSX>
int launch_bsx(hipStream_t s, int storage, int op, bool hard, const SweepGrid *d_grids,
               const int2 *d_items, int64_t n, double *d_partial, int iter) {
  switch (storage) {
    case MSBG_STORAGE_U16: return launch_st<BSX, MSBG_STORAGE_U16>(s, op, hard, d_grids, d_items, n, d_partial, iter);
    case MSBG_STORAGE_F32: return launch_st<BSX, MSBG_STORAGE_F32>(s, op, hard, d_grids, d_items, n, d_partial, iter);
    case MSBG_STORAGE_F16: return launch_st<BSX, MSBG_STORAGE_F16>(s, op, hard, d_grids, d_items, n, d_partial, iter);
    case MSBG_STORAGE_BF16: return launch_st<BSX, MSBG_STORAGE_BF16>(s, op, hard, d_grids, d_items, n, d_partial, iter);
    case MSBG_STORAGE_F64: return launch_st<BSX, MSBG_STORAGE_F64>(s, op, hard, d_grids, d_items, n, d_partial, iter);
  }
  set_error("applyChannelPdeFast: bad storage %d", storage);
  return MSBG_EINVAL;
}

// The Jacobi mode's last sweep of an odd numIter copies the swept blocks
// (msbg4.cpp:481-487): pages and their side-cars.
template <typename R>
__global__ void k_copy_items(const SweepGrid *__restrict__ g, const int2 *__restrict__ items,
                             int64_t pe, int64_t fs, int64_t es) {
  const int64_t page = items[blockIdx.x].y;
  const R *sp = (const R *)g->pages + page * pe;
  R *dp = (R *)g->dst_pages + page * pe;
  for (int64_t i = threadIdx.x; i < pe; i += blockDim.x) dp[i] = sp[i];
  const R *sf = (const R *)g->side + page * fs;
  R *df = (R *)g->dst_side + page * fs;
  for (int64_t i = threadIdx.x; i < fs; i += blockDim.x) df[i] = sf[i];
  const R *se = (const R *)g->side + g->npages * fs + page * es;
  R *de = (R *)g->dst_side + g->npages * fs + page * es;
  for (int64_t i = threadIdx.x; i < es; i += blockDim.x) de[i] = se[i];
}

}  // namespace

int launch_pde_items(hipStream_t s, int bsx, int storage, int op, bool hard,
                     const SweepGrid *d_grids, const int2 *d_items, int64_t n, double *d_partial,
                     int iter) {
  switch (bsx) {
    case 8: return launch_bsx<8>(s, storage, op, hard, d_grids, d_items, n, d_partial, iter);
    case 16: return launch_bsx<16>(s, storage, op, hard, d_grids, d_items, n, d_partial, iter);
    case 32: return launch_bsx<32>(s, storage, op, hard, d_grids, d_items, n, d_partial, iter);
  }
  set_error("applyChannelPdeFast: block size %d not supported (8, 16, 32)", bsx);
  return MSBG_EUNSUPP;
}

int launch_copy_items(hipStream_t s, int bsx, int storage, const SweepGrid *d_grid,
                      const int2 *d_items, int64_t n) {
  if (n <= 0) return 0;
  const int64_t pe = (int64_t)bsx * bsx * bsx, fs = 2 * (int64_t)bsx * bsx, es = 4 * bsx;
  const int esz = storage_bytes(storage);
  MSBG_CHECK(n < 0x7fffffff, MSBG_EINVAL, "copy: too many blocks");
  if (esz == 2)
    hipLaunchKernelGGL(k_copy_items<uint16_t>, dim3((unsigned)n), dim3(256), 0, s, d_grid, d_items,
                       pe, fs, es);
  else if (esz == 4)
    hipLaunchKernelGGL(k_copy_items<float>, dim3((unsigned)n), dim3(256), 0, s, d_grid, d_items,
                       pe, fs, es);
  else
    hipLaunchKernelGGL(k_copy_items<double>, dim3((unsigned)n), dim3(256), 0, s, d_grid, d_items,
                       pe, fs, es);
  MSBG_HIP(hipGetLastError());
  return 0;
}

}  // namespace msbg
