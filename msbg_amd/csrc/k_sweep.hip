// 8-colour in-place mean-curvature-flow sweep: one colour pass per launch.
//
// Reference: MultiresSparseGrid::applyChannelPdeFast<T>, laplTyp 4 in colour
// mode (msbg4.cpp:60-957): per block, fill a (bsx+2)^3 float halo
// (fillHaloBlock_ with OPT_BC_NEUMANN, halo.cpp:267-479), evaluate the
// 19-point MCF stencil for every voxel from that halo (Jacobi inside the
// block), clamp to [0,1] and store in place. Blocks of one colour never
// touch each other's halos, so a pass is race-free.
//
// GPU mapping: one workgroup per block. The workgroup marches through the
// block's z-planes; each step stages one halo plane ((bsx+2)^2 values,
// interior from the block's page with 4-wide coalesced loads, the ring from
// the neighbour pages and side-car faces named by the halo index table)
// into a double-buffered LDS plane, and every thread pulls the patch around
// its voxel column (4x2 for 32^3 blocks) into registers. Three patches
// (z-1, z, z+1) rotate in registers, and plane loads run two planes ahead.
// All loads of a plane happen before the plane is overwritten, which keeps
// the in-place update exact.
//
// Reference paths are relative to the reference's src/ directory.

#include <hip/hip_runtime.h>

#include <type_traits>

#include "internal.h"
#include "storage.h"
#include "sweep_core.h"

namespace msbg {

namespace {

using namespace sweep;

template <int BSX, int ST>
int ensure_side(msbg_grid *g, ChanLevel &cl, int64_t fromPage = 0) {
  using R = typename Stor<ST>::raw;
  if (cl.side_valid) return 0;
  const int64_t per = 2 * BSX * BSX + 4 * BSX;
  MSBG_TRY(cl.d_side.ensure((size_t)(cl.npages * per * (int64_t)sizeof(R))));
  MSBG_TRY(cl.d_consts.ensure(8 * sizeof(R)));
  R c[8];
  for (int k = 0; k < 4; k++) {
    c[k] = raw_value<ST>(cl.empty_raw);
    c[4 + k] = raw_value<ST>(cl.full_raw);
  }
  MSBG_HIP(hipMemcpyAsync(cl.d_consts.p, c, sizeof c, hipMemcpyHostToDevice, g->stream));
  if (cl.npages > fromPage)
    hipLaunchKernelGGL((k_refresh_side<BSX, R>), dim3((unsigned)(cl.npages - fromPage)), dim3(256),
                       0, g->stream, (const R *)cl.d_pages.p, (R *)cl.d_side.p, cl.npages,
                       fromPage);
  MSBG_HIP(hipGetLastError());
  MSBG_HIP(hipStreamSynchronize(g->stream));
  cl.side_valid = true;
  return 0;
}

template <int BSX, int ST>
int prepare_t(msbg_grid *g, ChanLevel &cl, int64_t fromPage) {
  return ensure_side<BSX, ST>(g, cl, fromPage);
}

template <int BSX, int ST, bool EXACT, int PX, int PY, bool C01 = false>
int launch_one(hipStream_t s, const SweepGrid *d_grids, const int2 *d_items, int64_t n,
               double *d_partial, int iter) {
  constexpr int NT = Tile<BSX, PX, PY>::NT;
  constexpr int B = Pack<NT>::BPW;
  for (int64_t off = 0; off < n; off += 0x40000000) {
    const int cnt = (int)(n - off < 0x40000000 ? n - off : 0x40000000);
    if (d_partial)
      hipLaunchKernelGGL((k_sweep<BSX, ST, EXACT, PX, PY, OP_MCF, false, B, C01, true>),
                         dim3((unsigned)((cnt + B - 1) / B)), dim3(NT * B), 0, s, d_grids,
                         d_items + off, d_partial + off, iter, cnt);
    else
      hipLaunchKernelGGL((k_sweep<BSX, ST, EXACT, PX, PY, OP_MCF, false, B, C01, false>),
                         dim3((unsigned)((cnt + B - 1) / B)), dim3(NT * B), 0, s, d_grids,
                         d_items + off, nullptr, iter, cnt);
    MSBG_HIP(hipGetLastError());
  }
  return 0;
}

template <int BSX, int ST>
int launch_st(hipStream_t s, bool exact, bool c01, int tile, const SweepGrid *d_grids,
              const int2 *d_items, int64_t n, double *d_partial, int iter) {
  if constexpr (ST == MSBG_STORAGE_F64) exact = true;  // the FP64 run is always exact
  if (exact)  // the exact path keeps the default column shape
    return BSX >= 32 ? launch_one<BSX, ST, true, 4, 2>(s, d_grids, d_items, n, d_partial, iter)
                     : launch_one<BSX, ST, true, 2, 2>(s, d_grids, d_items, n, d_partial, iter);
  if constexpr (ST != MSBG_STORAGE_F64) {
    if (tile == 2)
      return c01 ? launch_one<BSX, ST, false, 4, 4, true>(s, d_grids, d_items, n, d_partial, iter)
                 : launch_one<BSX, ST, false, 4, 4>(s, d_grids, d_items, n, d_partial, iter);
    if (tile == 1)
      return c01 ? launch_one<BSX, ST, false, 4, 2, true>(s, d_grids, d_items, n, d_partial, iter)
                 : launch_one<BSX, ST, false, 4, 2>(s, d_grids, d_items, n, d_partial, iter);
    return c01 ? launch_one<BSX, ST, false, 2, 2, true>(s, d_grids, d_items, n, d_partial, iter)
               : launch_one<BSX, ST, false, 2, 2>(s, d_grids, d_items, n, d_partial, iter);
  }
  return 0;
}

template <int BSX>
int launch_bsx(hipStream_t s, int storage, bool exact, bool c01, int tile,
               const SweepGrid *d_grids, const int2 *d_items, int64_t n, double *d_partial,
               int iter) {
  switch (storage) {
    case MSBG_STORAGE_U16: return launch_st<BSX, MSBG_STORAGE_U16>(s, exact, c01, tile, d_grids, d_items, n, d_partial, iter);
    case MSBG_STORAGE_F32: return launch_st<BSX, MSBG_STORAGE_F32>(s, exact, c01, tile, d_grids, d_items, n, d_partial, iter);
    case MSBG_STORAGE_F16: return launch_st<BSX, MSBG_STORAGE_F16>(s, exact, c01, tile, d_grids, d_items, n, d_partial, iter);
    case MSBG_STORAGE_BF16: return launch_st<BSX, MSBG_STORAGE_BF16>(s, exact, c01, tile, d_grids, d_items, n, d_partial, iter);
    case MSBG_STORAGE_F64: return launch_st<BSX, MSBG_STORAGE_F64>(s, exact, c01, tile, d_grids, d_items, n, d_partial, iter);
    case MSBG_STORAGE_U8: return launch_st<BSX, MSBG_STORAGE_U8>(s, exact, c01, tile, d_grids, d_items, n, d_partial, iter);
  }
  set_error("sweep: bad storage %d", storage);
  return MSBG_EINVAL;
}

template <int BSX>
int prepare_bsx(msbg_grid *g, Channel &ch) {
  ChanLevel &cl = ch.lv[0];
  switch (ch.storage) {
    case MSBG_STORAGE_U16: return prepare_t<BSX, MSBG_STORAGE_U16>(g, cl, 0);
    case MSBG_STORAGE_F32: return prepare_t<BSX, MSBG_STORAGE_F32>(g, cl, 0);
    case MSBG_STORAGE_F16: return prepare_t<BSX, MSBG_STORAGE_F16>(g, cl, 0);
    case MSBG_STORAGE_BF16: return prepare_t<BSX, MSBG_STORAGE_BF16>(g, cl, 0);
    case MSBG_STORAGE_F64: return prepare_t<BSX, MSBG_STORAGE_F64>(g, cl, 0);
    case MSBG_STORAGE_U8: return prepare_t<BSX, MSBG_STORAGE_U8>(g, cl, 0);
  }
  set_error("sweep: bad storage %d", ch.storage);
  return MSBG_EINVAL;
}

}  // namespace

// Voxel columns of the fast path. 32^3 blocks: 4x4 for FP16 pages (one wave
// per block, FP16 patches), 4x2 otherwise; measured on the 4096^3 scene
// (FP16, fast): 3.25 ms per colour pass with 4x4 against 3.38 ms with 4x2,
// 3.36 ms with 8x2 and 3.84 ms with 4x8 (register pressure). 16^3 blocks:
// 4x2 when the launch holds at least 4096 of them (two per workgroup, about
// 8 workgroups per CU): on the scene's level 1 (16^3, FP16) 6.1e11
// voxel-updates/s against 5.8e11 with 2x2, while small grids swept one at a
// time lose 6 % with half the workgroups; with the later kernel 6.7e11
// against 6.1e11 with 4x4 (four blocks per wave).
int default_tile(int bsx, int storage, int64_t n) {
  if (bsx >= 32) return storage == MSBG_STORAGE_F16 ? 2 : 1;
  return bsx == 16 && n >= 4096 ? 1 : 0;
}

int sweep_prepare_level(msbg_grid *g, ChanLevel &cl, int storage, int64_t fromPage) {
  auto st = [&](auto bsxc) -> int {
    constexpr int B = decltype(bsxc)::value;
    switch (storage) {
      case MSBG_STORAGE_U16: return prepare_t<B, MSBG_STORAGE_U16>(g, cl, fromPage);
      case MSBG_STORAGE_F32: return prepare_t<B, MSBG_STORAGE_F32>(g, cl, fromPage);
      case MSBG_STORAGE_F16: return prepare_t<B, MSBG_STORAGE_F16>(g, cl, fromPage);
      case MSBG_STORAGE_BF16: return prepare_t<B, MSBG_STORAGE_BF16>(g, cl, fromPage);
      case MSBG_STORAGE_F64: return prepare_t<B, MSBG_STORAGE_F64>(g, cl, fromPage);
      case MSBG_STORAGE_U8: return prepare_t<B, MSBG_STORAGE_U8>(g, cl, fromPage);
    }
    set_error("sweep: bad storage %d", storage);
    return MSBG_EINVAL;
  };
  switch (cl.bsx) {
    case 8: return st(std::integral_constant<int, 8>{});
    case 16: return st(std::integral_constant<int, 16>{});
    case 32: return st(std::integral_constant<int, 32>{});
  }
  set_error("sweep: block size %d not supported by the sweep kernel (8, 16, 32)", cl.bsx);
  return MSBG_EUNSUPP;
}

int sweep_prepare(msbg_grid *g, Channel &ch) {
  switch (g->bsx0) {
    case 8: return prepare_bsx<8>(g, ch);
    case 16: return prepare_bsx<16>(g, ch);
    case 32: return prepare_bsx<32>(g, ch);
  }
  set_error("sweep: block size %d not supported by the sweep kernel (8, 16, 32)", g->bsx0);
  return MSBG_EUNSUPP;
}

SweepGrid sweep_desc(Channel &ch, float dt, int doConstr01) {
  ChanLevel &cl = ch.lv[0];
  SweepGrid d{};
  d.pages = cl.d_pages.p;
  d.side = cl.d_side.p;
  d.consts = cl.d_consts.p;
  d.nbr = cl.d_nbr.p;
  d.npages = cl.npages;
  d.dt = dt;
  d.lo = doConstr01 ? 0.0f : -__builtin_inff();
  d.hi = doConstr01 ? 1.0f : __builtin_inff();
  d.dst_pages = d.pages;
  d.dst_side = d.side;
  return d;
}

int launch_sweep_items(hipStream_t s, int bsx, int storage, bool exact, bool c01, int tile,
                       const SweepGrid *d_grids, const int2 *d_items, int64_t n,
                       double *d_partial, int iter) {
  if (tile < 0) tile = default_tile(bsx, storage, n);
  switch (bsx) {
    case 8: return launch_bsx<8>(s, storage, exact, c01, tile, d_grids, d_items, n, d_partial, iter);
    case 16: return launch_bsx<16>(s, storage, exact, c01, tile, d_grids, d_items, n, d_partial, iter);
    case 32: return launch_bsx<32>(s, storage, exact, c01, tile, d_grids, d_items, n, d_partial, iter);
  }
  set_error("sweep: block size %d not supported by the sweep kernel (8, 16, 32)", bsx);
  return MSBG_EUNSUPP;
}

}  // namespace msbg
