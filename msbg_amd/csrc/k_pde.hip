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
  constexpr int B = Pack<NT>::BPW;
  for (int64_t off = 0; off < n; off += 0x40000000) {
    const int cnt = (int)(n - off < 0x40000000 ? n - off : 0x40000000);
    hipLaunchKernelGGL((k_sweep<BSX, ST, true, PX, PY, OP, HARD, B>),
                       dim3((unsigned)((cnt + B - 1) / B)), dim3(NT * B), 0, s, d_grids,
                       d_items + off, d_partial ? d_partial + off : nullptr, iter, cnt);
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
    case MSBG_STORAGE_U8: return launch_st<BSX, MSBG_STORAGE_U8>(s, op, hard, d_grids, d_items, n, d_partial, iter);
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

// ---- bi-laplacian (op 2, halo 2, msbg4.cpp:637-680) ---------------------
// One workgroup per block; every voxel reads its 25 stencil values through
// the block's 27 neighbour codes (value page, const EMPTY/FULL, or the
// Neumann clamp into the own block for NULL blocks and the outside) and
// writes the new value to a scratch page; a second kernel commits the
// scratch pages and their side-cars, so the update reads old values only.
template <int BSX, int ST>
__global__ __launch_bounds__(256) void k_bilap(const SweepGrid *__restrict__ grids,
                                               const int2 *__restrict__ items, uint8_t *scratch,
                                               double *__restrict__ partial, int hard) {
#pragma clang fp contract(off)
  using S = Stor<ST>;
  using R = typename S::raw;
  constexpr int M = BSX - 1;
  constexpr int64_t PE = (int64_t)BSX * BSX * BSX;
  const int2 item = items[blockIdx.x];
  const SweepGrid &G = grids[item.x];
  __shared__ int codes[27];
  __shared__ double red[4];
  const int tid = threadIdx.x;
  for (int k = tid; k < 27; k += 256) codes[k] = G.nbr[(int64_t)item.y * 27 + k];
  __syncthreads();
  const R *pages = (const R *)G.pages;
  const R *own = pages + (int64_t)item.y * PE;
  const R *consts = (const R *)G.consts;
  R *out = (R *)scratch + (int64_t)blockIdx.x * PE;
  auto F = [&](int x, int y, int z) -> float {
    const int kx = x < 0 ? 0 : (x > M ? 2 : 1), ky = y < 0 ? 0 : (y > M ? 2 : 1),
              kz = z < 0 ? 0 : (z > M ? 2 : 1);
    const int c = codes[kz * 9 + ky * 3 + kx];
    if (c >= 0) return (float)S::dec(pages[(int64_t)c * PE + ((z & M) * BSX + (y & M)) * BSX + (x & M)]);
    if (c == MSBG_BLOCK_EMPTY) return (float)S::dec(consts[0]);
    if (c == MSBG_BLOCK_FULL) return (float)S::dec(consts[4]);
    x = x < 0 ? 0 : (x > M ? M : x);
    y = y < 0 ? 0 : (y > M ? M : y);
    z = z < 0 ? 0 : (z > M ? M : z);
    return (float)S::dec(own[(z * BSX + y) * BSX + x]);
  };
  const float dt = G.dt;
  double acc = 0.0;
  for (int v = tid; v < PE; v += 256) {
    const int x = v % BSX, y = (v / BSX) % BSX, z = v / (BSX * BSX);
    const float F0 = F(x, y, z);
    const float H =
        42.f * F0 -
        12.f * (F(x + 1, y, z) + F(x, y + 1, z) + F(x, y, z + 1) + F(x - 1, y, z) +
                F(x, y - 1, z) + F(x, y, z - 1)) +
        1.f * (F(x + 2, y, z) + F(x, y + 2, z) + F(x, y, z + 2) + F(x - 2, y, z) +
               F(x, y - 2, z) + F(x, y, z - 2)) +
        2.f * (F(x + 1, y, z + 1) + F(x, y + 1, z + 1) + F(x - 1, y, z + 1) +
               F(x, y - 1, z + 1) + F(x + 1, y, z - 1) + F(x, y + 1, z - 1) +
               F(x - 1, y, z - 1) + F(x, y - 1, z - 1) + F(x + 1, y + 1, z) +
               F(x - 1, y + 1, z) + F(x - 1, y - 1, z) + F(x + 1, y - 1, z));
    const float GX = 0.5f * (F(x + 1, y, z) - F(x - 1, y, z)),
                GY = 0.5f * (F(x, y + 1, z) - F(x, y - 1, z)),
                GZ = 0.5f * (F(x, y, z + 1) - F(x, y, z - 1));
    const float Gm = sqrtf(GX * GX + GY * GY + GZ * GZ);
    float Dv = -dt * H * Gm;
    Dv = Dv > -0.1f ? Dv : -0.1f;
    Dv = Dv < 0.1f ? Dv : 0.1f;
    float ph = F0 + Dv;
    if (hard) {
      const float f0 = (float)S::dec(((const R *)G.phi0)[(int64_t)item.y * PE + v]);
      float lo = f0 - G.tol1, hi = f0 + G.tol2;
      lo = 0.0f > lo ? 0.0f : lo;
      hi = 1.0f < hi ? 1.0f : hi;
      ph = ph < hi ? ph : hi;
      ph = ph > lo ? ph : lo;
    } else if (G.lo == 0.0f) {
      ph = ph > 0.0f ? ph : 0.0f;
      ph = ph < 1.0f ? ph : 1.0f;
    }
    const R e = S::enc(ph);
    out[v] = e;
    if (partial) {
      const double d = (double)S::dec(e) - (double)F0;
      acc += d * d;
    }
  }
  if (partial) {
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
  }
}

// scratch pages -> destination pages, and their side-car faces and edges
template <int BSX, typename R>
__global__ void k_commit(const SweepGrid *__restrict__ g, const int2 *__restrict__ items,
                         const uint8_t *scratch) {
  constexpr int M = BSX - 1, PS = BSX * BSX, FS = 2 * PS, ES = 4 * BSX;
  constexpr int64_t PE = (int64_t)PS * BSX;
  const int64_t page = items[blockIdx.x].y;
  const R *src = (const R *)scratch + (int64_t)blockIdx.x * PE;
  R *dp = (R *)g->dst_pages + page * PE;
  for (int64_t i = threadIdx.x; i < PE; i += blockDim.x) dp[i] = src[i];
  R *faces = (R *)g->dst_side + page * FS;
  R *edges = (R *)g->dst_side + g->npages * FS + page * ES;
  for (int t = threadIdx.x; t < FS + ES; t += blockDim.x) {
    if (t < FS) {
      const int sd = t / PS, z = (t % PS) / BSX, y = t % BSX;
      faces[t] = src[z * PS + y * BSX + (sd ? M : 0)];
    } else {
      const int u = t - FS, k = u / BSX, z = u % BSX;
      edges[u] = src[z * PS + ((k & 2) ? M : 0) * BSX + ((k & 1) ? M : 0)];
    }
  }
}

template <int BSX, int ST>
int bilap_t(hipStream_t s, const SweepGrid *d_grid, const int2 *d_items, int64_t n,
            double *d_partial, bool hard, uint8_t *scratch) {
  using R = typename Stor<ST>::raw;
  hipLaunchKernelGGL((k_bilap<BSX, ST>), dim3((unsigned)n), dim3(256), 0, s, d_grid, d_items,
                     scratch, d_partial, hard ? 1 : 0);
  MSBG_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_commit<BSX, R>), dim3((unsigned)n), dim3(256), 0, s, d_grid, d_items,
                     (const uint8_t *)scratch);
  MSBG_HIP(hipGetLastError());
  return 0;
}

template <int BSX>
int bilap_bsx(hipStream_t s, int storage, const SweepGrid *d_grid, const int2 *d_items,
              int64_t n, double *d_partial, bool hard, uint8_t *scratch) {
  switch (storage) {
    case MSBG_STORAGE_U16: return bilap_t<BSX, MSBG_STORAGE_U16>(s, d_grid, d_items, n, d_partial, hard, scratch);
    case MSBG_STORAGE_F32: return bilap_t<BSX, MSBG_STORAGE_F32>(s, d_grid, d_items, n, d_partial, hard, scratch);
    case MSBG_STORAGE_F16: return bilap_t<BSX, MSBG_STORAGE_F16>(s, d_grid, d_items, n, d_partial, hard, scratch);
    case MSBG_STORAGE_BF16: return bilap_t<BSX, MSBG_STORAGE_BF16>(s, d_grid, d_items, n, d_partial, hard, scratch);
  }
  set_error("applyChannelPdeFast: laplTyp 2 with storage %d is not supported", storage);
  return MSBG_EUNSUPP;
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

int launch_bilap_items(hipStream_t s, int bsx, int storage, bool hard, const SweepGrid *d_grid,
                       const int2 *d_items, int64_t n, double *d_partial, uint8_t *scratch) {
  MSBG_CHECK(n < 0x7fffffff, MSBG_EINVAL, "bi-laplacian: too many blocks");
  if (n <= 0) return 0;
  switch (bsx) {
    case 8: return bilap_bsx<8>(s, storage, d_grid, d_items, n, d_partial, hard, scratch);
    case 16: return bilap_bsx<16>(s, storage, d_grid, d_items, n, d_partial, hard, scratch);
    case 32: return bilap_bsx<32>(s, storage, d_grid, d_items, n, d_partial, hard, scratch);
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
  if (esz == 1)
    hipLaunchKernelGGL(k_copy_items<uint8_t>, dim3((unsigned)n), dim3(256), 0, s, d_grid, d_items,
                       pe, fs, es);
  else if (esz == 2)
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
