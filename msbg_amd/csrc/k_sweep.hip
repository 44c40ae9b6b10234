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

#include "internal.h"
#include "storage.h"

namespace msbg {

namespace {

// compile-time configuration of the sweep kernel
#ifndef MSBG_SWEEP_ALG
#define MSBG_SWEEP_ALG 1  // fold -2 f0 of fxx/fyy/fzz into H = Np/G - 4 f0
#endif
#ifndef MSBG_SWEEP_PD
#define MSBG_SWEEP_PD 1  // plane prefetch distance (1 or 2)
#endif
#ifndef MSBG_SWEEP_WAVES
#define MSBG_SWEEP_WAVES 0  // minimum waves per SIMD requested (0 = none)
#endif


// ---- vector raw loads/stores -------------------------------------------
__device__ inline void load4(const uint16_t *p, uint16_t v[4]) {
  uint2 u = *(const uint2 *)p;
  v[0] = (uint16_t)(u.x & 0xffffu);
  v[1] = (uint16_t)(u.x >> 16);
  v[2] = (uint16_t)(u.y & 0xffffu);
  v[3] = (uint16_t)(u.y >> 16);
}
__device__ inline void load4(const float *p, float v[4]) {
  float4 u = *(const float4 *)p;
  v[0] = u.x;
  v[1] = u.y;
  v[2] = u.z;
  v[3] = u.w;
}
__device__ inline void load4(const double *p, double v[4]) {
  double2 a = ((const double2 *)p)[0], b = ((const double2 *)p)[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}
__device__ inline void store2(uint16_t *p, uint16_t a, uint16_t b) {
  *(uint32_t *)p = (uint32_t)a | ((uint32_t)b << 16);
}
__device__ inline void store2(float *p, float a, float b) { *(float2 *)p = make_float2(a, b); }
__device__ inline void store2(double *p, double a, double b) {
  *(double2 *)p = make_double2(a, b);
}

// ---- the MCF update --------------------------------------------------------
// Patches: (PX+2) x (PY+2) values of one plane around the thread's PX x PY
// voxel column, row stride W = PX+2; halo-local x = i-1, y = j-1.

// Reference evaluation order (msbg4.cpp:559-637, 813-825), one rounding per
// operation: bit-identical to the strict CPU restatement.
template <typename C, int W>
__device__ inline C mcf_exact(const C *D, const C *M, const C *U, int c, C dt, bool c01) {
#pragma clang fp contract(off)
  const C h = (C)0.5, two = (C)2;
  C f0 = M[c], f1 = M[c - 1], f2 = M[c + 1], f3 = M[c - W], f4 = M[c + W], f5 = D[c], f6 = U[c];
  C fx = h * (f2 - f1), fy = h * (f4 - f3), fz = h * (f6 - f5);
  C fxx = f1 + f2 - two * f0, fyy = f3 + f4 - two * f0, fzz = f5 + f6 - two * f0;
  C fy_right = (M[c + W + 1] - M[c - W + 1]) * h, fy_left = (M[c + W - 1] - M[c - W - 1]) * h;
  C fz_right = (U[c + 1] - D[c + 1]) * h, fz_left = (U[c - 1] - D[c - 1]) * h;
  C fz_up = (U[c + W] - D[c + W]) * h, fz_down = (U[c - W] - D[c - W]) * h;
  C fyx = (fy_right - fy_left) * h, fzx = (fz_right - fz_left) * h, fzy = (fz_up - fz_down) * h;
  C H = (fy * fy + fz * fz) * fxx + (fx * fx + fz * fz) * fyy + (fx * fx + fy * fy) * fzz -
        two * (fx * fy * fyx + fx * fz * fzx + fy * fz * fzy);
  C g = fx * fx + fy * fy + fz * fz;
  // the threshold is the float literal 1e-7f, also for the FP64 run
  H = g > (C)1e-7f ? H / g : (C)0;
  C Dv = dt * H;
  Dv = Dv > (C)-0.1f ? Dv : (C)-0.1f;  // Vec8f max(a,b) = a > b ? a : b
  Dv = Dv < (C)0.1f ? Dv : (C)0.1f;    // Vec8f min(a,b) = a < b ? a : b
  C phi = f0 + Dv;
  if (c01) {
    phi = phi > (C)0 ? phi : (C)0;
    phi = phi < (C)1 ? phi : (C)1;
  }
  return phi;
}

// Fast evaluation of the same formula for a PX x PY column step. With
// d* = 2 f* (central differences), c** = 4 f** (mixed derivatives) and
// s** = the two neighbours' sum along an axis, the reference's
// H / |grad|^2 equals
//   Np / G - 4 f0,  Np = A sxx + B syy + C szz - 0.5 (dx dy cyx + dx dz czx + dy dz czy),
//   A = dy^2+dz^2, B = dx^2+dz^2, C = dx^2+dy^2, G = dx^2+dy^2+dz^2,
// because A+B+C = 2G folds the -2 f0 of fxx, fyy, fzz into one term; the
// threshold |grad|^2 > 1e-7 becomes G > 4e-7. Central differences shared by
// neighbouring voxels of the column are formed once.
template <int PX, int PY>
__device__ inline void mcf_fast(const float *D, const float *M, const float *U, float dt, float lo,
                                float hi, float *out) {
  constexpr int W = PX + 2;
  float colY[PY + 2][PX + 2], dzc[PY + 2][PX + 2];
#pragma unroll
  for (int j = 1; j <= PY; j++)
#pragma unroll
    for (int i = 0; i < W; i++) colY[j][i] = M[(j + 1) * W + i] - M[(j - 1) * W + i];
#pragma unroll
  for (int j = 0; j < PY + 2; j++)
#pragma unroll
    for (int i = 0; i < W; i++)
      if ((j >= 1 && j <= PY) || (i >= 1 && i <= PX)) dzc[j][i] = U[j * W + i] - D[j * W + i];
  const float m4dt = -4.0f * dt;
#pragma unroll
  for (int b = 0; b < PY; b++)
#pragma unroll
    for (int a = 0; a < PX; a++) {
      const int i = a + 1, j = b + 1, c = j * W + i;
      const float f0 = M[c];
      const float dx = M[c + 1] - M[c - 1], dy = colY[j][i], dz = dzc[j][i];
      const float sxx = M[c - 1] + M[c + 1], syy = M[c - W] + M[c + W], szz = D[c] + U[c];
      const float cyx = colY[j][i + 1] - colY[j][i - 1];
      const float czx = dzc[j][i + 1] - dzc[j][i - 1];
      const float czy = dzc[j + 1][i] - dzc[j - 1][i];
      const float dz2 = dz * dz, dx2 = dx * dx;
      const float A = __builtin_fmaf(dy, dy, dz2), B = dx2 + dz2, Cq = __builtin_fmaf(dy, dy, dx2);
      const float G = A + dx2;
      float p1 = dy * cyx;
      p1 = __builtin_fmaf(dz, czx, p1);
      float t = (dy * dz) * czy;
      t = __builtin_fmaf(dx, p1, t);
      const float r = __builtin_amdgcn_rcpf(G);
#if MSBG_SWEEP_ALG
      float Np = A * sxx;
      Np = __builtin_fmaf(B, syy, Np);
      Np = __builtin_fmaf(Cq, szz, Np);
      Np = __builtin_fmaf(-0.5f, t, Np);
      float Dv = __builtin_fmaf(Np * dt, r, m4dt * f0);  // dt * (Np/G - 4 f0)
#else
      float N = A * __builtin_fmaf(-2.0f, f0, sxx);
      N = __builtin_fmaf(B, __builtin_fmaf(-2.0f, f0, syy), N);
      N = __builtin_fmaf(Cq, __builtin_fmaf(-2.0f, f0, szz), N);
      N = __builtin_fmaf(-0.5f, t, N);
      float Dv = (N * dt) * r;
#endif
      Dv = G > 4.0f * 1e-7f ? Dv : 0.0f;
      Dv = __builtin_amdgcn_fmed3f(Dv, -0.1f, 0.1f);
      // doConstrZeroOne: [lo,hi] = [0,1], else (-inf,inf)
      out[b * PX + a] = __builtin_amdgcn_fmed3f(f0 + Dv, lo, hi);
    }
}

// ---- side-car refresh ---------------------------------------------------
// faces [page][2][z][y] (x = 0 and x = bsx-1), then edges [page][4][z]
// (edge k at x = (k&1) ? bsx-1 : 0, y = (k&2) ? bsx-1 : 0)
template <int BSX, typename R>
__global__ void k_refresh_side(const R *__restrict__ pages, R *__restrict__ side, int64_t npages) {
  constexpr int M = BSX - 1, PS = BSX * BSX, FS = 2 * PS, ES = 4 * BSX;
  constexpr int64_t PE = (int64_t)PS * BSX;
  const int64_t p = blockIdx.x;
  const R *pg = pages + p * PE;
  R *faces = side + p * FS;
  R *edges = side + npages * FS + p * ES;
  for (int t = threadIdx.x; t < FS + ES; t += blockDim.x) {
    if (t < FS) {
      int sd = t / PS, z = (t % PS) / BSX, y = t % BSX;
      faces[t] = pg[z * PS + y * BSX + (sd ? M : 0)];
    } else {
      int u = t - FS, k = u / BSX, z = u % BSX;
      edges[u] = pg[z * PS + ((k & 2) ? M : 0) * BSX + ((k & 1) ? M : 0)];
    }
  }
}

// ---- the pass kernel ----------------------------------------------------
// One source of halo values for one ring cell (or the interior of the
// planes outside the block): pointers for the plane below the block
// (zz = -1), for plane 0 of the block (advanced by `stride` per plane) and
// for the plane above (zz = bsx). Const blocks point into a small buffer of
// the const value with stride 0.
template <typename R>
struct Src {
  const R *m, *p0, *p;
  int stride;
  __device__ inline const R *at(int zz, int n) const {
    return zz < 0 ? m : (zz >= n ? p : p0 + zz * stride);
  }
};

// Voxel column per thread (PX x PY): the default is 4 x 2 for 32^3 blocks
// (128 threads) and 2 x 2 for smaller blocks. Measured on MI355X (FP16
// pages, 32^3 blocks): 4x2 with one plane of prefetch ran 10% faster than
// 2x2; 2x4, 4x4, 8x2 (register pressure) and a two-plane prefetch (lower
// occupancy) were slower. Batches pick per block size with an autotuner.
template <int BSX, int PX, int PY>
struct Tile {
  static constexpr int NT = (BSX / PX) * (BSX / PY);
};

#if MSBG_SWEEP_WAVES > 0
#define MSBG_SWEEP_LB(nt) __launch_bounds__(nt, MSBG_SWEEP_WAVES)
#else
#define MSBG_SWEEP_LB(nt) __launch_bounds__(nt)
#endif
template <int BSX, int ST, bool EXACT, int PX, int PY>
__global__ MSBG_SWEEP_LB((Tile<BSX, PX, PY>::NT)) void k_sweep(
    const SweepGrid *__restrict__ grids, const int2 *__restrict__ items,
    double *__restrict__ partial) {
  using S = Stor<ST>;
  using R = typename S::raw;
  using C = typename S::ct;
  constexpr int NT = Tile<BSX, PX, PY>::NT;
  constexpr int W = PX + 2, PN = (PX + 2) * (PY + 2);  // patch row stride / size
  constexpr int P = BSX + 2;                            // LDS row pitch
  constexpr int PS = BSX * BSX;
  constexpr int64_t PE = (int64_t)BSX * BSX * BSX;
  constexpr int FS = 2 * PS, ES = 4 * BSX;
  constexpr int NR = 4 * BSX + 4;  // ring cells of one halo plane
  constexpr int RPT = (NR + NT - 1) / NT;
  constexpr int LPT = PS / (4 * NT);  // 4-wide interior chunks per thread and plane
  constexpr int M = BSX - 1;
  static_assert(LPT * 4 * NT == PS, "plane does not split into 4-wide chunks");

  // work item: which grid, which page of it
  const int2 item = items[blockIdx.x];
  const SweepGrid G = grids[item.x];
  const int32_t *__restrict__ nbrTab = G.nbr;
  uint8_t *pagesRaw = G.pages;
  uint8_t *sideRaw = G.side;
  const uint8_t *constsRaw = G.consts;
  const int64_t npages = G.npages;
  const float dtf = G.dt, lo01 = G.lo, hi01 = G.hi;

  __shared__ C buf[2][P * P];
  __shared__ int codes[27];
  __shared__ double red[(NT + 63) / 64];

  const int tid = threadIdx.x;
  const int page = item.y;
  // 27 codes; small blocks have fewer than 27 threads
  for (int k = tid; k < 27; k += NT) codes[k] = nbrTab[(int64_t)page * 27 + k];
  __syncthreads();

  R *pages = (R *)pagesRaw;
  R *own = pages + (int64_t)page * PE;
  R *faces = (R *)sideRaw;                 // [npages][2][z][y]
  R *edges = faces + npages * FS;          // [npages][4][z]
  R *ownFaces = faces + (int64_t)page * FS;
  R *ownEdges = edges + (int64_t)page * ES;
  const R *consts = (const R *)constsRaw;  // [0..3] empty, [4..7] full
  const C dt = (C)dtf;

  // compute mapping: PX x PY column
  const int x0 = PX * (tid % (BSX / PX)), y0 = PY * (tid / (BSX / PX));

  // interior of the planes below/above the block (uniform over the block)
  auto zsrc = [&](int dzi, bool top) -> const R * {
    int c = codes[dzi * 9 + 4];
    if (c >= 0) return pages + (int64_t)c * PE + (top ? 0 : M) * PS;
    if (c == MSBG_BLOCK_EMPTY) return consts;
    if (c == MSBG_BLOCK_FULL) return consts + 4;
    return own + (top ? M : 0) * PS;  // Neumann: the block's own boundary plane
  };
  const R *zsrcM = zsrc(0, false), *zsrcP = zsrc(2, true);
  const bool constM = codes[4] == MSBG_BLOCK_EMPTY || codes[4] == MSBG_BLOCK_FULL;
  const bool constP = codes[22] == MSBG_BLOCK_EMPTY || codes[22] == MSBG_BLOCK_FULL;

  // ring cells of this thread and their sources
  Src<R> rs[RPT];
  int rlds[RPT];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    int r = tid + q * NT;
    if (r >= NR) r = NR - 1;  // loads a valid cell; only r < NR is stored
    int x, y;
    if (r < BSX) {
      x = -1;
      y = r;
    } else if (r < 2 * BSX) {
      x = BSX;
      y = r - BSX;
    } else if (r < 3 * BSX) {
      x = r - 2 * BSX;
      y = -1;
    } else if (r < 4 * BSX) {
      x = r - 3 * BSX;
      y = BSX;
    } else {
      int k = r - 4 * BSX;
      x = (k & 1) ? BSX : -1;
      y = (k & 2) ? BSX : -1;
    }
    rlds[q] = (y + 1) * P + x + 1;
    const int dxi = x < 0 ? 0 : (x >= BSX ? 2 : 1), dyi = y < 0 ? 0 : (y >= BSX ? 2 : 1);
    const bool xside = dxi != 1, yside = dyi != 1;
    // source pointer for the neighbour at (dxi, dyi, dzi) and plane zz
    auto one = [&](int dzi, int zz, int *stride) -> const R * {
      int c = codes[dzi * 9 + dyi * 3 + dxi];
      int zc = zz < 0 ? 0 : (zz > M ? M : zz);
      if (c == MSBG_BLOCK_EMPTY) {
        *stride = 0;
        return consts;
      }
      if (c == MSBG_BLOCK_FULL) {
        *stride = 0;
        return consts + 4;
      }
      if (c >= 0) {
        int zn = zz & M;  // plane inside the neighbour
        if (xside && yside) {  // corner: the neighbour's opposite z-edge
          int k = (x < 0 ? 1 : 0) + (y < 0 ? 2 : 0);
          *stride = 1;
          return edges + (int64_t)c * ES + k * BSX + zn;
        }
        if (xside) {  // x-face: the neighbour's opposite face, [z][y]
          *stride = BSX;
          return faces + (int64_t)c * FS + (x < 0 ? 1 : 0) * PS + zn * BSX + y;
        }
        *stride = PS;  // y-face: a contiguous row of the neighbour page
        return pages + (int64_t)c * PE + zn * PS + (y < 0 ? M : 0) * BSX + x;
      }
      // NULL block or outside the domain: the block's own clamped voxel
      if (xside && yside) {
        int k = (x < 0 ? 0 : 1) + (y < 0 ? 0 : 2);
        *stride = 1;
        return ownEdges + k * BSX + zc;
      }
      if (xside) {
        *stride = BSX;
        return ownFaces + (x < 0 ? 0 : 1) * PS + zc * BSX + y;
      }
      *stride = PS;
      return own + zc * PS + (y < 0 ? 0 : M) * BSX + x;
    };
    int sdummy, st0;
    rs[q].m = one(0, -1, &sdummy);
    rs[q].p0 = one(1, 0, &st0);
    rs[q].p = one(2, BSX, &sdummy);
    rs[q].stride = st0;
  }

  // two register sets of staged plane values (prefetch distance 2 planes)
  struct Staged {
    C in[LPT * 4];
    C ring[RPT];
  };
  auto dec = [](R r) -> C { return (C)S::dec(r); };

  // chunk k of this thread: 4 consecutive x of one row
  auto chunk_off = [&](int k) -> int {
    int e = (tid + k * NT) * 4;
    return e;  // = y * BSX + x with x multiple of 4
  };

  auto load_plane = [&](int zz, Staged &st) {
#pragma unroll
    for (int k = 0; k < LPT; k++) {
      const int e = chunk_off(k);
      const R *src = zz < 0 ? zsrcM + (constM ? 0 : e)
                            : (zz >= BSX ? zsrcP + (constP ? 0 : e) : own + zz * PS + e);
      R r[4];
      load4(src, r);
#pragma unroll
      for (int u = 0; u < 4; u++) st.in[k * 4 + u] = dec(r[u]);
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) st.ring[q] = dec(*rs[q].at(zz, BSX));
  };

  auto put_plane = [&](C *b, const Staged &st) {
#pragma unroll
    for (int k = 0; k < LPT; k++) {
      const int e = chunk_off(k);
      C *row = b + (e / BSX + 1) * P + (e % BSX) + 1;
#pragma unroll
      for (int u = 0; u < 4; u++) row[u] = st.in[k * 4 + u];
    }
#pragma unroll
    for (int q = 0; q < RPT; q++)
      if (tid + q * NT < NR) b[rlds[q]] = st.ring[q];
  };

  auto get_patch = [&](const C *b, C (&p)[PN]) {
#pragma unroll
    for (int j = 0; j < PY + 2; j++)
#pragma unroll
      for (int i = 0; i < PX + 2; i++) p[j * W + i] = b[(y0 + j) * P + x0 + i];
  };

  double acc = 0.0;
  const bool wFace0 = x0 == 0, wFace1 = x0 == BSX - PX, wY0 = y0 == 0, wY1 = y0 == BSX - PY;

  auto compute = [&](int z, const C (&Dp)[PN], const C (&Mp)[PN], const C (&Up)[PN]) {
    C o[PX * PY];
    if constexpr (EXACT || sizeof(C) == 8) {
      const bool c01 = lo01 == 0.0f;
#pragma unroll
      for (int b = 0; b < PY; b++)
#pragma unroll
        for (int a = 0; a < PX; a++)
          o[b * PX + a] = mcf_exact<C, W>(Dp, Mp, Up, (b + 1) * W + a + 1, dt, c01);
    } else {
      mcf_fast<PX, PY>(Dp, Mp, Up, dt, lo01, hi01, o);
    }
    R e[PX * PY];
#pragma unroll
    for (int k = 0; k < PX * PY; k++) e[k] = S::enc(o[k]);
#pragma unroll
    for (int b = 0; b < PY; b++)
#pragma unroll
      for (int a = 0; a < PX; a += 2)
        store2(own + z * PS + (y0 + b) * BSX + x0 + a, e[b * PX + a], e[b * PX + a + 1]);
    // keep the side-car faces and edges equal to the page
    if (wFace0) {
#pragma unroll
      for (int b = 0; b < PY; b++) ownFaces[z * BSX + y0 + b] = e[b * PX];
    }
    if (wFace1) {
#pragma unroll
      for (int b = 0; b < PY; b++) ownFaces[PS + z * BSX + y0 + b] = e[b * PX + PX - 1];
    }
    if (wFace0 && wY0) ownEdges[0 * BSX + z] = e[0];
    if (wFace1 && wY0) ownEdges[1 * BSX + z] = e[PX - 1];
    if (wFace0 && wY1) ownEdges[2 * BSX + z] = e[(PY - 1) * PX];
    if (wFace1 && wY1) ownEdges[3 * BSX + z] = e[PY * PX - 1];
    if (partial) {
#pragma unroll
      for (int b = 0; b < PY; b++)
#pragma unroll
        for (int a = 0; a < PX; a++) {
          double d = (double)S::dec(e[b * PX + a]) - (double)Mp[(b + 1) * W + a + 1];
          acc += d * d;
        }
    }
  };

  C pA[PN], pB[PN], pC[PN];
#if MSBG_SWEEP_PD == 2
  Staged s0, s1;  // plane zz lives in s[zz & 1]

  // prologue: planes -1 and 0 into pA, pB; planes 1 and 2 prefetched
  load_plane(-1, s1);
  load_plane(0, s0);
  put_plane(buf[1], s1);
  load_plane(1, s1);
  __syncthreads();
  get_patch(buf[1], pA);
  put_plane(buf[0], s0);
  load_plane(2, s0);
  __syncthreads();
  get_patch(buf[0], pB);

  // step z: stage plane z+1, prefetch z+3 into its register set, compute z
  auto step = [&](int z, Staged &st, C (&Dp)[PN], C (&Mp)[PN], C (&Up)[PN]) {
    put_plane(buf[(z + 1) & 1], st);
    if (z + 3 <= BSX) load_plane(z + 3, st);
    __syncthreads();
    get_patch(buf[(z + 1) & 1], Up);
    compute(z, Dp, Mp, Up);
  };

  // unrolled by 6: patches rotate with period 3, register sets with period 2
  for (int z = 0; z < BSX; z += 6) {
    step(z, s1, pA, pB, pC);
    if (z + 1 < BSX) step(z + 1, s0, pB, pC, pA);
    if (z + 2 < BSX) step(z + 2, s1, pC, pA, pB);
    if (z + 3 < BSX) step(z + 3, s0, pA, pB, pC);
    if (z + 4 < BSX) step(z + 4, s1, pB, pC, pA);
    if (z + 5 < BSX) step(z + 5, s0, pC, pA, pB);
  }
#else
  Staged s0;

  // prologue: planes -1 and 0 into pA, pB; plane 1 prefetched
  load_plane(-1, s0);
  put_plane(buf[1], s0);
  load_plane(0, s0);
  __syncthreads();
  get_patch(buf[1], pA);
  put_plane(buf[0], s0);
  load_plane(1, s0);
  __syncthreads();
  get_patch(buf[0], pB);

  // step z: stage plane z+1 (prefetched), prefetch z+2, compute plane z
  auto step = [&](int z, C (&Dp)[PN], C (&Mp)[PN], C (&Up)[PN]) {
    put_plane(buf[(z + 1) & 1], s0);
    if (z + 2 <= BSX) load_plane(z + 2, s0);
    __syncthreads();
    get_patch(buf[(z + 1) & 1], Up);
    compute(z, Dp, Mp, Up);
  };

  for (int z = 0; z < BSX; z += 3) {
    step(z, pA, pB, pC);
    if (z + 1 < BSX) step(z + 1, pB, pC, pA);
    if (z + 2 < BSX) step(z + 2, pC, pA, pB);
  }
#endif

  if (partial) {
    // deterministic block reduction: wave shuffles, then waves in order
    constexpr int WV = NT < 64 ? NT : 64;
    for (int off = WV / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, WV);
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
      double sum = 0.0;
      for (int w = 0; w < (NT + 63) / 64; w++) sum += red[w];
      partial[blockIdx.x] = sum;
    }
  }
}

template <int BSX, int ST>
int ensure_side(msbg_grid *g, ChanLevel &cl) {
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
  if (cl.npages)
    hipLaunchKernelGGL((k_refresh_side<BSX, R>), dim3((unsigned)cl.npages), dim3(256), 0, g->stream,
                       (const R *)cl.d_pages.p, (R *)cl.d_side.p, cl.npages);
  MSBG_HIP(hipGetLastError());
  MSBG_HIP(hipStreamSynchronize(g->stream));
  cl.side_valid = true;
  return 0;
}

template <int BSX, int ST>
int prepare_t(msbg_grid *g, ChanLevel &cl) {
  return ensure_side<BSX, ST>(g, cl);
}

template <int BSX, int ST, bool EXACT, int PX, int PY>
int launch_one(hipStream_t s, const SweepGrid *d_grids, const int2 *d_items, int64_t n,
               double *d_partial) {
  constexpr int NT = Tile<BSX, PX, PY>::NT;
  for (int64_t off = 0; off < n; off += 0x7fffffff) {
    unsigned cnt = (unsigned)(n - off < 0x7fffffff ? n - off : 0x7fffffff);
    hipLaunchKernelGGL((k_sweep<BSX, ST, EXACT, PX, PY>), dim3(cnt), dim3(NT), 0, s, d_grids,
                       d_items + off, d_partial ? d_partial + off : nullptr);
    MSBG_HIP(hipGetLastError());
  }
  return 0;
}

template <int BSX, int ST>
int launch_st(hipStream_t s, bool exact, int tile, const SweepGrid *d_grids, const int2 *d_items,
              int64_t n, double *d_partial) {
  if constexpr (ST == MSBG_STORAGE_F64) exact = true;  // the FP64 run is always exact
  if (exact)  // the exact path keeps the default column shape
    return BSX >= 32 ? launch_one<BSX, ST, true, 4, 2>(s, d_grids, d_items, n, d_partial)
                     : launch_one<BSX, ST, true, 2, 2>(s, d_grids, d_items, n, d_partial);
  if constexpr (ST != MSBG_STORAGE_F64) {
    if (tile == 1) return launch_one<BSX, ST, false, 4, 2>(s, d_grids, d_items, n, d_partial);
    return launch_one<BSX, ST, false, 2, 2>(s, d_grids, d_items, n, d_partial);
  }
  return 0;
}

template <int BSX>
int launch_bsx(hipStream_t s, int storage, bool exact, int tile, const SweepGrid *d_grids,
               const int2 *d_items, int64_t n, double *d_partial) {
  switch (storage) {
    case MSBG_STORAGE_U16: return launch_st<BSX, MSBG_STORAGE_U16>(s, exact, tile, d_grids, d_items, n, d_partial);
    case MSBG_STORAGE_F32: return launch_st<BSX, MSBG_STORAGE_F32>(s, exact, tile, d_grids, d_items, n, d_partial);
    case MSBG_STORAGE_F16: return launch_st<BSX, MSBG_STORAGE_F16>(s, exact, tile, d_grids, d_items, n, d_partial);
    case MSBG_STORAGE_BF16: return launch_st<BSX, MSBG_STORAGE_BF16>(s, exact, tile, d_grids, d_items, n, d_partial);
    case MSBG_STORAGE_F64: return launch_st<BSX, MSBG_STORAGE_F64>(s, exact, tile, d_grids, d_items, n, d_partial);
  }
  set_error("sweep: bad storage %d", storage);
  return MSBG_EINVAL;
}

template <int BSX>
int prepare_bsx(msbg_grid *g, Channel &ch) {
  ChanLevel &cl = ch.lv[0];
  switch (ch.storage) {
    case MSBG_STORAGE_U16: return prepare_t<BSX, MSBG_STORAGE_U16>(g, cl);
    case MSBG_STORAGE_F32: return prepare_t<BSX, MSBG_STORAGE_F32>(g, cl);
    case MSBG_STORAGE_F16: return prepare_t<BSX, MSBG_STORAGE_F16>(g, cl);
    case MSBG_STORAGE_BF16: return prepare_t<BSX, MSBG_STORAGE_BF16>(g, cl);
    case MSBG_STORAGE_F64: return prepare_t<BSX, MSBG_STORAGE_F64>(g, cl);
  }
  set_error("sweep: bad storage %d", ch.storage);
  return MSBG_EINVAL;
}

}  // namespace

int default_tile(int bsx) { return bsx >= 32 ? 1 : 0; }

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
  return d;
}

int launch_sweep_items(hipStream_t s, int bsx, int storage, bool exact, int tile,
                       const SweepGrid *d_grids, const int2 *d_items, int64_t n,
                       double *d_partial) {
  if (tile < 0) tile = default_tile(bsx);
  switch (bsx) {
    case 8: return launch_bsx<8>(s, storage, exact, tile, d_grids, d_items, n, d_partial);
    case 16: return launch_bsx<16>(s, storage, exact, tile, d_grids, d_items, n, d_partial);
    case 32: return launch_bsx<32>(s, storage, exact, tile, d_grids, d_items, n, d_partial);
  }
  set_error("sweep: block size %d not supported by the sweep kernel (8, 16, 32)", bsx);
  return MSBG_EUNSUPP;
}

}  // namespace msbg
