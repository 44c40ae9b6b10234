// Core of the sweep kernels: vector loads/stores, the MCF update (exact
// and fast), the other applyChannelPdeFast stencils, side-car refresh and
// the z-marching pass kernel, templated on block size, storage, arithmetic,
// voxel column, operator and hard constraints. Included by the files that
// instantiate the kernels.
//
// Reference paths are relative to the reference's src/ directory.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "internal.h"
#include "storage.h"

namespace msbg {
namespace sweep {

// ---- vector raw loads/stores -------------------------------------------
// Page, side-car and const pointers come out of the grid descriptor as
// generic pointers; casting them to the global address space at each access
// makes the compiler emit global (not flat) loads and stores, which wait on
// the vector-memory counter only.
#define MSBG_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ inline T ldg(const T *p) {
  return *(const MSBG_GLOBAL T *)p;
}
template <typename T>
__device__ inline void stg(T *p, T v) {
  *(MSBG_GLOBAL T *)p = v;
}

typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));

__device__ inline void load4(const uint16_t *p, uint16_t v[4]) {
  const uint64_t u = *(const MSBG_GLOBAL uint64_t *)p;
  v[0] = (uint16_t)(u & 0xffffu);
  v[1] = (uint16_t)((u >> 16) & 0xffffu);
  v[2] = (uint16_t)((u >> 32) & 0xffffu);
  v[3] = (uint16_t)(u >> 48);
}
__device__ inline void load4(const uint8_t *p, uint8_t v[4]) {
  const uint32_t u = *(const MSBG_GLOBAL uint32_t *)p;
  v[0] = (uint8_t)(u & 0xffu);
  v[1] = (uint8_t)((u >> 8) & 0xffu);
  v[2] = (uint8_t)((u >> 16) & 0xffu);
  v[3] = (uint8_t)(u >> 24);
}
__device__ inline void load4(const float *p, float v[4]) {
  const f32x4_t u = *(const MSBG_GLOBAL f32x4_t *)p;
  v[0] = u.x;
  v[1] = u.y;
  v[2] = u.z;
  v[3] = u.w;
}
__device__ inline void load4(const double *p, double v[4]) {
  const f64x2_t a = ((const MSBG_GLOBAL f64x2_t *)p)[0], b = ((const MSBG_GLOBAL f64x2_t *)p)[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}
__device__ inline void store2(uint16_t *p, uint16_t a, uint16_t b) {
  *(MSBG_GLOBAL uint32_t *)p = (uint32_t)a | ((uint32_t)b << 16);
}
__device__ inline void store2(uint8_t *p, uint8_t a, uint8_t b) {
  *(MSBG_GLOBAL uint16_t *)p = (uint16_t)((uint32_t)a | ((uint32_t)b << 8));
}
__device__ inline void store2(float *p, float a, float b) {
  *(MSBG_GLOBAL f32x2_t *)p = f32x2_t{a, b};
}
__device__ inline void store2(double *p, double a, double b) {
  *(MSBG_GLOBAL f64x2_t *)p = f64x2_t{a, b};
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

// ---- the fast MCF update ----------------------------------------------------
// One plane's (PX+2) x (PY+2) patch around a thread's PX x PY voxel column
// (halo-local x = i-1, y = j-1). Float patches hold decoded values; FP16
// patches keep the storage bits, two to a register, so that the arithmetic
// can take them as FP16 operands of v_fma_mix_f32 without a conversion.
typedef uint32_t __attribute__((may_alias)) u32a;
typedef uint64_t __attribute__((may_alias)) u64a;
typedef uint16_t __attribute__((may_alias)) u16a;
typedef _Float16 h2x_t __attribute__((ext_vector_type(2)));

template <typename L, int PX, int PY>
struct Patch {
  static constexpr int W = PX + 2;
  L v[(PY + 2) * W];
  __device__ inline L at(int j, int i) const { return v[j * W + i]; }
};
template <int PX, int PY>
struct Patch<_Float16, PX, PY> {
  static constexpr int WW = (PX + 2) / 2;
  uint32_t w[PY + 2][WW];
  __device__ inline _Float16 at(int j, int i) const {
    return __builtin_bit_cast(_Float16, (uint16_t)(w[j][i >> 1] >> ((i & 1) * 16)));
  }
};

// a + b and a - b of two patch values, one rounding in FP32. FP16 operands
// go straight into v_fma_mix_f32 as a * 1.0 + (+-b), each widened exactly
// inside the instruction. The instruction is written out because the
// compiler, seeing one FP16 value used by several sums, converts it once and
// then spends a separate v_cvt_f32_f16 on every value of the patch.
__device__ inline float vadd(float a, float b) { return a + b; }
__device__ inline float vsub(float a, float b) { return a - b; }

// a + s * b for the FP16 halves ha / hb of the words a / b (s = +-1)
#define MSBG_MIX(SIGN, OPSEL)                                                        \
  asm("v_fma_mix_f32 %0, %1, 1.0, " SIGN "%2 " OPSEL " op_sel_hi:[1,0,1]" : "=v"(r) \
      : "v"(a), "v"(b))
__device__ inline float mix_add(uint32_t a, int ha, uint32_t b, int hb) {
  float r;
  if (ha == 0 && hb == 0) MSBG_MIX("", "");
  else if (ha == 1 && hb == 0) MSBG_MIX("", "op_sel:[1,0,0]");
  else if (ha == 0 && hb == 1) MSBG_MIX("", "op_sel:[0,0,1]");
  else MSBG_MIX("", "op_sel:[1,0,1]");
  return r;
}
__device__ inline float mix_sub(uint32_t a, int ha, uint32_t b, int hb) {
  float r;
  if (ha == 0 && hb == 0) MSBG_MIX("-", "");
  else if (ha == 1 && hb == 0) MSBG_MIX("-", "op_sel:[1,0,0]");
  else if (ha == 0 && hb == 1) MSBG_MIX("-", "op_sel:[0,0,1]");
  else MSBG_MIX("-", "op_sel:[1,0,1]");
  return r;
}
#undef MSBG_MIX

// f0 * -4 + w and dt * m + f0 (clamped to [0, 1] when C01) with f0 the FP16
// half h of the word a
__device__ inline float mix_m4(uint32_t a, int h, float w) {
  float r;
  if (h == 0)
    asm("v_fma_mix_f32 %0, %1, -4.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(a), "v"(w));
  else
    asm("v_fma_mix_f32 %0, %1, -4.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r)
        : "v"(a), "v"(w));
  return r;
}
template <bool C01>
__device__ inline float mix_step(float dt, float m, uint32_t a, int h) {
  float r;
  if (C01) {
    if (h == 0)
      asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,0,1] clamp" : "=v"(r)
          : "v"(dt), "v"(m), "v"(a));
    else
      asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[0,0,1] op_sel_hi:[0,0,1] clamp" : "=v"(r)
          : "v"(dt), "v"(m), "v"(a));
  } else {
    if (h == 0)
      asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,0,1]" : "=v"(r) : "v"(dt), "v"(m), "v"(a));
    else
      asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "=v"(r)
          : "v"(dt), "v"(m), "v"(a));
  }
  return r;
}

// sums and differences of patch values A(ja, ia) and B(jb, ib)
template <typename L, int PX, int PY>
__device__ inline float pdiff(const Patch<L, PX, PY> &A, int ja, int ia,
                              const Patch<L, PX, PY> &B, int jb, int ib) {
  return vsub((float)A.at(ja, ia), (float)B.at(jb, ib));
}
template <typename L, int PX, int PY>
__device__ inline float psum(const Patch<L, PX, PY> &A, int ja, int ia,
                             const Patch<L, PX, PY> &B, int jb, int ib) {
  return vadd((float)A.at(ja, ia), (float)B.at(jb, ib));
}
template <int PX, int PY>
__device__ inline float pdiff(const Patch<_Float16, PX, PY> &A, int ja, int ia,
                              const Patch<_Float16, PX, PY> &B, int jb, int ib) {
  return mix_sub(A.w[ja][ia >> 1], ia & 1, B.w[jb][ib >> 1], ib & 1);
}
template <int PX, int PY>
__device__ inline float psum(const Patch<_Float16, PX, PY> &A, int ja, int ia,
                             const Patch<_Float16, PX, PY> &B, int jb, int ib) {
  return mix_add(A.w[ja][ia >> 1], ia & 1, B.w[jb][ib >> 1], ib & 1);
}
// w - 4 f0 and the limited step dt * m + f0 for f0 = M(j, i)
template <typename L, int PX, int PY>
__device__ inline float pm4(const Patch<L, PX, PY> &M, int j, int i, float w) {
  return __builtin_fmaf((float)M.at(j, i), -4.0f, w);
}
template <int PX, int PY>
__device__ inline float pm4(const Patch<_Float16, PX, PY> &M, int j, int i, float w) {
  return mix_m4(M.w[j][i >> 1], i & 1, w);
}
template <bool C01, typename L, int PX, int PY>
__device__ inline float pstep(const Patch<L, PX, PY> &M, int j, int i, float dt, float m,
                              float lo, float hi) {
  const float phi = __builtin_fmaf(dt, m, (float)M.at(j, i));
  return C01 ? __builtin_amdgcn_fmed3f(phi, 0.0f, 1.0f) : __builtin_amdgcn_fmed3f(phi, lo, hi);
}
template <bool C01, int PX, int PY>
__device__ inline float pstep(const Patch<_Float16, PX, PY> &M, int j, int i, float dt, float m,
                              float lo, float hi) {
  const float phi = mix_step<C01>(dt, m, M.w[j][i >> 1], i & 1);
  return C01 ? phi : __builtin_amdgcn_fmed3f(phi, lo, hi);
}

// Fast evaluation of the reference's update (msbg4.cpp:559-637) for a PX x PY
// column step. With d* = 2 f* (central differences), c** = 4 f** (mixed
// derivatives) and s** = the two neighbours' sum along an axis, the
// reference's H / |grad|^2 equals
//   Np / G - 4 f0,  Np = A sxx + B syy + C szz - 0.5 (dx dy cyx + dx dz czx + dy dz czy),
//   A = dy^2+dz^2, B = dx^2+dz^2, C = dx^2+dy^2, G = dx^2+dy^2+dz^2,
// because A+B+C = 2G folds the -2 f0 of fxx, fyy, fzz into one term; the
// threshold |grad|^2 > 1e-7 becomes G > 4e-7. The central differences dy and
// dz of the column and its x-neighbours (colY, dzc) are formed once per step
// and shared by neighbouring voxels. The limiter is applied as
// dt * clamp(u, +-0.1/dt) (lim = 0.1/dt), so that the last step is one fma
// onto f0 whose [0,1] clamp is the instruction's clamp bit.
template <int PX, int PY, bool C01, typename PT>
__device__ inline void mcf_fast(const PT &D, const PT &M, const PT &U, float dt, float lim,
                                float lo, float hi, float *out) {
  constexpr int W = PX + 2;
  float colY[PY][W], dzc[PY + 2][W];
#pragma unroll
  for (int b = 0; b < PY; b++)
#pragma unroll
    for (int i = 0; i < W; i++) colY[b][i] = pdiff(M, b + 2, i, M, b, i);
#pragma unroll
  for (int j = 0; j < PY + 2; j++)
#pragma unroll
    for (int i = 0; i < W; i++)
      if ((j >= 1 && j <= PY) || (i >= 1 && i <= PX)) dzc[j][i] = pdiff(U, j, i, D, j, i);
#pragma unroll
  for (int b = 0; b < PY; b++)
#pragma unroll
    for (int a = 0; a < PX; a++) {
      const int i = a + 1, j = b + 1;
      const float sxx = psum(M, j, i - 1, M, j, i + 1);
      const float syy = psum(M, j - 1, i, M, j + 1, i);
      const float szz = psum(D, j, i, U, j, i);
      const float dx = pdiff(M, j, i + 1, M, j, i - 1), dy = colY[b][i], dz = dzc[j][i];
      const float cyx = colY[b][i + 1] - colY[b][i - 1];
      const float czx = dzc[j][i + 1] - dzc[j][i - 1];
      const float czy = dzc[j + 1][i] - dzc[j - 1][i];
      // five operations for A, B, C and G: C = G - dz^2
      const float dz2 = dz * dz;
      const float A = __builtin_fmaf(dy, dy, dz2), B = __builtin_fmaf(dx, dx, dz2);
      const float G = __builtin_fmaf(dx, dx, A), Cq = G - dz2;
      float p1 = dy * cyx;
      p1 = __builtin_fmaf(dz, czx, p1);
      float t = (dy * dz) * czy;
      t = __builtin_fmaf(dx, p1, t);
      float Np = A * sxx;
      Np = __builtin_fmaf(B, syy, Np);
      Np = __builtin_fmaf(Cq, szz, Np);
      Np = __builtin_fmaf(-0.5f, t, Np);
      const float r = __builtin_amdgcn_rcpf(G);
      float u = pm4(M, j, i, Np * r);
      u = G > 4.0f * 1e-7f ? u : 0.0f;
      const float m = __builtin_amdgcn_fmed3f(u, -lim, lim);
      out[b * PX + a] = pstep<C01>(M, j, i, dt, m, lo, hi);
    }
}

// ---- the other stencils of applyChannelPdeFast (7-point) ----------------
// Reference evaluation order, one rounding per operation, as the oracle.
constexpr int OP_LAPL = 1, OP_DIFF = 3, OP_MCF = 4, OP_PM = 14, OP_PMJ = 15, OP_DIFFJ = 17;

// FMA_FAST_EXP (fastmath.h:583-628), including its repeated PX2 term
__device__ inline float ref_fast_exp(float x0) {
#pragma clang fp contract(off)
  const float MAXLOGF = 88.72283905206835f, MINLOGF = -88.f;
  const float C1F = 0.693359375f, C2F = -2.12194440e-4f;
  const float PX1 = 1.9875691500E-4f, PX2 = 1.3981999507E-3f, PX3 = 8.3334519073E-3f,
              PX4 = 4.1665795894E-2f, PX5 = 1.6666665459E-1f, PX6 = 5.0000001201E-1f;
  const float LOG2EF = 1.44269504088896341f;
  float x = x0;
  float z = floorf(LOG2EF * x + 0.5f);
  x -= z * C1F;
  x -= z * C2F;
  const int n = (int)z;
  const float x2 = x * x;
  z = x * PX1;
  z += PX2;
  z *= x;
  z += PX2;
  z *= x;
  z += PX3;
  z *= x;
  z += PX4;
  z *= x;
  z += PX5;
  z *= x;
  z += PX6;
  z *= x2;
  z += x + 1.0f;
  z *= __uint_as_float((uint32_t)(n + 0x7f) << 23);
  if (x0 < MINLOGF) z = 0.0f;
  if (x0 > MAXLOGF) z = __builtin_inff();
  return z;
}

// G_ of msbg4.cpp:761-764
__device__ inline float pm_weight(float f, float alpha) {
#pragma clang fp contract(off)
  if (alpha < 0) {
    const float s = fabsf(f) * (1.0f / -alpha);
    return ref_fast_exp(-(s * s));
  }
  const float s = fabsf(f) / alpha;
  return 1.0f / (1.0f + s * s);
}

template <int OP, int W>
__device__ inline float pde7_exact(const float *D, const float *M, const float *U, int c, float dt,
                                   float taubin, float alpha, int iter) {
#pragma clang fp contract(off)
  const float F0 = M[c], F1 = M[c - 1], F2 = M[c + 1], F3 = M[c - W], F4 = M[c + W], F5 = D[c],
              F6 = U[c];
  float phi = F0;
  if constexpr (OP == OP_LAPL) {  // msbg4.cpp:682-712
    const float L = F1 + F2 + F3 + F4 + F5 + F6 - 6.0f * F0;
    const float GX = 0.5f * (F2 - F1), GY = 0.5f * (F4 - F3), GZ = 0.5f * (F6 - F5);
    const float G = sqrtf(GX * GX + GY * GY + GZ * GZ);
    const float Dv = dt * L * G;
    if (taubin != 0.0f)
      phi = (iter & 1) ? phi + Dv : phi - taubin * Dv;
    else
      phi += Dv;
  } else if constexpr (OP == OP_DIFF) {  // msbg4.cpp:713-735
    const float Sm = (1.0f / 6.0f) * (F1 + F2 + F3 + F4 + F5 + F6);
    const float Dv = dt * (Sm - phi);
    phi += Dv;
  } else if constexpr (OP == OP_PM || OP == OP_PMJ) {  // msbg4.cpp:736-790
    const float g1 = F1 - F0, g2 = F2 - F0, g3 = F3 - F0, g4 = F4 - F0, g5 = F5 - F0,
                g6 = F6 - F0;
    const float G1 = pm_weight(g1, alpha), G2 = pm_weight(g2, alpha), G3 = pm_weight(g3, alpha),
                G4 = pm_weight(g4, alpha), G5 = pm_weight(g5, alpha), G6 = pm_weight(g6, alpha);
    if constexpr (OP == OP_PMJ) {
      const float sum = G1 * F1 + G2 * F2 + G3 * F3 + G4 * F4 + G5 * F5 + G6 * F6;
      const float wsum = G1 + G2 + G3 + G4 + G5 + G6;
      phi = (phi + dt * sum) * 1.0f / (1.0f + dt * wsum);
    } else {
      const float sum = G1 * g1 + G2 * g2 + G3 * g3 + G4 * g4 + G5 * g5 + G6 * g6;
      phi += dt * sum;
    }
  } else if constexpr (OP == OP_DIFFJ) {  // msbg4.cpp:791-805
    phi = (phi + dt * (F1 + F2 + F3 + F4 + F5 + F6)) * 1.0f / (1.0f + dt * 6.0f);
  }
  return phi;
}

// ---- side-car refresh ---------------------------------------------------
// faces [page][2][z][y] (x = 0 and x = bsx-1), then edges [page][4][z]
// (edge k at x = (k&1) ? bsx-1 : 0, y = (k&2) ? bsx-1 : 0)
// One thread per row (z, y) of the page: its first and last voxel go to the
// x-faces [side][z][y], and rows y = 0 and y = M also give the z-edges
// [k][z], k = (x == M) + 2 (y == M). Both reads share the row's cache line.
// Workgroup i refreshes page first + i of the pool's npages.
template <int BSX, typename R>
__global__ void k_refresh_side(const R *__restrict__ pages, R *__restrict__ side, int64_t npages,
                               int64_t first) {
  constexpr int M = BSX - 1, PS = BSX * BSX, FS = 2 * PS, ES = 4 * BSX;
  constexpr int64_t PE = (int64_t)PS * BSX;
  const int64_t p = first + blockIdx.x;
  const R *pg = pages + p * PE;
  R *faces = side + p * FS;
  R *edges = side + npages * FS + p * ES;
  for (int r = threadIdx.x; r < PS; r += blockDim.x) {
    const R a = pg[r * BSX], b = pg[r * BSX + M];
    faces[r] = a;
    faces[PS + r] = b;
    const int z = r / BSX, y = r % BSX;
    if (y == 0 || y == M) {
      const int k = y ? 2 : 0;
      edges[k * BSX + z] = a;
      edges[(k + 1) * BSX + z] = b;
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
};

// Voxel column per thread (PX x PY): the fast default is 4 x 2 for 32^3
// blocks (128 threads), for 16^3 blocks in launches of at least 4096 (32
// threads, two blocks per wave) and 2 x 2 otherwise; exact arithmetic keeps
// 2 x 2 below 32^3. Measured on MI355X (FP16 pages, 32^3
// blocks): 4x2 with one plane of prefetch ran 10% faster than 2x2; 2x4,
// 4x4, 8x2 (register pressure) and a two-plane prefetch (lower occupancy)
// were slower. At 16^3, 4x2 ran 6% faster than 2x2. Batches pick per block
// size with an autotuner.
template <int BSX, int PX, int PY>
struct Tile {
  static constexpr int NT = (BSX / PX) * (BSX / PY);
};

// Blocks per workgroup: a block with fewer columns than a wave (8^3 blocks
// have 16 with 2x2 columns) shares its workgroup with others, each group of
// NT threads working on its own item with its own LDS planes.
template <int NT>
struct Pack {
  static constexpr int BPW = NT >= 64 ? 1 : 64 / NT;
};

// OP: the stencil (OP_MCF is the hot path); HARD: hard constraints around
// the field at entry (G.phi0); iter: the sweep step (Taubin's parity; the
// UNORM8 store's stochastic rounding draws from it);
// nItems: the number of items of this launch (the last workgroup may hold
// fewer than BPW); C01 (fast path only): every grid of the launch clamps to
// [0,1] (doConstrZeroOne), which the last fma of the update does with its
// clamp bit; RES: the launch may accumulate the FP64 residual (partial !=
// nullptr). Launches without a residual take RES = false, which leaves the
// residual's conversions and branches out of the z-loop.
//
// The fast MCF path (OP_MCF, not EXACT, not HARD, FP32 arithmetic) on FP16
// pages keeps the planes in LDS and the patches in registers as FP16 bits
// (LDS rows padded to a multiple of four elements so that a patch row of a
// 4-wide column is one 8-byte and one 4-byte read): half the registers of
// decoded patches, which keeps three waves per SIMD, at the price of sums and
// differences in v_fma_mix_f32 instead of plain FP32 adds.
// One colour pass of one block (item: grid, page) by NT threads (tid), the
// sub-th group of BPW in its workgroup; inactive groups (past the end of the
// item list) work on a valid item but store nothing.
template <int BSX, int ST, bool EXACT, int PX, int PY, int OP, bool HARD, int BPW, bool C01>
__device__ __forceinline__ void sweep_block(const SweepGrid *__restrict__ grids, const int2 item,
                                            double *__restrict__ const partial, const int iter,
                                            const int itemIdx, const bool active, const int sub,
                                            const int tid) {
  using S = Stor<ST>;
  using R = typename S::raw;
  using C = typename S::ct;
  constexpr bool FAST = OP == OP_MCF && !EXACT && !HARD && sizeof(C) == 4;
  constexpr bool H16 = FAST && ST == MSBG_STORAGE_F16;
  using L = std::conditional_t<H16, _Float16, C>;  // element of the LDS planes and patches
  using PT = Patch<L, PX, PY>;
  constexpr int NT = Tile<BSX, PX, PY>::NT;
  constexpr int W = PX + 2;          // patch row stride
  constexpr int P = H16 ? BSX + 4 : BSX + 2;  // LDS row pitch
  constexpr int PL = (BSX + 2) * P;  // LDS plane size
  constexpr int PS = BSX * BSX;
  constexpr int64_t PE = (int64_t)BSX * BSX * BSX;
  constexpr int FS = 2 * PS, ES = 4 * BSX;
  constexpr int NR = 4 * BSX + 4;  // ring cells of one halo plane
  constexpr int RPT = (NR + NT - 1) / NT;
  constexpr int LPT = PS / (4 * NT);  // 4-wide interior chunks per thread and plane
  constexpr int M = BSX - 1;
  static_assert(LPT * 4 * NT == PS, "plane does not split into 4-wide chunks");
  static_assert(!H16 || PX % 2 == 0, "FP16 patches hold pairs of x");

  static_assert(BPW == 1 || BPW * NT <= 64, "packed blocks must share one wave");
  const SweepGrid G = grids[item.x];
  const int32_t *__restrict__ nbrTab = G.nbr;
  uint8_t *pagesRaw = G.pages;
  uint8_t *sideRaw = G.side;
  const uint8_t *constsRaw = G.consts;
  const int64_t npages = G.npages;
  const float dtf = G.dt, lo01 = G.lo, hi01 = G.hi;

  // 8^3 blocks stage all BSX + 2 planes at once (WHOLE); larger blocks march
  // z through a ring of three planes
  constexpr bool WHOLE = BSX <= 8;
  constexpr int NBUF = WHOLE ? BSX + 2 : 3;
  __shared__ __align__(16) L bufs[BPW][NBUF][PL];
  __shared__ int codesAll[BPW][27];
  __shared__ double red[(NT + 63) / 64];
  L(&buf)[NBUF][PL] = bufs[sub];
  int *codes = codesAll[sub];

  const int page = item.y;
  // 27 codes; small blocks have fewer than 27 threads
  for (int k = tid; k < 27; k += NT) codes[k] = ldg(nbrTab + (int64_t)page * 27 + k);
  __syncthreads();

  R *pages = (R *)pagesRaw;
  R *own = pages + (int64_t)page * PE;
  R *faces = (R *)sideRaw;                 // [npages][2][z][y]
  R *edges = faces + npages * FS;          // [npages][4][z]
  R *ownFaces = faces + (int64_t)page * FS;
  R *ownEdges = edges + (int64_t)page * ES;
  const R *consts = (const R *)constsRaw;  // [0..3] empty, [4..7] full
  const C dt = (C)dtf;
  // destination of the new values (== the source in the in-place modes)
  R *dOwn = (R *)G.dst_pages + (int64_t)page * PE;
  R *dFaces = (R *)G.dst_side + (int64_t)page * FS;
  R *dEdges = (R *)G.dst_side + npages * FS + (int64_t)page * ES;

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
    // source pointer of the cell (xx, yy) of plane zz (each in [-1, BSX]) in
    // the block with code c, or nullptr for NULL and outside-the-domain codes
    auto resolve = [&](int c, int xx, int yy, int zz, int *stride) -> const R * {
      if (c == MSBG_BLOCK_EMPTY) {
        *stride = 0;
        return consts;
      }
      if (c == MSBG_BLOCK_FULL) {
        *stride = 0;
        return consts + 4;
      }
      if (c < 0) return nullptr;
      const bool xs = xx < 0 || xx > M, ys = yy < 0 || yy > M;
      const int zn = zz & M;  // plane inside that block
      if (xs && ys) {  // corner: the neighbour's opposite z-edge
        const int k = (xx < 0 ? 1 : 0) + (yy < 0 ? 2 : 0);
        *stride = 1;
        return edges + (int64_t)c * ES + k * BSX + zn;
      }
      if (xs) {  // x-face: the neighbour's opposite face, [z][y]
        *stride = BSX;
        return faces + (int64_t)c * FS + (xx < 0 ? 1 : 0) * PS + zn * BSX + yy;
      }
      *stride = PS;  // a row of that block's page
      return pages + (int64_t)c * PE + zn * PS + (yy & M) * BSX + (xx & M);
    };
    // source pointer for the neighbour at (dxi, dyi, dzi) and plane zz
    auto one = [&](int dzi, int zz, int *stride) -> const R * {
      int c = codes[dzi * 9 + dyi * 3 + dxi];
      const R *src;
      if (c == MSBG_BLOCK_OUT && G.outClip) {
        // getOutOfBlockValue with OPT_BC_COARSE_LEVEL (halo.h:288-300):
        // coordinates outside the domain are clipped to it first, which
        // lands in the block itself or an in-domain neighbour
        int xx = x, yy = y, z2 = zz, dx2 = dxi, dy2 = dyi, dz2 = dzi;
        if (dxi != 1 && codes[12 + dxi] == MSBG_BLOCK_OUT) {
          dx2 = 1;
          xx = x < 0 ? 0 : M;
        }
        if (dyi != 1 && codes[10 + 3 * dyi] == MSBG_BLOCK_OUT) {
          dy2 = 1;
          yy = y < 0 ? 0 : M;
        }
        if (dzi != 1 && codes[dzi * 9 + 4] == MSBG_BLOCK_OUT) {
          dz2 = 1;
          z2 = zz < 0 ? 0 : M;
        }
        src = resolve(codes[dz2 * 9 + dy2 * 3 + dx2], xx, yy, z2, stride);
      } else {
        src = resolve(c, x, y, zz, stride);
      }
      if (src) return src;
      // NULL block or outside the domain: the block's own clamped voxel
      const int zc = zz < 0 ? 0 : (zz > M ? M : zz);
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
  // ring sources of the next interior plane, advanced by one plane per load
  const R *rcur[RPT];
#pragma unroll
  for (int q = 0; q < RPT; q++) rcur[q] = rs[q].p0;

  // one plane's values held between its load and its LDS store: decoded,
  // or the FP16 bits (two per register) on the FP16 fast path
  struct Staged {
    std::conditional_t<H16, uint32_t, C> in[H16 ? LPT * 2 : LPT * 4];
    std::conditional_t<H16, uint16_t, C> ring[RPT];
  };
  auto dec = [](R r) -> C { return (C)S::dec(r); };

  // chunk k of this thread: 4 consecutive x of one row, e = y * BSX + x
  auto chunk_off = [&](int k) -> int { return (tid + k * NT) * 4; };

  // plane zz of the block (0 <= zz < BSX, loaded in order): page rows from
  // the block's own page at a uniform base, ring cells at their cursors
  auto load_inner = [&](int zz, Staged &st) {
#pragma unroll
    for (int k = 0; k < LPT; k++) {
      const R *src = own + zz * PS + chunk_off(k);
      if constexpr (H16) {
        const uint64_t u = *(const MSBG_GLOBAL uint64_t *)src;
        st.in[2 * k] = (uint32_t)u;
        st.in[2 * k + 1] = (uint32_t)(u >> 32);
      } else {
        R r[4];
        load4(src, r);
#pragma unroll
        for (int u = 0; u < 4; u++) st.in[k * 4 + u] = dec(r[u]);
      }
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      if constexpr (H16)
        st.ring[q] = ldg(rcur[q]);
      else
        st.ring[q] = dec(ldg(rcur[q]));
      rcur[q] += rs[q].stride;
    }
  };
  // the plane below (top = false) or above (top = true) the block
  auto load_outer = [&](bool top, Staged &st) {
    const R *base = top ? zsrcP : zsrcM;
    const bool cst = top ? constP : constM;
#pragma unroll
    for (int k = 0; k < LPT; k++) {
      const R *src = base + (cst ? 0 : chunk_off(k));
      if constexpr (H16) {
        const uint64_t u = *(const MSBG_GLOBAL uint64_t *)src;
        st.in[2 * k] = (uint32_t)u;
        st.in[2 * k + 1] = (uint32_t)(u >> 32);
      } else {
        R r[4];
        load4(src, r);
#pragma unroll
        for (int u = 0; u < 4; u++) st.in[k * 4 + u] = dec(r[u]);
      }
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const R *src = top ? rs[q].p : rs[q].m;
      if constexpr (H16)
        st.ring[q] = ldg(src);
      else
        st.ring[q] = dec(ldg(src));
    }
  };

  auto put_plane = [&](L *b, const Staged &st) {
#pragma unroll
    for (int k = 0; k < LPT; k++) {
      const int e = chunk_off(k);
      L *row = b + (e / BSX + 1) * P + (e % BSX) + 1;
      if constexpr (H16) {
        // x..x+3 land at the odd LDS columns x+1..x+4: one 2-byte, one
        // 4-byte and one 2-byte store
        const uint32_t w0 = st.in[2 * k], w1 = st.in[2 * k + 1];
        *(u16a *)row = (uint16_t)w0;
        *(u32a *)(row + 1) = __builtin_amdgcn_alignbit(w1, w0, 16);
        *(u16a *)(row + 3) = (uint16_t)(w1 >> 16);
      } else {
#pragma unroll
        for (int u = 0; u < 4; u++) row[u] = st.in[k * 4 + u];
      }
    }
#pragma unroll
    for (int q = 0; q < RPT; q++)
      if (tid + q * NT < NR) {
        if constexpr (H16)
          *(u16a *)(b + rlds[q]) = st.ring[q];
        else
          b[rlds[q]] = st.ring[q];
      }
  };

  auto get_patch = [&](const L *b, PT &p) {
#pragma unroll
    for (int j = 0; j < PY + 2; j++) {
      const L *row = b + (y0 + j) * P + x0;
      if constexpr (H16) {
        if constexpr (PX % 4 == 0) {  // 8-byte aligned rows
#pragma unroll
          for (int k = 0; k + 1 < PT::WW; k += 2) {
            const uint64_t q = *(const u64a *)(row + 2 * k);
            p.w[j][k] = (uint32_t)q;
            p.w[j][k + 1] = (uint32_t)(q >> 32);
          }
          if constexpr (PT::WW % 2) p.w[j][PT::WW - 1] = *(const u32a *)(row + 2 * (PT::WW - 1));
        } else {
#pragma unroll
          for (int k = 0; k < PT::WW; k++) p.w[j][k] = *(const u32a *)(row + 2 * k);
        }
      } else {
#pragma unroll
        for (int i = 0; i < PX + 2; i++) p.v[j * W + i] = row[i];
      }
    }
  };

  double acc = 0.0;
  const bool wFace0 = x0 == 0, wFace1 = x0 == BSX - PX, wY0 = y0 == 0, wY1 = y0 == BSX - PY;
  const float lim = FAST ? 0.1f / dtf : 0.0f;  // limiter bound of the fast path, in units of dt

  // encode and store the new values of plane z, keep the side-cars equal to
  // the page, accumulate the residual
  auto store = [&](int z, const C (&o)[PX * PY], const PT &Mp) {
    R e[PX * PY];
    if constexpr (H16) {
#pragma unroll
      for (int k = 0; k < PX * PY; k += 2) {  // v_cvt_pk_f16_f32
        const uint32_t pk =
            __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{o[k], o[k + 1]}, h2x_t));
        e[k] = (R)pk;
        e[k + 1] = (R)(pk >> 16);
      }
    } else if constexpr (ST == MSBG_STORAGE_U8) {
#pragma unroll
      for (int b = 0; b < PY; b++)
#pragma unroll
        for (int a = 0; a < PX; a++) {
          const uint32_t vid = (uint32_t)((z * BSX + y0 + b) * BSX + x0 + a);
          e[b * PX + a] = u8_from_float_sr(o[b * PX + a], sr_uniform((uint32_t)iter, (uint32_t)page, vid));
        }
    } else {
#pragma unroll
      for (int k = 0; k < PX * PY; k++) e[k] = S::enc(o[k]);
    }
#pragma unroll
    for (int b = 0; b < PY; b++)
#pragma unroll
      for (int a = 0; a < PX; a += 2)
        store2(dOwn + z * PS + (y0 + b) * BSX + x0 + a, e[b * PX + a], e[b * PX + a + 1]);
    if (wFace0) {
#pragma unroll
      for (int b = 0; b < PY; b++) stg(dFaces + z * BSX + y0 + b, e[b * PX]);
    }
    if (wFace1) {
#pragma unroll
      for (int b = 0; b < PY; b++) stg(dFaces + PS + z * BSX + y0 + b, e[b * PX + PX - 1]);
    }
    if (wFace0 && wY0) stg(dEdges + 0 * BSX + z, e[0]);
    if (wFace1 && wY0) stg(dEdges + 1 * BSX + z, e[PX - 1]);
    if (wFace0 && wY1) stg(dEdges + 2 * BSX + z, e[(PY - 1) * PX]);
    if (wFace1 && wY1) stg(dEdges + 3 * BSX + z, e[PY * PX - 1]);
    if (partial) {
#pragma unroll
      for (int b = 0; b < PY; b++)
#pragma unroll
        for (int a = 0; a < PX; a++) {
          double d = (double)S::dec(e[b * PX + a]) - (double)(C)Mp.at(b + 1, a + 1);
          acc += d * d;
        }
    }
  };

  // plane z's update from the patches of planes z-1, z, z+1
  auto compute = [&](int z, const PT &Dp, const PT &Mp, const PT &Up) {
    if (BPW > 1 && !active) return;
    C o[PX * PY];
    if constexpr (FAST) {
      mcf_fast<PX, PY, C01>(Dp, Mp, Up, dtf, lim, lo01, hi01, o);
    } else if constexpr (OP == OP_MCF) {
      const bool c01x = !HARD && lo01 == 0.0f;
#pragma unroll
      for (int b = 0; b < PY; b++)
#pragma unroll
        for (int a = 0; a < PX; a++)
          o[b * PX + a] = mcf_exact<C, W>(Dp.v, Mp.v, Up.v, (b + 1) * W + a + 1, dt, c01x);
    } else {
      static_assert(sizeof(C) == 4, "the 7-point stencils run in FP32");
#pragma unroll
      for (int b = 0; b < PY; b++)
#pragma unroll
        for (int a = 0; a < PX; a++) {
          float ph = pde7_exact<OP, W>(Dp.v, Mp.v, Up.v, (b + 1) * W + a + 1, dtf, G.taubin,
                                       G.alpha, iter);
          if (!HARD && lo01 == 0.0f) {  // doConstrZeroOne: min(max(phi,0),1)
            ph = ph > 0.0f ? ph : 0.0f;
            ph = ph < 1.0f ? ph : 1.0f;
          }
          o[b * PX + a] = ph;
        }
    }
    if constexpr (HARD) {  // msbg4.cpp:813-818
      const R *p0 = (const R *)G.phi0 + (int64_t)page * PE + z * PS;
#pragma unroll
      for (int b = 0; b < PY; b++)
#pragma unroll
        for (int a = 0; a < PX; a++) {
          const C f0 = (C)S::dec(ldg(p0 + (y0 + b) * BSX + x0 + a));
          C lo = f0 - (C)G.tol1, hi = f0 + (C)G.tol2;
          lo = (C)0 > lo ? (C)0 : lo;  // Vec8f max(0, phi0 - tol1)
          hi = (C)1 < hi ? (C)1 : hi;  // Vec8f min(1, phi0 + tol2)
          C ph = o[b * PX + a];
          ph = ph < hi ? ph : hi;
          ph = ph > lo ? ph : lo;
          o[b * PX + a] = ph;
        }
    }
    store(z, o, Mp);
  };

  PT pA, pB, pC;
  if constexpr (WHOLE) {
    // An 8^3 block's planes -1..BSX are loaded together, so that the block
    // waits for one load latency instead of one per few planes, then stored
    // to LDS; every plane is computed from LDS after one barrier (the same
    // values as the ring: the block's planes as they were at the start)
    Staged st[BSX + 2];
    load_outer(false, st[0]);
#pragma unroll
    for (int z = 0; z < BSX; z++) load_inner(z, st[z + 1]);
    load_outer(true, st[BSX + 1]);
#pragma unroll
    for (int q = 0; q < BSX + 2; q++) put_plane(buf[q], st[q]);
    __syncthreads();
    get_patch(buf[0], pA);
    get_patch(buf[1], pB);
#pragma unroll
    for (int z = 0; z < BSX; z += 3) {
      get_patch(buf[z + 2], pC);
      compute(z, pA, pB, pC);
      if (z + 1 < BSX) {
        get_patch(buf[z + 3], pA);
        compute(z + 1, pB, pC, pA);
      }
      if (z + 2 < BSX) {
        get_patch(buf[z + 4], pB);
        compute(z + 2, pC, pA, pB);
      }
    }
  } else {
    // Plane p of the march (-1..BSX) is loaded PFD steps before its LDS store
    // and staged one step before its patch is read: the patch reads of a step
    // go to a plane stored by the previous step and are issued before this
    // step's stores, so they do not queue behind them. PFD is three for the
    // FP16 kernels of one wave per workgroup (registers to spare at two waves
    // per SIMD; a step is about as long as a loaded HBM round trip), one
    // otherwise. Plane p is held in stg[(p + 3) % 3] between its load and its
    // LDS store and staged in buf[(p + 3) % 3], so that every step of the
    // three-step loop below has fixed registers and LDS offsets.
    constexpr int PFD = (H16 && NT * BPW >= 64) ? 3 : 1;
    static_assert(PFD >= 1 && PFD <= 3, "prefetch distance");
    Staged stg[3];

    // prologue: planes -1, 0, 1 staged, 2..1+PFD in flight, patches of -1, 0
    load_outer(false, stg[2]);
    load_inner(0, stg[0]);
    load_inner(1, stg[1]);
    put_plane(buf[2], stg[2]);
    load_inner(2, stg[2]);
    put_plane(buf[0], stg[0]);
    if (PFD >= 2) load_inner(3, stg[0]);
    put_plane(buf[1], stg[1]);
    if (PFD >= 3) load_inner(4, stg[1]);
    __syncthreads();
    get_patch(buf[2], pA);
    get_patch(buf[0], pB);
    // step 0 stages plane 2 over plane -1: every wave's patch of plane -1
    // must be read first (later steps overwrite planes read two barriers back)
    __syncthreads();

    // step z: read the patch of plane z+1 (staged in buf[nb]), stage plane
    // z+2 (ps) and prefetch plane z+2+PFD into its registers (ld 1: a plane of
    // the block, 2: the plane above it, 0: none), compute plane z. Plane z+2
    // overwrites plane z-1, whose patch was read two barriers earlier.
    auto step = [&](int z, int nb, bool ps, int ld, const PT &Dp, const PT &Mp, PT &Up) {
      const int ns = (nb + 1) % 3;  // slot and buffer of plane z+2
      get_patch(buf[nb], Up);
      if (ps) put_plane(buf[ns], stg[ns]);
      if (ld == 1)
        load_inner(z + 2 + PFD, stg[(ns + PFD) % 3]);
      else if (ld == 2)
        load_outer(true, stg[(ns + PFD) % 3]);
      __syncthreads();
      compute(z, Dp, Mp, Up);
    };

    // steps whose prefetch is a plane of the block
    constexpr int ZM = ((BSX - 2 - PFD) / 3) * 3;
    for (int z = 0; z < ZM; z += 3) {
      step(z, 1, true, 1, pA, pB, pC);
      step(z + 1, 2, true, 1, pB, pC, pA);
      step(z + 2, 0, true, 1, pC, pA, pB);
    }
  #pragma unroll
    for (int t = 0; t < BSX - ZM; t++) {  // the last 3..6 steps
      const int z = ZM + t;
      const int pl = z + 2 + PFD;
      const int ld = pl < BSX ? 1 : (pl == BSX ? 2 : 0);
      const bool ps = z + 2 <= BSX;
      if (t % 3 == 0)
        step(z, 1, ps, ld, pA, pB, pC);
      else if (t % 3 == 1)
        step(z, 2, ps, ld, pB, pC, pA);
      else
        step(z, 0, ps, ld, pC, pA, pB);
    }
  }

  if (partial) {
    // deterministic block reduction: wave shuffles, then waves in order
    constexpr int WV = NT < 64 ? NT : 64;
    for (int off = WV / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, WV);
    if constexpr (BPW > 1) {
      if (tid == 0 && active) partial[itemIdx] = acc;  // one group per item
    } else {
      if ((tid & 63) == 0) red[tid >> 6] = acc;
      __syncthreads();
      if (tid == 0) {
        double sum = 0.0;
        for (int w = 0; w < (NT + 63) / 64; w++) sum += red[w];
        partial[itemIdx] = sum;
      }
    }
  }
}

template <int BSX, int ST, bool EXACT, int PX, int PY, int OP = OP_MCF, bool HARD = false,
          int BPW = 1, bool C01 = false, bool RES = true>
__global__ __launch_bounds__((Tile<BSX, PX, PY>::NT * BPW)) void k_sweep(
    const SweepGrid *__restrict__ grids, const int2 *__restrict__ items,
    double *__restrict__ partialArg, int iter, int nItems) {
  constexpr int NT = Tile<BSX, PX, PY>::NT;
  // work item: which grid, which page of it. A group past the end of the
  // list works on the last item without storing anything, so that it still
  // reaches every barrier.
  const int sub = BPW > 1 ? (int)threadIdx.x / NT : 0;
  const int tid = BPW > 1 ? (int)threadIdx.x % NT : (int)threadIdx.x;
  const int itemIdx = (int)blockIdx.x * BPW + sub;
  const bool active = itemIdx < nItems;
  const int2 item = items[active ? itemIdx : nItems - 1];
  sweep_block<BSX, ST, EXACT, PX, PY, OP, HARD, BPW, C01>(grids, item, RES ? partialArg : nullptr,
                                                          iter, itemIdx, active, sub, tid);
}

// ---- the persistent batched pass -------------------------------------------
// The work queue of a batch group (BatchQueue) holds every (sweep, colour,
// grid, block) item of the run in that order. Item i is sweep s = i / N8
// (N8 = items per sweep), entry r = i - s N8 of the colour-sorted item list,
// whose colour c satisfies off[c] <= r < off[c+1].

// Workgroups (one wave each) take runs of kQueueRun items in queue order.
// Before sweeping a block of pass (s, c) of grid g, a workgroup waits until
// every block of g's earlier passes is done: done[g] >= s * pre[g][8] +
// pre[g][c] (the reference's colour order per grid, msbg4.cpp:384-900; grids
// are independent, so one grid's next pass overlaps the others' current
// one). Every item a wait depends on sits earlier in the queue, in an earlier
// run or published before the waiting item within its own run, and was taken
// by a running workgroup whose own waits point further back, so the waits end;
// each is still bounded by waitTicks of wall-clock time, past which it sets
// *fail and goes on (the run is then reported as failed, and no workgroup
// can hang). Waits poll with a sleep between polls, one 64-byte line per
// grid's counter. Stores are published with an agent-scope release before
// the count and read after an agent-scope acquire.
//
// The loop has no lane-dependent branch: every lane executes each atomic,
// adding 1 on lane 0 and 0 elsewhere, and all control (the item taken, the
// loop's exit, the wait) goes through readfirstlane. A lane-0-only region
// here lets the compiler thread lane 0's path from one iteration's count to
// the next iteration's fetch past the other lanes, which then loop on a
// stale item.
template <int BSX, int ST, bool EXACT, int PX, int PY, bool C01>
__global__ __launch_bounds__(64) void k_sweep_queue(const SweepGrid *__restrict__ grids,
                                                    const BatchQueue q) {
  static_assert(Tile<BSX, PX, PY>::NT == 64, "one wave per workgroup");
  const bool lane0 = threadIdx.x == 0;
  const unsigned n8 = (unsigned)q.off[8];
  const unsigned total = n8 * (unsigned)q.sweeps;
  // A workgroup takes kQueueRun consecutive items at a time. Its finished
  // items of one grid and pass are published together (one release, one
  // count), before an item of another grid or pass and at the run's end, so
  // no item waits on an unpublished item of its own run. An item whose grid
  // count this workgroup has already waited for and acquired skips the wait.
  int ae = -1;
  unsigned aneed = 0;
  for (;;) {
    const unsigned base =
        __builtin_amdgcn_readfirstlane(atomicAdd(q.head, lane0 ? kQueueRun : 0u));
    if (base >= total) break;
    const unsigned end = base + kQueueRun < total ? base + kQueueRun : total;
    int pe = -1, ps = -1, pc = -1;
    unsigned pn = 0;
    auto publish = [&]() {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_fetch_add(q.done + (int64_t)pe * kDoneStride, lane0 ? pn : 0u,
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      pn = 0;
    };
    for (unsigned i = base; i < end; i++) {
      const int s = (int)(i / n8), r = (int)(i - (unsigned)s * n8);
      int c = 0;
      while (c < 7 && r >= q.off[c + 1]) c++;
      const int e = __builtin_amdgcn_readfirstlane(q.items[r].x);
      const int page = __builtin_amdgcn_readfirstlane(q.items[r].y);
      if (pn && (e != pe || s != ps || c != pc)) publish();
      const int *pre = q.pre + (int64_t)e * 9;
      const unsigned need = __builtin_amdgcn_readfirstlane((unsigned)s * (unsigned)pre[8] +
                                                           (unsigned)pre[c]);
      if (e != ae || need > aneed) {
        unsigned *const done = q.done + (int64_t)e * kDoneStride;
        const uint64_t t0 = wall_clock64();
        for (;;) {
          const unsigned d = __builtin_amdgcn_readfirstlane(
              __hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
          if (d >= need) break;
          if ((uint64_t)wall_clock64() - t0 > q.waitTicks) {
            atomicOr(q.fail, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(16);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ae = e;
        aneed = need;
      }
      // the lane index through an opaque move, so that the block's per-lane
      // set-up is computed per item instead of being hoisted out of the loop
      // and held across it (which costs registers the z-march needs)
      int lane;
      asm volatile("v_mov_b32 %0, %1" : "=v"(lane) : "v"((int)threadIdx.x));
      sweep_block<BSX, ST, EXACT, PX, PY, OP_MCF, false, 1, C01>(
          grids, make_int2(e, page), nullptr, s, 0, true, 0, lane);
      pe = e;
      ps = s;
      pc = c;
      pn++;
    }
    if (pn) publish();
  }
}

}  // namespace sweep
}  // namespace msbg
