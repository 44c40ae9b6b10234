// Kernels of the adaptive pressure relax (msbg3.cpp).
//
// k_ps_block: one block per BSX*BSX threads (several small blocks share a
// 256-thread workgroup). The halo block of IprocessBlockLaplacian
// (msbg3.cpp:530-1030) is built in LDS, premultiplied by h, together with the
// diagonal, the right-hand side and the per-vector gates. The relax then runs
// the reference's _mgSmBlockIter sub-iterations (msbg3.cpp:1100-1470):
//   - every sub-iteration but the last is a lexicographic sweep over SW-wide
//     vectors (SW = 8 for 16^3 and 8^3 blocks, = BSX below), forward or
//     backward by the sub-iteration's parity. Inside a vector the update is
//     Jacobi (all lanes read the old values of their in-vector neighbours);
//     across vectors it is Gauss-Seidel. On the GPU one thread owns one (x,y)
//     column and the sweep runs as a wavefront: vector (i,y,z) is updated at
//     step i+y+z (mirrored for the backward order). Its earlier neighbours
//     (i-1,y,z), (i,y-1,z), (i,y,z-1) were written at the step before, its
//     later ones are written at the step after, so one barrier per step keeps
//     the sequential order exactly; the in-vector neighbours come from the old
//     values by cross-lane shuffles.
//   - the last sub-iteration reads the buffer without writing it (Jacobi) and
//     stores Y/h to the destination page.
// The multiply (msbg3.cpp:1271-1320) is one read-only pass of the same kind.
//
// Every operation rounds as in the CPU oracle (no contraction; fmaf where the
// oracle has it), so the result is bit-exact.
#include <hip/hip_runtime.h>

#include "internal.h"
#include "ps.h"

namespace msbg {
namespace {

__device__ __forceinline__ int vox(int bsx, int x, int y, int z) { return x + bsx * (y + bsx * z); }

// sum of a*x over the coarse-fine face cells of procNeighMixedRes2
// (msbg3.cpp:59-160), with a = 1/3 (finer neighbour, four cells) or 2/3
// (coarser neighbour), times the face area when the block cuts a solid
template <int BSX>
__device__ double mixed_sum(const PsBlockArgs &a, const float *X2, int L, int L2, int b, int b2,
                            int d, int s, int vx, int vy, int vz, bool cuts) {
#pragma clang fp contract(off)
  double sum = 0;
  if (L2 < L) {
    constexpr int BH = 2 * BSX;
    const float *fa = a.fa[L2][d] + (int64_t)(s ? b2 : b) * BH * BH * BH;
    const float a0 = (float)(1.0 * (1. / (4 * (3. / 4.))));
#pragma unroll
    for (int k = 0; k < 4; k++) {
      int p = k & 1, q = k >> 1, x2, y2, z2, fx, fy, fz;
      if (d == 0) { x2 = s ? 0 : BH - 1; y2 = 2 * vy + p; z2 = 2 * vz + q; fx = 0; fy = y2; fz = z2; }
      else if (d == 1) { y2 = s ? 0 : BH - 1; x2 = 2 * vx + p; z2 = 2 * vz + q; fx = x2; fy = 0; fz = z2; }
      else { z2 = s ? 0 : BH - 1; x2 = 2 * vx + p; y2 = 2 * vy + q; fx = x2; fy = y2; fz = 0; }
      float c = a0;
      if (cuts) c = c * fa[vox(BH, fx, fy, fz)];
      sum += (double)(c * X2[vox(BH, x2, y2, z2)]);
    }
  } else {
    constexpr int BL = BSX / 2 > 0 ? BSX / 2 : 1;
    int x2, y2, z2, fx = vx, fy = vy, fz = vz;
    if (d == 0) { x2 = s ? 0 : BL - 1; y2 = vy / 2; z2 = vz / 2; if (s) fx = 0; }
    else if (d == 1) { y2 = s ? 0 : BL - 1; x2 = vx / 2; z2 = vz / 2; if (s) fy = 0; }
    else { z2 = s ? 0 : BL - 1; x2 = vx / 2; y2 = vy / 2; if (s) fz = 0; }
    const float *fa = a.fa[L][d] + (int64_t)(s ? b2 : b) * BSX * BSX * BSX;
    float c = (float)(1.0 * 2 * (1. / (4 * (3. / 4.))));
    if (cuts) c = c * fa[vox(BSX, fx, fy, fz)];
    sum += (double)(c * X2[vox(BL, x2, y2, z2)]);
  }
  return sum;
}

// N consecutive floats / uint16 as one aligned vector load
template <int N>
struct alignas(4 * N) F32v {
  float v[N];
};
template <int N>
struct alignas(2 * N) U16v {
  uint16_t v[N];
};
template <int N>
__device__ __forceinline__ F32v<N> load_f32v(const float *p) {
  return *reinterpret_cast<const F32v<N> *>(__builtin_assume_aligned(p, 4 * N));
}
template <int N>
__device__ __forceinline__ U16v<N> load_u16v(const uint16_t *p) {
  return *reinterpret_cast<const U16v<N> *>(__builtin_assume_aligned(p, 2 * N));
}

// the old value of the lane below / above inside a 16-lane row
__device__ __forceinline__ float lane_below(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_above(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x101, 0xf, 0xf, true));
}

// face weights of the six faces, stored per cell as k/64 in one byte each:
// every face area of a block level (at most 3 halvings of a 0/1 face) is a
// multiple of 1/64, so the encoding is exact
__device__ __forceinline__ uint32_t w_code(float w) { return (uint32_t)(w * 64.0f); }
__device__ __forceinline__ float w_decode(uint64_t c, int f) {
  return (float)(uint32_t)((c >> (8 * f)) & 0xffu) * (1.0f / 64.0f);
}

template <int BSX, int MODE, bool CUTS>
__global__ __launch_bounds__(256) void k_ps_block(PsBlockArgs a, const PsItem *items, int n) {
#pragma clang fp contract(off)
  constexpr int SW = BSX >= 8 ? 8 : BSX;
  constexpr int NV = BSX / SW;  // vectors per row = vectors per thread in the set-up
  constexpr int TPB = BSX * BSX;
  constexpr int NB = 256 / TPB;
  constexpr int HS = BSX + 2;
  constexpr int NC = BSX * BSX * BSX;
  constexpr int NVEC = NC / SW;
  __shared__ float s_buf[NB][HS * HS * HS];
  __shared__ float s_D[NB][NC];
  __shared__ float s_B[NB][MODE == PS_MODE_MULTIPLY ? 1 : NC];
  __shared__ uint8_t s_g[NB][NVEC];
  __shared__ uint64_t s_w[CUTS ? NB : 1][CUTS ? NC : 1];
  __shared__ double s_red[MODE == PS_MODE_MULTIPLY ? 256 : 1];

  const int slot = threadIdx.x / TPB, t = threadIdx.x % TPB;
  const int x = t % BSX, y = t / BSX;
  const int item = blockIdx.x * NB + slot;
  const bool valid = item < n;
  PsItem it0;
  if (valid) {
    it0 = items[item];
  } else {
    it0.off = 0, it0.b = 0, it0.L = 0;
#pragma unroll
    for (int f = 0; f < 6; f++) it0.nbr[f] = -1, it0.noff[f] = 0, it0.nL[f] = 0;
  }
  const int b = it0.b;
  const int L = it0.L;
  const float h = (float)((double)(1 << L) * a.dx0);
  const float HR = 1.f / h;
  const int64_t off = it0.off;
  float *buf = s_buf[slot];
  auto H = [](int xx, int yy, int zz) { return (xx + 1) + HS * ((yy + 1) + HS * (zz + 1)); };

  int nb2[6], kind[6];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const int b2 = it0.nbr[f];
    nb2[f] = b2;
    kind[f] = b2 < 0 ? 1 : (it0.nL[f] == L ? 0 : 2);
  }

  // interior, whole vectors per thread: gates, X*h (0 in vectors without
  // fluid), D, B and, for blocks that cut a solid, the face weights
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const int vec = t + k * TPB;
    const int vid0 = vec * SW;
    const int x0 = vid0 % BSX, yy = (vid0 / BSX) % BSX, zz = vid0 / (BSX * BSX);
    U16v<SW> f;
    F32v<SW> X, D, B;
    if (valid) {
      f = load_u16v<SW>(a.flags + off + vid0);
      X = load_f32v<SW>(a.X + off + vid0);
      D = load_f32v<SW>(a.D + off + vid0);
      if constexpr (MODE != PS_MODE_MULTIPLY) B = load_f32v<SW>(a.B + off + vid0);
    } else {
#pragma unroll
      for (int l = 0; l < SW; l++) f.v[l] = kPsCellSolid, X.v[l] = D.v[l] = B.v[l] = 0.0f;
    }
    int any = 0, fl = 0;
#pragma unroll
    for (int l = 0; l < SW; l++) {
      any |= (f.v[l] & (kPsCellSolid | kPsCellPartialSolid)) != 0;
      fl |= (f.v[l] & (kPsCellSolid | kPsCellVoid)) == 0;
    }
    const int gate = any | ((!fl) << 1);
    s_g[slot][vec] = (uint8_t)gate;
#pragma unroll
    for (int l = 0; l < SW; l++) {
      buf[H(x0 + l, yy, zz)] = (gate & 2) ? 0.0f : X.v[l] * h;
      s_D[slot][vid0 + l] = D.v[l];
      if constexpr (MODE != PS_MODE_MULTIPLY) s_B[slot][vid0 + l] = B.v[l];
    }
    if constexpr (CUTS) {
      if (valid) {
        const float *fu = a.fa[L][0] + (int64_t)b * NC, *fv = a.fa[L][1] + (int64_t)b * NC,
                    *fw = a.fa[L][2] + (int64_t)b * NC;
        const F32v<SW> U = load_f32v<SW>(fu + vid0), V = load_f32v<SW>(fv + vid0),
                       W = load_f32v<SW>(fw + vid0);
        const float unext = x0 + SW < BSX ? fu[vid0 + SW] : 0.0f;
        F32v<SW> V1, W1;
        if (yy < BSX - 1) V1 = load_f32v<SW>(fv + vid0 + BSX);
        if (zz < BSX - 1) W1 = load_f32v<SW>(fw + vid0 + BSX * BSX);
        // upper-border weights: the neighbour's lower face, 1/0 for an
        // open/solid domain border, 1 on a mixed-resolution face
        auto border = [&](int d, int p, int q) -> float {
          const int kd = kind[2 * d + 1];
          if (kd == 1) return a.openRight[d] ? 1.0f : 0.0f;
          if (kd == 2) return 1.0f;
          const int c = d == 0 ? vox(BSX, 0, p, q) : d == 1 ? vox(BSX, p, 0, q) : vox(BSX, p, q, 0);
          return a.fa[L][d][(int64_t)nb2[2 * d + 1] * NC + c];
        };
#pragma unroll
        for (int l = 0; l < SW; l++) {
          const int xl = x0 + l;
          const float w0 = (xl == 0 && kind[0] == 2) ? 1.0f : U.v[l];
          const float w1 = xl < BSX - 1 ? (l < SW - 1 ? U.v[l + 1] : unext) : border(0, yy, zz);
          const float w2 = (yy == 0 && kind[2] == 2) ? 1.0f : V.v[l];
          const float w3 = yy < BSX - 1 ? V1.v[l] : border(1, xl, zz);
          const float w4 = (zz == 0 && kind[4] == 2) ? 1.0f : W.v[l];
          const float w5 = zz < BSX - 1 ? W1.v[l] : border(2, xl, yy);
          s_w[slot][vid0 + l] = (uint64_t)w_code(w0) | ((uint64_t)w_code(w1) << 8) |
                                ((uint64_t)w_code(w2) << 16) | ((uint64_t)w_code(w3) << 24) |
                                ((uint64_t)w_code(w4) << 32) | ((uint64_t)w_code(w5) << 40);
        }
      }
    }
  }

  // border slabs: same level (X*h), outside the domain (0) or mixed
  // resolution (procNeighMixedRes2 sum times h); the same-level loads are
  // all issued before any is used
  float fv[6];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const int d = f >> 1, s = f & 1;
    int q[3];
    q[d] = s ? 0 : BSX - 1;
    if (d == 0) { q[1] = x; q[2] = y; }
    else if (d == 1) { q[0] = x; q[2] = y; }
    else { q[0] = x; q[1] = y; }
    fv[f] = kind[f] == 0 ? a.X[it0.noff[f] + vox(BSX, q[0], q[1], q[2])] : 0.0f;
  }
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const int d = f >> 1, s = f & 1;
    int c[3];
    c[d] = s ? BSX - 1 : 0;
    if (d == 0) { c[1] = x; c[2] = y; }
    else if (d == 1) { c[0] = x; c[2] = y; }
    else { c[0] = x; c[1] = y; }
    int o[3] = {c[0], c[1], c[2]};
    o[d] += s ? 1 : -1;
    float val = 0.0f;
    if (kind[f] == 0) {
      val = fv[f] * h;
    } else if (kind[f] == 2) {
      const double sum = mixed_sum<BSX>(a, a.X + it0.noff[f], L, it0.nL[f], b, nb2[f], d, s, c[0],
                                        c[1], c[2], CUTS);
      val = (float)(sum * (double)h);
    }
    buf[H(o[0], o[1], o[2])] = val;
  }
  __syncthreads();

  const float omega = a.omega, om1 = 1.0f - omega;
  // stencil sum: plain, or with the face weights where the vector is gated
  auto stencil = [&](int gate, uint64_t wc, float F1, float F2, float F3, float F4, float F5,
                     float F6) -> float {
    if (CUTS && (gate & 1)) {
      float S = 0.f;
      S = __builtin_fmaf(F1, w_decode(wc, 0), S);
      S = __builtin_fmaf(F2, w_decode(wc, 1), S);
      S = __builtin_fmaf(F3, w_decode(wc, 2), S);
      S = __builtin_fmaf(F4, w_decode(wc, 3), S);
      S = __builtin_fmaf(F5, w_decode(wc, 4), S);
      S = __builtin_fmaf(F6, w_decode(wc, 5), S);
      return S;
    }
    return F1 + F2 + F3 + F4 + F5 + F6;
  };
  auto wcode_at = [&](int v) -> uint64_t {
    if constexpr (CUTS) return s_w[slot][v];
    else return 0;
  };

  if constexpr (MODE == PS_MODE_RELAX) {
    const int i = x / SW, lane = x % SW;
    for (int it = 0; it < a.biter; it++) {
      const bool rev = a.rev == (it & 1);
      if (it < a.biter - 1) {
        // wavefront over the column owned by this thread: the cell k along
        // the sweep direction is updated at step st0 + k; the thread keeps
        // the new value below, the old value and the old value above in
        // registers
        constexpr int T = (NV - 1) + 2 * (BSX - 1) + 1;
        const int dz = rev ? -1 : 1;
        const int zs = rev ? BSX - 1 : 0;
        const int st0 = rev ? (NV - 1 - i) + (BSX - 1 - y) : i + y;
        float prv = buf[H(x, y, zs - dz)];
        float cur = buf[H(x, y, zs)];
        float nxt = buf[H(x, y, zs + dz)];
        int v = vox(BSX, x, y, zs);
        float Bc = s_B[slot][v], Dc = s_D[slot][v];
        int gc = s_g[slot][v / SW];
        uint64_t wc = wcode_at(v);
        for (int st = 0; st < T; st++) {
          const int k = st - st0;
          const int z = zs + dz * k;
          const float lsh = lane_below(cur), rsh = lane_above(cur);
          if (k >= 0 && k < BSX) {
            // every LDS read of the step is issued before the step's write:
            // the neighbours, the old value two cells on and the next cell's
            // coefficients, so that their latencies overlap
            float *p = buf + H(x, y, z);
            const float e1 = p[-1], e2 = p[1], F3 = p[-HS], F4 = p[HS];
            const float nx2 = k + 2 <= BSX ? p[2 * dz * HS * HS] : 0.0f;
            const int vn = k + 1 < BSX ? vox(BSX, x, y, z + dz) : v;
            const float Bn = s_B[slot][vn], Dn = s_D[slot][vn];
            const int gn = s_g[slot][vn / SW];
            const uint64_t wn = wcode_at(vn);
            float Yn = cur;
            if (!(gc & 2)) {
              const float F1 = lane > 0 ? lsh : e1;
              const float F2 = lane < SW - 1 ? rsh : e2;
              const float F5 = rev ? nxt : prv, F6 = rev ? prv : nxt;
              const float S = stencil(gc, wc, F1, F2, F3, F4, F5, F6);
              Yn = __builtin_fmaf(om1, cur, (omega * (Bc + S)) * Dc);
              *p = Yn;
            }
            prv = Yn;
            cur = nxt;
            nxt = nx2;
            v = vn, Bc = Bn, Dc = Dn, gc = gn, wc = wn;
          }
          __syncthreads();
        }
      } else {
        for (int z = 0; z < BSX; z++) {
          const int v = vox(BSX, x, y, z);
          const int gate = s_g[slot][v / SW];
          if (gate & 2) continue;
          const float F0 = buf[H(x, y, z)];
          const float S = stencil(gate, wcode_at(v), buf[H(x - 1, y, z)], buf[H(x + 1, y, z)],
                                  buf[H(x, y - 1, z)], buf[H(x, y + 1, z)], buf[H(x, y, z - 1)],
                                  buf[H(x, y, z + 1)]);
          const float Y = __builtin_fmaf(om1, F0, (omega * (s_B[slot][v] + S)) * s_D[slot][v]);
          if (valid) a.Y[off + v] = Y * HR;
        }
      }
    }
  } else {
    double dot = 0;
    for (int z = 0; z < BSX; z++) {
      const int v = vox(BSX, x, y, z);
      const int gate = s_g[slot][v / SW];
      if (gate & 2) continue;
      const float F0 = buf[H(x, y, z)];
      const float S = stencil(gate, wcode_at(v), buf[H(x - 1, y, z)], buf[H(x + 1, y, z)],
                              buf[H(x, y - 1, z)], buf[H(x, y + 1, z)], buf[H(x, y, z - 1)],
                              buf[H(x, y, z + 1)]);
      const float D = s_D[slot][v];
      float r = D != 0.0f ? (1.f / D) * F0 - S : 0.0f;
      if constexpr (MODE == PS_MODE_RESIDUAL) r = s_B[slot][v] - r;
      else dot += (double)r * (double)(F0 * HR);
      if (valid) a.Y[off + v] = r;
    }
    if constexpr (MODE == PS_MODE_MULTIPLY) {
      s_red[threadIdx.x] = dot;
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s_red[threadIdx.x] += s_red[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) a.dot[blockIdx.x] = s_red[0];
    }
  }
}

// dense level, one thread per 4-cell x group (msbg3.cpp:166-407)
template <int MODE>
__global__ __launch_bounds__(256) void k_ps_dense(PsDenseArgs a) {
#pragma clang fp contract(off)
  const int ngx = (a.sx + 3) / 4;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (int64_t)ngx * a.sy * a.sz) return;
  const int x0 = (int)(gid % ngx) * 4, y = (int)((gid / ngx) % a.sy), z = (int)(gid / ((int64_t)ngx * a.sy));
  const int64_t ys = a.sx + 2, zs = ys * (a.sy + 2);
  const int64_t p0 = (x0 + 1) + ys * (y + 1) + zs * (z + 1);
  const int nl = a.sx - x0 < 4 ? a.sx - x0 : 4;
  int anyfluid = 0, gated = 0;
  for (int l = 0; l < nl; l++) {
    const uint16_t f = a.flags[p0 + l];
    anyfluid |= (f & (kPsCellSolid | kPsCellVoid)) == 0;
    gated |= (f & (kPsCellSolid | kPsCellPartialSolid)) != 0;
  }
  if (!anyfluid) return;
  for (int l = 0; l < nl; l++) {
    const int64_t p = p0 + l;
    float F0 = a.X[p], F1 = a.X[p - 1], F2 = a.X[p + 1], F3 = a.X[p - ys], F4 = a.X[p + ys],
          F5 = a.X[p - zs], F6 = a.X[p + zs];
    if (gated) {
      F1 = F1 * a.fa[0][p];
      F2 = F2 * a.fa[0][p + 1];
      F3 = F3 * a.fa[1][p];
      F4 = F4 * a.fa[1][p + ys];
      F5 = F5 * a.fa[2][p];
      F6 = F6 * a.fa[2][p + zs];
    }
    const float S = (F1 + F2 + F3 + F4 + F5 + F6) * a.h, D = a.D[p];
    const bool fluid = (a.flags[p] & (kPsCellSolid | kPsCellVoid)) == 0;
    float r;
    if (MODE == PS_MODE_RELAX) {
      r = __builtin_fmaf(a.om1, F0, ((a.B[p] + S) * D) * a.hromega);
      if (!fluid) r = 0.0f;
    } else {
      const float N = 1.0f / D;
      r = (a.h * N) * F0 - S;
      if (!fluid) r = 0.0f;
      if (MODE == PS_MODE_RESIDUAL) r = a.B[p] - r;
    }
    a.Y[p] = r;
  }
}

__device__ __forceinline__ bool is_fluid(uint16_t f) { return (f & (kPsCellSolid | kPsCellVoid)) == 0; }

// block -> block restriction, one workgroup per block
__global__ __launch_bounds__(256) void k_ps_restrict_bb(PsXferArgs a, const float *src, float *dst) {
#pragma clang fp contract(off)
  const int b = blockIdx.x;
  const int Lf = a.Lf[b], Lc = a.Lc[b], bf = a.bsx0 >> Lf, bc = a.bsx0 >> Lc;
  const float *S = src + a.offF[b];
  float *Dd = dst + a.offC[b];
  const uint16_t *fc = a.flagsC + a.offC[b];
  for (int v = threadIdx.x; v < bc * bc * bc; v += blockDim.x) {
    const int x = v % bc, y = (v / bc) % bc, z = v / (bc * bc);
    float s;
    if (Lc == Lf) {
      s = S[v];
    } else {
      s = 0.0f;
#pragma unroll
      for (int k = 0; k < 8; k++)
        s = s + S[vox(bf, 2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + (k >> 2))];
    }
    Dd[v] = is_fluid(fc[v]) ? s : 0.0f;
  }
}

// block/dense -> dense restriction, one thread per coarse cell
__global__ __launch_bounds__(256) void k_ps_restrict_d(PsXferArgs a, const float *src, float *dst) {
#pragma clang fp contract(off)
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (int64_t)a.csx * a.csy * a.csz) return;
  const int x = (int)(gid % a.csx), y = (int)((gid / a.csx) % a.csy), z = (int)(gid / ((int64_t)a.csx * a.csy));
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int X = 2 * x + (k & 1), Y = 2 * y + ((k >> 1) & 1), Z = 2 * z + (k >> 2);
    int64_t p;
    if (a.kind == 1) {  // every block is at the last block level, 2^3 cells
      const int b = (X >> 1) + a.nbx * ((Y >> 1) + a.nby * (Z >> 1));
      p = a.offF[b] + vox(2, X & 1, Y & 1, Z & 1);
    } else {
      p = (X + 1) + (int64_t)(a.fsx + 2) * ((Y + 1) + (int64_t)(a.fsy + 2) * (Z + 1));
    }
    s = s + src[p];
  }
  const int64_t q = (x + 1) + (int64_t)(a.csx + 2) * ((y + 1) + (int64_t)(a.csy + 2) * (z + 1));
  dst[q] = is_fluid(a.flagsC[q]) ? s : 0.0f;
}

// block -> block prolongation (fine fluid cells add the parent's value)
__global__ __launch_bounds__(256) void k_ps_prolong_bb(PsXferArgs a, const float *src, float *dst) {
#pragma clang fp contract(off)
  const int b = blockIdx.x;
  const int Lf = a.Lf[b], Lc = a.Lc[b], bf = a.bsx0 >> Lf, bc = a.bsx0 >> Lc;
  const float *S = src + a.offC[b];
  float *Dd = dst + a.offF[b];
  const uint16_t *ff = a.flagsF + a.offF[b];
  for (int v = threadIdx.x; v < bf * bf * bf; v += blockDim.x) {
    if (!is_fluid(ff[v])) continue;
    const int x = v % bf, y = (v / bf) % bf, z = v / (bf * bf);
    const float e = Lc == Lf ? S[v] : S[vox(bc, x / 2, y / 2, z / 2)];
    Dd[v] = Dd[v] + e;
  }
}

// dense -> block/dense prolongation, one thread per fine cell
__global__ __launch_bounds__(256) void k_ps_prolong_d(PsXferArgs a, const float *src, float *dst) {
#pragma clang fp contract(off)
  const int fsx = a.kind == 1 ? 2 * a.nbx : a.fsx, fsy = a.kind == 1 ? 2 * a.nby : a.fsy,
            fsz = a.kind == 1 ? 2 * a.nbz : a.fsz;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (int64_t)fsx * fsy * fsz) return;
  const int X = (int)(gid % fsx), Y = (int)((gid / fsx) % fsy), Z = (int)(gid / ((int64_t)fsx * fsy));
  int64_t p;
  if (a.kind == 1) {
    const int b = (X >> 1) + a.nbx * ((Y >> 1) + a.nby * (Z >> 1));
    p = a.offF[b] + vox(2, X & 1, Y & 1, Z & 1);
  } else {
    p = (X + 1) + (int64_t)(fsx + 2) * ((Y + 1) + (int64_t)(fsy + 2) * (Z + 1));
  }
  if (!is_fluid(a.flagsF[p])) return;
  const int64_t q = (X / 2 + 1) + (int64_t)(a.csx + 2) * ((Y / 2 + 1) + (int64_t)(a.csy + 2) * (Z / 2 + 1));
  dst[p] = dst[p] + src[q];
}

__global__ void k_ps_fill(float *p, int64_t n, float v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_ps_copy(const float *s, float *d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}

template <int BSX, bool CUTS>
void launch_block_bsx(hipStream_t s, int mode, const PsBlockArgs &a, const PsItem *list, int n,
                      int wg) {
  if (mode == PS_MODE_RELAX)
    hipLaunchKernelGGL((k_ps_block<BSX, PS_MODE_RELAX, CUTS>), dim3(wg), dim3(256), 0, s, a, list, n);
  else if (mode == PS_MODE_MULTIPLY)
    hipLaunchKernelGGL((k_ps_block<BSX, PS_MODE_MULTIPLY, CUTS>), dim3(wg), dim3(256), 0, s, a, list, n);
  else
    hipLaunchKernelGGL((k_ps_block<BSX, PS_MODE_RESIDUAL, CUTS>), dim3(wg), dim3(256), 0, s, a, list, n);
}

int grid_1d(int64_t n) { return (int)((n + 255) / 256); }

}  // namespace

int ps_block_workgroups(int bsx, int n) {
  if (bsx < 2 || bsx > 16 || n <= 0) return 0;
  const int nb = 256 / (bsx * bsx);
  return (n + nb - 1) / nb;
}

int launch_ps_block(hipStream_t s, int bsx, int mode, bool cuts, const PsBlockArgs &a,
                    const PsItem *list, int n, int *nWorkgroups) {
  const int wg = ps_block_workgroups(bsx, n);
  if (nWorkgroups) *nWorkgroups = wg;
  if (n <= 0) return 0;
  switch (bsx * 2 + (cuts ? 1 : 0)) {
    case 32: launch_block_bsx<16, false>(s, mode, a, list, n, wg); break;
    case 33: launch_block_bsx<16, true>(s, mode, a, list, n, wg); break;
    case 16: launch_block_bsx<8, false>(s, mode, a, list, n, wg); break;
    case 17: launch_block_bsx<8, true>(s, mode, a, list, n, wg); break;
    case 8: launch_block_bsx<4, false>(s, mode, a, list, n, wg); break;
    case 9: launch_block_bsx<4, true>(s, mode, a, list, n, wg); break;
    case 4: launch_block_bsx<2, false>(s, mode, a, list, n, wg); break;
    case 5: launch_block_bsx<2, true>(s, mode, a, list, n, wg); break;
    default: set_error("pressure relax: block edge %d unsupported", bsx); return MSBG_EUNSUPP;
  }
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_ps_dense(hipStream_t s, int mode, const PsDenseArgs &a) {
  const int64_t n = (int64_t)((a.sx + 3) / 4) * a.sy * a.sz;
  if (mode == PS_MODE_RELAX) hipLaunchKernelGGL(k_ps_dense<PS_MODE_RELAX>, dim3(grid_1d(n)), dim3(256), 0, s, a);
  else if (mode == PS_MODE_MULTIPLY) hipLaunchKernelGGL(k_ps_dense<PS_MODE_MULTIPLY>, dim3(grid_1d(n)), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_ps_dense<PS_MODE_RESIDUAL>, dim3(grid_1d(n)), dim3(256), 0, s, a);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_ps_restrict(hipStream_t s, const PsXferArgs &a, const float *src, float *dst) {
  if (a.kind == 0)
    hipLaunchKernelGGL(k_ps_restrict_bb, dim3(a.nbx * a.nby * a.nbz), dim3(256), 0, s, a, src, dst);
  else
    hipLaunchKernelGGL(k_ps_restrict_d, dim3(grid_1d((int64_t)a.csx * a.csy * a.csz)), dim3(256), 0, s, a, src, dst);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_ps_prolongate(hipStream_t s, const PsXferArgs &a, const float *src, float *dst) {
  if (a.kind == 0) {
    hipLaunchKernelGGL(k_ps_prolong_bb, dim3(a.nbx * a.nby * a.nbz), dim3(256), 0, s, a, src, dst);
  } else {
    const int64_t n = a.kind == 1 ? (int64_t)8 * a.nbx * a.nby * a.nbz : (int64_t)a.fsx * a.fsy * a.fsz;
    hipLaunchKernelGGL(k_ps_prolong_d, dim3(grid_1d(n)), dim3(256), 0, s, a, src, dst);
  }
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_ps_fill(hipStream_t s, float *p, int64_t n, float v) {
  if (n <= 0) return 0;
  int g = grid_1d(n);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_ps_fill, dim3(g), dim3(256), 0, s, p, n, v);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_ps_copy(hipStream_t s, const float *src, float *dst, int64_t n) {
  if (n <= 0) return 0;
  int g = grid_1d(n);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_ps_copy, dim3(g), dim3(256), 0, s, src, dst, n);
  MSBG_HIP(hipGetLastError());
  return 0;
}

}  // namespace msbg
