// Kernels of the learned warm start: a 3-stage 3D U-Net (16/32/64 channels)
// on 32^3 blocks with a 2-voxel halo, BF16 activations and weights, FP32
// accumulation on the gfx950 matrix cores.
//
// The network has no counterpart in the reference: it predicts the change of
// a block over K plain MCF sweeps (msbg4.cpp:60-957) so that the plain solver
// starts closer to its tolerance.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16:
//   forward / backward-data: D[cout][voxel] = W[cout][k] * X[k][voxel], with
//     k = (tap, input channel); a workgroup stages a (TZ+2)x(TY+2)x(XP+2)
//     input slab of up to 32 channels in LDS and each wave owns TN voxel
//     tiles of 16 consecutive x times all cout tiles;
//   backward-weights: dW[tap][cout][cin] = dZ^T[cout][k] * X_shift[k][cin],
//     k = (output voxel, 8 samples); both operands come from transposed
//     copies whose fastest index is the sample, so every fragment is one
//     aligned 16-byte load for any tap shift.
// Operand lane layout of the 16x16x32 instruction (checked on the device):
// A lane l holds row l%16, k = 8(l/16)..+7; B lane l holds column l%16,
// the same k; D lane l holds column l%16, rows 4(l/16)..+3.

#include <hip/hip_runtime.h>

#include <algorithm>

#include <cstdint>

#include "cnn_kernels.h"
#include "internal.h"
#include "storage.h"

namespace msbg {
namespace cnn {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ inline uint16_t f2bf(float f) { return bf16_bits_from_float(f); }
// two floats as packed BF16 bits (a low), rounded to nearest even by one
// v_cvt_pk_bf16_f32. With FP32 denormals kept (the kernels' mode), this
// equals f2bf for every non-NaN value; f2bf's NaN test made the software
// rounding a divergent branch per value, and the network's forward ran
// 10 % slower with it.
__device__ inline uint32_t pk_bf16(float a, float b) {
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  const bf16x2 v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, v);
}

__device__ inline void unpack8(const uint4 &u, float (&f)[8]) {
  const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    f[2 * i] = __uint_as_float(w[i] << 16);
    f[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
  }
}

__device__ inline uint4 pack8(const float (&f)[8]) {
  uint32_t w[4];
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = pk_bf16(f[2 * i], f[2 * i + 1]);
  return make_uint4(w[0], w[1], w[2], w[3]);
}

__host__ __device__ constexpr int chan_chunk(int cin) {
  return cin <= 32 ? cin : (cin % 32 == 0 ? 32 : 16);
}
__host__ __device__ constexpr int ksteps(int cin) { return (27 * chan_chunk(cin) + 31) / 32; }

template <int CIN, int COUT, int DOUT, int PAD>
struct ConvShape {
  static constexpr int DIN = DOUT + 2 - 2 * PAD;
  static constexpr int XP = (DOUT + 15) / 16 * 16;  // x extent padded to voxel tiles
  static constexpr int TZ = 2;
  static constexpr int TY = DOUT == 32 ? 4 : (DOUT > 32 ? 2 : 8);
  static constexpr int CC = chan_chunk(CIN);
  static constexpr int NCH = CIN / CC;
  static constexpr int KS = ksteps(CIN);
  static constexpr int TM = COUT / 16;
  static constexpr int SZ = TZ + 2, SY = TY + 2, SX = XP + 2;
  static constexpr int VS8 = CC / 8 + 1;  // LDS voxel stride in uint4 (one pad)
  static constexpr int NTILE = TZ * TY * XP / 16;
  static constexpr int TN = NTILE / 4;
  static constexpr int SLAB = SZ * SY * SX * VS8;  // uint4 elements
  // Whether a thread issues all its slab loads before its LDS stores. The
  // decoder's first 32^3 layer (upsample + skip, 48 -> 16) and the 64 -> 64
  // bottom layer stage one element at a time: batched, they lost a wave per
  // SIMD and ran 9 % and 35 % slower on MI355X, while the other layers ran
  // 6-24 % faster.
  static constexpr bool BATCH = !(CIN == 48 && COUT == 16 && DOUT == 32) &&
                                !(CIN == 64 && COUT == 64);
  static_assert(CIN % CC == 0 && CC % 8 == 0, "channel chunking");
  static_assert(COUT % 16 == 0, "cout tiles");
  static_assert(DOUT % TZ == 0 && DOUT % TY == 0, "spatial tiles");
  static_assert(NTILE % 4 == 0, "voxel tiles per wave");
};

// CA > 0: the input is the decoder's concatenation, read in place: its
// first CA channels are the coarser activation x (edge DIN/2, nearest
// upsampling), the other CIN - CA are the skip activation xb (edge DIN). A
// channel chunk never straddles the two.
// HEAD: the network's 1x1x1 head (16 -> 1, head = {hw[16], hb}) applied in
// the store: yh[n][v] = hb + sum_c hw[c] * bf16(y[n][v][c]) in channel order,
// the same fma chain as k_head, and y (the BF16 activation) is stored only
// when not null.
template <int CIN, int COUT, int DOUT, int PAD, int CA = 0, bool HEAD = false>
__global__ __launch_bounds__(256) void k_conv(const uint16_t *__restrict__ x,
                                              const uint16_t *__restrict__ xb,
                                              const uint4 *__restrict__ wsw,
                                              const float *__restrict__ bias, int relu,
                                              uint16_t *__restrict__ y,
                                              const float *__restrict__ head,
                                              float *__restrict__ yh) {
  using K = ConvShape<CIN, COUT, DOUT, PAD>;
  static_assert(CA % K::CC == 0 && CA < CIN, "concat boundary inside a channel chunk");
  static_assert(!HEAD || (COUT == 16 && K::XP == DOUT), "head: 16 channels, no idle lanes");
  constexpr int DIN = K::DIN, TZ = K::TZ, TY = K::TY, XP = K::XP, CC = K::CC, TM = K::TM,
                TN = K::TN, SY = K::SY, SX = K::SX, VS8 = K::VS8;
  __shared__ uint4 slab[K::SLAB];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane >> 4, col = lane & 15;
  const int n = blockIdx.y;
  constexpr int NYT = DOUT / TY;
  const int z0 = (int)(blockIdx.x / NYT) * TZ, y0 = (int)(blockIdx.x % NYT) * TY;
  const uint16_t *xs = x + (int64_t)n * DIN * DIN * DIN * CIN;

  int tz[TN], ty[TN], tx[TN];
#pragma unroll
  for (int t = 0; t < TN; t++) {
    const int T = wave * TN + t;
    tz[t] = T / (TY * XP / 16);
    ty[t] = (T / (XP / 16)) % TY;
    tx[t] = (T % (XP / 16)) * 16 + col;
  }
  f32x4 acc[TM][TN];
#pragma unroll
  for (int m = 0; m < TM; m++)
#pragma unroll
    for (int t = 0; t < TN; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};

  // Layers with one output tile (COUT 16) and one channel chunk hold their
  // weight fragments in registers, loaded before the slab is staged so that
  // their latency overlaps it, instead of one dependent load ahead of every
  // k-step (6-8 % on the 32^3 16 -> 16 layers; with three chunks the extra
  // registers cost the 48 -> 16 decoder layer a wave per SIMD and 22 %).
  constexpr bool WREG = TM == 1 && K::NCH == 1 && K::KS <= 16;
  bf16x8 aw[WREG ? K::KS : 1];
  for (int ch = 0; ch < K::NCH; ch++) {
    if (ch) __syncthreads();
    if constexpr (WREG) {
#pragma unroll
      for (int s = 0; s < K::KS; s++)
        aw[s] = __builtin_bit_cast(bf16x8, wsw[(int64_t)(ch * K::KS + s) * 64 + lane]);
    }
    // the slab element g of this chunk: (voxel, 8 channels), zero outside
    // the input
    constexpr int G8 = CC / 8, NG = K::SZ * SY * SX * G8, NIT = (NG + 255) / 256;
    auto fetch = [&](int g) {
      const int c8 = g % G8, sv = g / G8;
      const int sx = sv % SX, sy = (sv / SX) % SY, sz = sv / (SX * SY);
      const int iz = z0 - PAD + sz, iy = y0 - PAD + sy, ix = sx - PAD;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (iz >= 0 && iz < DIN && iy >= 0 && iy < DIN && ix >= 0 && ix < DIN) {
        if constexpr (CA == 0) {
          v = *(const uint4 *)(xs + (((int64_t)iz * DIN + iy) * DIN + ix) * CIN + ch * CC +
                               c8 * 8);
        } else if (ch * CC < CA) {  // upsampled coarse channels
          constexpr int DH = DIN / 2;
          const uint16_t *xa = x + (int64_t)n * DH * DH * DH * CA;
          v = *(const uint4 *)(xa + (((int64_t)(iz >> 1) * DH + (iy >> 1)) * DH + (ix >> 1)) * CA +
                               ch * CC + c8 * 8);
        } else {  // skip channels
          const uint16_t *xk = xb + (int64_t)n * DIN * DIN * DIN * (CIN - CA);
          v = *(const uint4 *)(xk + (((int64_t)iz * DIN + iy) * DIN + ix) * (CIN - CA) +
                               ch * CC - CA + c8 * 8);
        }
      }
      return v;
    };
    if constexpr (K::BATCH) {
      // every load of the slab is issued before the first LDS store, so that
      // a workgroup waits for one HBM round trip per chunk instead of one
      // per element it stages
      uint4 stage[NIT];
#pragma unroll
      for (int it = 0; it < NIT; it++) {
        const int g = tid + it * 256;
        stage[it] = g < NG ? fetch(g) : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (int it = 0; it < NIT; it++) {
        const int g = tid + it * 256;
        if (g < NG) slab[(g / G8) * VS8 + g % G8] = stage[it];
      }
    } else {
      for (int g = tid; g < NG; g += 256) slab[(g / G8) * VS8 + g % G8] = fetch(g);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < K::KS; s++) {
      const int kk = s * 32 + 8 * q;
      const int tap = kk / CC, c8 = (kk % CC) / 8;
      const bool ok = tap < 27;
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
      bf16x8 a[TM];
      if constexpr (WREG) {
        a[0] = aw[s];
      } else {
#pragma unroll
        for (int m = 0; m < TM; m++)
          a[m] = __builtin_bit_cast(bf16x8, wsw[((int64_t)(ch * K::KS + s) * TM + m) * 64 + lane]);
      }
#pragma unroll
      for (int t = 0; t < TN; t++) {
        uint4 bv = make_uint4(0, 0, 0, 0);
        if (ok) bv = slab[(((tz[t] + dz) * SY + ty[t] + dy) * SX + tx[t] + dx) * VS8 + c8];
        const bf16x8 b = __builtin_bit_cast(bf16x8, bv);
#pragma unroll
        for (int m = 0; m < TM; m++)
          acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b, acc[m][t], 0, 0, 0);
      }
    }
  }

  float hw[4] = {0.f, 0.f, 0.f, 0.f}, hb = 0.f;
  if constexpr (HEAD) {
#pragma unroll
    for (int r = 0; r < 4; r++) hw[r] = head[4 * q + r];
    hb = head[16];
  }
#pragma unroll
  for (int t = 0; t < TN; t++) {
    if (!HEAD && tx[t] >= DOUT) continue;
    const int64_t v = ((int64_t)(z0 + tz[t]) * DOUT + (y0 + ty[t])) * DOUT + tx[t];
#pragma unroll
    for (int m = 0; m < TM; m++) {
      const int co = m * 16 + 4 * q;
      float fv[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        float f = acc[m][t][r] + (bias ? bias[co + r] : 0.f);
        if (relu) f = f > 0.f ? f : 0.f;
        fv[r] = f;
      }
      const uint32_t p0 = pk_bf16(fv[0], fv[1]), p1 = pk_bf16(fv[2], fv[3]);
      const uint32_t h[4] = {p0 & 0xffffu, p0 >> 16, p1 & 0xffffu, p1 >> 16};
      if (!HEAD || y)
        *(uint2 *)(y + ((int64_t)n * DOUT * DOUT * DOUT + v) * COUT + co) = make_uint2(p0, p1);
      if constexpr (HEAD) {
        // lane group q holds channels 4q..4q+3 of voxel v; the running sum
        // passes from group q to group q+1 (16 lanes up), so that the chain
        // runs over channels 0..15 in order and ends in group 3
        float run = 0.f;
#pragma unroll
        for (int g = 0; g < 4; g++) {
          float r2 = run;
#pragma unroll
          for (int r = 0; r < 4; r++) r2 = __builtin_fmaf(hw[r], bf2f((uint16_t)h[r]), r2);
          run = q == g ? r2 : run;
          if (g < 3) {
            const float up = __shfl_up(run, 16, 64);
            run = q == g + 1 ? up : run;
          }
        }
        if (q == 3) yh[(int64_t)n * DOUT * DOUT * DOUT + v] = hb + run;
      }
    }
  }
}

// The single-channel first layer: k = tap (27 of 32 used), one channel, so
// the slab holds one BF16 per voxel and each lane gathers its 8 taps. A
// workgroup owns C1_TZ whole output planes; its waves stride over 16-voxel
// tiles of the planes' voxels in memory order, so that only the last tile
// of a workgroup has idle lanes (tiles of 16 x-consecutive voxels left 14 of
// every 48 lanes idle at the 34^3 edge).
constexpr int C1_TZ = 2;

template <int COUT, int DOUT, int PAD>
__global__ __launch_bounds__(256) void k_conv_c1(const uint16_t *__restrict__ x,
                                                 const uint4 *__restrict__ wsw,
                                                 const float *__restrict__ bias, int relu,
                                                 uint16_t *__restrict__ y) {
  constexpr int DIN = DOUT + 2 - 2 * PAD, TM = COUT / 16;
  constexpr int SZ = C1_TZ + 2, SY = DOUT + 2, SX = DOUT + 2;
  constexpr int NV = C1_TZ * DOUT * DOUT, NTILE = (NV + 15) / 16;
  static_assert(DOUT % C1_TZ == 0, "z tiles");
  __shared__ uint16_t slab[SZ * SY * SX];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane >> 4, col = lane & 15;
  const int n = blockIdx.y;
  const int z0 = (int)blockIdx.x * C1_TZ;
  const uint16_t *xs = x + (int64_t)n * DIN * DIN * DIN;
  // all loads of the slab are issued before the first LDS store (one HBM
  // round trip per workgroup, not one per staged voxel)
  constexpr int NG = SZ * SY * SX, NIT = (NG + 255) / 256;
  uint16_t stage[NIT];
#pragma unroll
  for (int it = 0; it < NIT; it++) {
    const int g = tid + it * 256;
    const int sx = g % SX, sy = (g / SX) % SY, sz = g / (SX * SY);
    const int iz = z0 - PAD + sz, iy = sy - PAD, ix = sx - PAD;
    uint16_t v = 0;
    if (g < NG && iz >= 0 && iz < DIN && iy >= 0 && iy < DIN && ix >= 0 && ix < DIN)
      v = xs[((int64_t)iz * DIN + iy) * DIN + ix];
    stage[it] = v;
  }
#pragma unroll
  for (int it = 0; it < NIT; it++)
    if (tid + it * 256 < NG) slab[tid + it * 256] = stage[it];
  bf16x8 a[TM];
#pragma unroll
  for (int m = 0; m < TM; m++) a[m] = __builtin_bit_cast(bf16x8, wsw[m * 64 + lane]);
  float bv[TM][4];
#pragma unroll
  for (int m = 0; m < TM; m++)
#pragma unroll
    for (int r = 0; r < 4; r++) bv[m][r] = bias ? bias[m * 16 + 4 * q + r] : 0.f;
  // this lane's 8 taps as slab offsets; taps 27..31 read offset 0 and are
  // masked to zero
  int toff[8];
  uint32_t tmask[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int tap = 8 * q + j;
    toff[j] = tap < 27 ? ((tap / 9) * SY + (tap / 3) % 3) * SX + tap % 3 : 0;
    tmask[j] = tap < 27 ? 0xffffu : 0u;
  }
  __syncthreads();

  uint16_t *yw = y + ((int64_t)n * DOUT * DOUT * DOUT + (int64_t)z0 * DOUT * DOUT) * COUT;
  for (int T = wave; T < NTILE; T += 4) {
    const int v = T * 16 + col;          // voxel of the workgroup's planes
    const int vc = v < NV ? v : NV - 1;  // idle lanes of the last tile read a valid voxel
    const int tz = vc / (DOUT * DOUT), ty = (vc / DOUT) % DOUT, tx = vc % DOUT;
    const int base = (tz * SY + ty) * SX + tx;
    uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const uint32_t t = slab[base + toff[j]] & tmask[j];
      w[j >> 1] |= t << (16 * (j & 1));
    }
    const bf16x8 b = __builtin_bit_cast(bf16x8, make_uint4(w[0], w[1], w[2], w[3]));
    f32x4 acc[TM];
#pragma unroll
    for (int m = 0; m < TM; m++)
      acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0,
                                                       0);
    if (v >= NV) continue;
    uint16_t *o = yw + (int64_t)v * COUT;
#pragma unroll
    for (int m = 0; m < TM; m++) {
      const int co = m * 16 + 4 * q;
      float fv[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        float f = acc[m][r] + bv[m][r];
        if (relu) f = f > 0.f ? f : 0.f;
        fv[r] = f;
      }
      *(uint2 *)(o + co) = make_uint2(pk_bf16(fv[0], fv[1]), pk_bf16(fv[2], fv[3]));
    }
  }
}

__global__ void k_swizzle(const float *__restrict__ w, int cin_true, int cin, int cout, int flip,
                          uint4 *__restrict__ out, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int cc = chan_chunk(cin), ks = ksteps(cin), tm = cout / 16;
  const int lane = (int)(i % 64);
  int64_t r = i / 64;
  const int m = (int)(r % tm);
  r /= tm;
  const int s = (int)(r % ks), ch = (int)(r / ks);
  const int co = m * 16 + (lane & 15);
  float f[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int kk = s * 32 + 8 * (lane >> 4) + j;
    const int tap = kk / cc, c = ch * cc + kk % cc;
    float v = 0.f;
    if (tap < 27) {
      if (!flip) {
        if (c < cin_true) v = w[((int64_t)co * cin_true + c) * 27 + tap];
      } else {
        // backward data: input channel c is the forward layer's output
        // channel, output co its input channel; taps are mirrored
        v = w[((int64_t)c * cout + co) * 27 + (26 - tap)];
      }
    }
    f[j] = v;
  }
  out[i] = pack8(f);
}

template <int CIN, int COUT, int DOUT>
__global__ __launch_bounds__(64) void k_wgrad(const uint16_t *__restrict__ dzt,
                                              const uint16_t *__restrict__ xtp, int lgNG,
                                              int64_t steps, int64_t spc,
                                              float *__restrict__ partial) {
  constexpr int DP = DOUT + 2;
  constexpr int64_t V = (int64_t)DOUT * DOUT * DOUT, VP = (int64_t)DP * DP * DP;
  constexpr int MT = COUT / 16, NT = CIN / 16;
  const int tap = blockIdx.x, chunk = blockIdx.y, lane = threadIdx.x;
  const int q = lane >> 4, col = lane & 15;
  const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
  const int64_t NS = (int64_t)8 << lgNG;  // samples
  const int64_t KG = V << lgNG;           // k-groups of 8 samples
  f32x4 acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int64_t s0 = chunk * spc, s1 = s0 + spc < steps ? s0 + spc : steps;
  for (int64_t s = s0; s < s1; s++) {
    const int64_t G = s * 4 + q;
    const bool ok = G < KG;
    const int64_t v = ok ? (G >> lgNG) : 0;
    const int ng = (int)(G & ((1 << lgNG) - 1));
    const int xx = (int)(v % DOUT), yy = (int)((v / DOUT) % DOUT), zz = (int)(v / (DOUT * DOUT));
    const int64_t vin = ((int64_t)(zz + dz) * DP + (yy + dy)) * DP + (xx + dx);
    bf16x8 a[MT], b[NT];
#pragma unroll
    for (int m = 0; m < MT; m++) {
      uint4 u = make_uint4(0, 0, 0, 0);
      if (ok) u = *(const uint4 *)(dzt + ((int64_t)(m * 16 + col) * V + v) * NS + ng * 8);
      a[m] = __builtin_bit_cast(bf16x8, u);
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
      uint4 u = make_uint4(0, 0, 0, 0);
      if (ok) u = *(const uint4 *)(xtp + ((int64_t)(t * 16 + col) * VP + vin) * NS + ng * 8);
      b[t] = __builtin_bit_cast(bf16x8, u);
    }
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
      for (int t = 0; t < NT; t++)
        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[t], acc[m][t], 0, 0, 0);
  }
  float *P = partial + ((int64_t)chunk * 27 + tap) * COUT * CIN;
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) P[(m * 16 + 4 * q + r) * CIN + t * 16 + col] = acc[m][t][r];
}

__global__ void k_wgrad_reduce(const float *__restrict__ partial, int chunks, int cin16,
                               int cin_true, int cout, float *__restrict__ dw) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)cout * cin_true * 27;
  if (i >= total) return;
  const int tap = (int)(i % 27), ci = (int)((i / 27) % cin_true), co = (int)(i / (27 * cin_true));
  float s = 0.f;
  for (int c = 0; c < chunks; c++)
    s += partial[(((int64_t)c * 27 + tap) * cout + co) * cin16 + ci];
  dw[i] = s;
}

// The weight-gradient layout: for channel ch and voxel v, the 8 samples of
// group g side by side ([ch][v][g][8]), from activations laid out
// [n][v][c]. One workgroup per row of XR voxels, slice of 16 channels and
// slice of up to 32 samples: the row is read with consecutive threads on
// consecutive channels of a voxel and consecutive voxels of a row
// (coalesced), staged in LDS, and written as 16-byte groups of 8 samples,
// consecutive threads on consecutive (voxel, group) pairs of one channel.
constexpr int kXC = 16;  // channels per workgroup
constexpr int kXS = 32;  // samples per workgroup
constexpr int kXR = 36;  // voxels of a row, at most

// activations x [n][din^3][c] into xtp [cpad][dp^3][n/8][8], dp = din + 2 pad,
// zero outside the input and for channels c..cpad-1
__global__ __launch_bounds__(256) void k_xpose_pad(const uint16_t *__restrict__ x, int n, int din,
                                                   int c, int pad, int cpad,
                                                   uint16_t *__restrict__ xtp) {
  __shared__ uint16_t T[kXS][kXR][kXC + 2];
  const int dp = din + 2 * pad;
  const int pz = (int)blockIdx.x / dp, py = (int)blockIdx.x % dp;
  const int c0 = (int)blockIdx.y * kXC, s0 = (int)blockIdx.z * kXS;
  const int ns = n - s0 < kXS ? n - s0 : kXS;
  const int iz = pz - pad, iy = py - pad;
  const bool rowIn = iz >= 0 && iy >= 0 && iz < din && iy < din;
  const int64_t vin = (int64_t)din * din * din;
  const int tid = (int)threadIdx.x;
  // load: (sample, voxel, channel), channel fastest
  const int nload = ns * din * kXC;
  for (int k = tid; k < nload; k += 256) {
    const int ch = k % kXC, xi = (k / kXC) % din, j = k / (kXC * din);
    uint16_t v = 0;
    if (rowIn && c0 + ch < c)
      v = x[((int64_t)(s0 + j) * vin + ((int64_t)iz * din + iy) * din + xi) * c + c0 + ch];
    T[j][xi][ch] = v;
  }
  __syncthreads();
  // store: (channel, voxel of the padded row, group), group fastest
  const int NG = n / 8, gs = ns / 8;
  const int64_t vp = (int64_t)dp * dp * dp;
  const int nstore = kXC * dp * gs;
  for (int k = tid; k < nstore; k += 256) {
    const int g = k % gs, px = (k / gs) % dp, ch = k / (gs * dp);
    if (c0 + ch >= cpad) continue;
    const int xi = px - pad;
    uint32_t w[4] = {0, 0, 0, 0};
    if (xi >= 0 && xi < din) {
#pragma unroll
      for (int j = 0; j < 8; j++) w[j >> 1] |= (uint32_t)T[g * 8 + j][xi][ch] << (16 * (j & 1));
    }
    const int64_t pv = ((int64_t)pz * dp + py) * dp + px;
    *(uint4 *)(xtp + ((int64_t)(c0 + ch) * vp + pv) * (NG * 8) + (s0 / 8 + g) * 8) =
        make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// ReLU backward in place (dy = 0 where y <= 0, when relu) and the same
// transpose of the masked gradient into dzt [c][v][n/8][8]; rows of d voxels
__global__ __launch_bounds__(256) void k_relu_bwd_xpose(uint16_t *__restrict__ dy,
                                                        const uint16_t *__restrict__ y, int n,
                                                        int d, int c, int relu,
                                                        uint16_t *__restrict__ dzt) {
  __shared__ uint16_t T[kXS][kXR][kXC + 2];
  const int rz = (int)blockIdx.x / d, ry = (int)blockIdx.x % d;
  const int c0 = (int)blockIdx.y * kXC, s0 = (int)blockIdx.z * kXS;
  const int ns = n - s0 < kXS ? n - s0 : kXS;
  const int64_t v = (int64_t)d * d * d;
  const int tid = (int)threadIdx.x;
  const int nload = ns * d * kXC;
  for (int k = tid; k < nload; k += 256) {
    const int ch = k % kXC, xi = (k / kXC) % d, j = k / (kXC * d);
    uint16_t g = 0;
    if (c0 + ch < c) {
      const int64_t idx = ((int64_t)(s0 + j) * v + ((int64_t)rz * d + ry) * d + xi) * c + c0 + ch;
      g = dy[idx];
      if (relu && !(bf2f(y[idx]) > 0.f)) {
        g = 0;
        dy[idx] = 0;
      }
    }
    T[j][xi][ch] = g;
  }
  __syncthreads();
  const int NG = n / 8, gs = ns / 8;
  const int nstore = kXC * d * gs;
  for (int k = tid; k < nstore; k += 256) {
    const int g = k % gs, xi = (k / gs) % d, ch = k / (gs * d);
    if (c0 + ch >= c) continue;
    uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 8; j++) w[j >> 1] |= (uint32_t)T[g * 8 + j][xi][ch] << (16 * (j & 1));
    const int64_t vv = ((int64_t)rz * d + ry) * d + xi;
    *(uint4 *)(dzt + ((int64_t)(c0 + ch) * v + vv) * (NG * 8) + (s0 / 8 + g) * 8) =
        make_uint4(w[0], w[1], w[2], w[3]);
  }
}

constexpr int kRowChunks = 64;

// partial[row][chunk]: chunk `blockIdx.y` of row `blockIdx.x`, fixed order
__global__ __launch_bounds__(256) void k_rowsum(const uint16_t *__restrict__ a, int64_t len,
                                                float *__restrict__ partial) {
  __shared__ float red[256];
  const int row = blockIdx.x, ch = blockIdx.y, tid = threadIdx.x;
  const int64_t per = (len + kRowChunks - 1) / kRowChunks;
  const int64_t b = ch * per, e = b + per < len ? b + per : len;
  const uint16_t *p = a + (int64_t)row * len;
  float s = 0.f;
  for (int64_t i = b + tid; i < e; i += 256) s += bf2f(p[i]);
  red[tid] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) red[tid] += red[tid + k];
    __syncthreads();
  }
  if (tid == 0) partial[(int64_t)row * kRowChunks + ch] = red[0];
}

__global__ void k_rowsum_final(const float *__restrict__ partial, int rows,
                               float *__restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float s = 0.f;
  for (int c = 0; c < kRowChunks; c++) s += partial[(int64_t)r * kRowChunks + c];
  out[r] = s;
}

__global__ void k_pool2(const uint16_t *__restrict__ x, int n, int d, int c,
                        uint16_t *__restrict__ y) {
  const int h = d / 2, g8 = c / 8;
  const int64_t vh = (int64_t)h * h * h, vd = (int64_t)d * d * d;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * vh * g8) return;
  const int c8 = (int)(i % g8);
  const int64_t v = (i / g8) % vh;
  const int64_t s = i / (g8 * vh);
  const int xx = (int)(v % h), yy = (int)((v / h) % h), zz = (int)(v / ((int64_t)h * h));
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 8; k++) {
    const int64_t cv = ((int64_t)(2 * zz + (k >> 2)) * d + 2 * yy + ((k >> 1) & 1)) * d + 2 * xx +
                       (k & 1);
    float f[8];
    unpack8(*(const uint4 *)(x + (s * vd + cv) * c + c8 * 8), f);
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] += f[j];
  }
#pragma unroll
  for (int j = 0; j < 8; j++) acc[j] *= 0.125f;
  *(uint4 *)(y + (s * vh + v) * c + c8 * 8) = pack8(acc);
}

__global__ void k_pool2_bwd(const uint16_t *__restrict__ dy, int n, int d, int c, int accumulate,
                            uint16_t *__restrict__ dx) {
  const int h = d / 2, g8 = c / 8;
  const int64_t vh = (int64_t)h * h * h, vd = (int64_t)d * d * d;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * vd * g8) return;
  const int c8 = (int)(i % g8);
  const int64_t v = (i / g8) % vd;
  const int64_t s = i / (g8 * vd);
  const int xx = (int)(v % d), yy = (int)((v / d) % d), zz = (int)(v / ((int64_t)d * d));
  const int64_t pv = ((int64_t)(zz / 2) * h + yy / 2) * h + xx / 2;
  float g[8], o[8];
  unpack8(*(const uint4 *)(dy + (s * vh + pv) * c + c8 * 8), g);
  uint4 *dst = (uint4 *)(dx + (s * vd + v) * c + c8 * 8);
  if (accumulate) unpack8(*dst, o);
#pragma unroll
  for (int j = 0; j < 8; j++) o[j] = accumulate ? o[j] + g[j] * 0.125f : g[j] * 0.125f;
  *dst = pack8(o);
}

__global__ void k_upcat(const uint16_t *__restrict__ a, int ca, const uint16_t *__restrict__ b,
                        int cb, int n, int d, uint16_t *__restrict__ out) {
  const int h = d / 2, c = ca + cb, g8 = c / 8;
  const int64_t vh = (int64_t)h * h * h, vd = (int64_t)d * d * d;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * vd * g8) return;
  const int c8 = (int)(i % g8);
  const int64_t v = (i / g8) % vd;
  const int64_t s = i / (g8 * vd);
  uint4 val;
  if (c8 * 8 < ca) {
    const int xx = (int)(v % d), yy = (int)((v / d) % d), zz = (int)(v / ((int64_t)d * d));
    const int64_t pv = ((int64_t)(zz / 2) * h + yy / 2) * h + xx / 2;
    val = *(const uint4 *)(a + (s * vh + pv) * ca + c8 * 8);
  } else {
    val = *(const uint4 *)(b + (s * vd + v) * cb + (c8 * 8 - ca));
  }
  *(uint4 *)(out + (s * vd + v) * c + c8 * 8) = val;
}

__global__ void k_upcat_bwd_a(const uint16_t *__restrict__ dout, int ca, int cb, int n, int d,
                              uint16_t *__restrict__ da) {
  const int h = d / 2, c = ca + cb, g8 = ca / 8;
  const int64_t vh = (int64_t)h * h * h, vd = (int64_t)d * d * d;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * vh * g8) return;
  const int c8 = (int)(i % g8);
  const int64_t v = (i / g8) % vh;
  const int64_t s = i / (g8 * vh);
  const int xx = (int)(v % h), yy = (int)((v / h) % h), zz = (int)(v / ((int64_t)h * h));
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 8; k++) {
    const int64_t cv = ((int64_t)(2 * zz + (k >> 2)) * d + 2 * yy + ((k >> 1) & 1)) * d + 2 * xx +
                       (k & 1);
    float f[8];
    unpack8(*(const uint4 *)(dout + (s * vd + cv) * c + c8 * 8), f);
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] += f[j];
  }
  *(uint4 *)(da + (s * vh + v) * ca + c8 * 8) = pack8(acc);
}

__global__ void k_upcat_bwd_b(const uint16_t *__restrict__ dout, int ca, int cb, int n, int d,
                              uint16_t *__restrict__ db) {
  const int c = ca + cb, g8 = cb / 8;
  const int64_t vd = (int64_t)d * d * d;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * vd * g8) return;
  const int c8 = (int)(i % g8);
  const int64_t sv = i / g8;
  *(uint4 *)(db + sv * cb + c8 * 8) = *(const uint4 *)(dout + sv * c + ca + c8 * 8);
}

constexpr int kHeadC = 16, kOut = 32;

__global__ void k_head(const uint16_t *__restrict__ d, int n, const float *__restrict__ head,
                       float *__restrict__ y) {
  constexpr int64_t V = (int64_t)kOut * kOut * kOut;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * V) return;
  float f[8], g[8];
  unpack8(*(const uint4 *)(d + i * kHeadC), f);
  unpack8(*(const uint4 *)(d + i * kHeadC + 8), g);
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < 8; j++) acc += head[j] * f[j];
#pragma unroll
  for (int j = 0; j < 8; j++) acc += head[8 + j] * g[j];
  y[i] = head[kHeadC] + acc;
}

__device__ inline float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_head_bwd(const float *__restrict__ dy,
                                                  const uint16_t *__restrict__ d, int n,
                                                  const float *__restrict__ head,
                                                  uint16_t *__restrict__ dd,
                                                  float *__restrict__ partial) {
  constexpr int64_t V = (int64_t)kOut * kOut * kOut;
  __shared__ float red[4][kHeadC + 1];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int tid = threadIdx.x;
  float g = 0.f, f[16] = {};
  if (i < n * V) {
    g = dy[i];
    float a[8], b[8];
    unpack8(*(const uint4 *)(d + i * kHeadC), a);
    unpack8(*(const uint4 *)(d + i * kHeadC + 8), b);
    float o[8], p[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      f[j] = a[j];
      f[8 + j] = b[j];
      o[j] = g * head[j];
      p[j] = g * head[8 + j];
    }
    *(uint4 *)(dd + i * kHeadC) = pack8(o);
    *(uint4 *)(dd + i * kHeadC + 8) = pack8(p);
  }
#pragma unroll
  for (int j = 0; j <= kHeadC; j++) {
    const float s = wave_sum(j < kHeadC ? g * f[j] : g);
    if ((tid & 63) == 0) red[tid >> 6][j] = s;
  }
  __syncthreads();
  if (tid <= kHeadC)
    partial[(int64_t)blockIdx.x * (kHeadC + 1) + tid] =
        ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
}

// out[c] = sum over rows of partial[row][c]: one block per column, fixed order
__global__ __launch_bounds__(256) void k_sum_partials(const float *__restrict__ partial, int rows,
                                                      int width, float *__restrict__ out) {
  __shared__ float red[256];
  const int c = blockIdx.x, tid = threadIdx.x;
  float s = 0.f;
  for (int r = tid; r < rows; r += 256) s += partial[(int64_t)r * width + c];
  red[tid] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) red[tid] += red[tid + k];
    __syncthreads();
  }
  if (tid == 0) out[c] = red[0];
}

__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ y,
                                             const float *__restrict__ t, int64_t count,
                                             float *__restrict__ dy, double *__restrict__ partial) {
  __shared__ double red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int tid = threadIdx.x;
  double e = 0.0;
  if (i < count) {
    const float d = y[i] - t[i];
    dy[i] = d * (2.0f / (float)count);
    e = (double)d * (double)d;
  }
  red[tid] = e;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) red[tid] += red[tid + k];
    __syncthreads();
  }
  if (tid == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_sum_f64(const double *__restrict__ partial, int n,
                                                 double *__restrict__ out) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int i = tid; i < n; i += 256) s += partial[i];
  red[tid] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) red[tid] += red[tid + k];
    __syncthreads();
  }
  if (tid == 0) *out = red[0];
}

// acc[0] += *sum * scale, acc[1] += 1 (the running loss of queued steps)
__global__ void k_loss_acc(const double *__restrict__ sum, double scale, double *__restrict__ acc) {
  if (threadIdx.x == 0) {
    acc[0] += *sum * scale;
    acc[1] += 1.0;
  }
}

// dst[i] = src[rows[i]], rows of `row16` 16-byte words
__global__ __launch_bounds__(256) void k_gather_rows(uint4 *__restrict__ dst,
                                                     const uint4 *__restrict__ src,
                                                     const int32_t *__restrict__ rows,
                                                     int64_t row16) {
  const int64_t r = blockIdx.y;
  const uint4 *s = src + (int64_t)rows[r] * row16;
  uint4 *d = dst + r * row16;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < row16;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

inline unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

template <int CIN, int COUT, int DOUT, int PAD, int CA = 0>
int conv_t(hipStream_t s, const uint16_t *x, int n, const uint4 *wsw, const float *bias, bool relu,
           uint16_t *y, const uint16_t *xb = nullptr) {
  using K = ConvShape<CIN, COUT, DOUT, PAD>;
  dim3 grid((DOUT / K::TZ) * (DOUT / K::TY), (unsigned)n);
  hipLaunchKernelGGL((k_conv<CIN, COUT, DOUT, PAD, CA>), grid, dim3(256), 0, s, x, xb, wsw, bias,
                     relu ? 1 : 0, y, nullptr, nullptr);
  MSBG_HIP(hipGetLastError());
  return 0;
}

template <int COUT, int DOUT, int PAD>
int conv_c1_t(hipStream_t s, const uint16_t *x, int n, const uint4 *wsw, const float *bias,
              bool relu, uint16_t *y) {
  dim3 grid(DOUT / C1_TZ, (unsigned)n);
  hipLaunchKernelGGL((k_conv_c1<COUT, DOUT, PAD>), grid, dim3(256), 0, s, x, wsw, bias,
                     relu ? 1 : 0, y);
  MSBG_HIP(hipGetLastError());
  return 0;
}

template <int CIN, int COUT, int DOUT>
int wgrad_t(hipStream_t s, const uint16_t *dzt, const uint16_t *xtp, int n, float *partial) {
  int lg = 0;
  while ((8 << lg) < n) lg++;
  const int64_t V = (int64_t)DOUT * DOUT * DOUT;
  const int64_t steps = ((V << lg) + 3) / 4;
  const int chunks = wgrad_chunks(CIN, COUT, DOUT, n);
  const int64_t spc = (steps + chunks - 1) / chunks;
  hipLaunchKernelGGL((k_wgrad<CIN, COUT, DOUT>), dim3(27, (unsigned)chunks), dim3(64), 0, s, dzt,
                     xtp, lg, steps, spc, partial);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int lg_groups(int n) {
  int lg = 0;
  while ((8 << lg) < n) lg++;
  return (8 << lg) == n ? lg : -1;
}

}  // namespace

int64_t swizzled_size(int cin, int cout) {
  return (int64_t)(cin / chan_chunk(cin)) * ksteps(cin) * (cout / 16) * 64;
}

int launch_swizzle(hipStream_t s, const float *w, int cin_true, int cin, int cout, bool flip,
                   uint4 *out) {
  MSBG_CHECK((cin == 1 || cin % 8 == 0) && cout % 16 == 0 && cin_true <= cin, MSBG_EINVAL,
             "swizzle: bad shape cin %d cout %d", cin, cout);
  const int64_t total = swizzled_size(cin, cout);
  hipLaunchKernelGGL(k_swizzle, dim3(blocks_for(total)), dim3(256), 0, s, w, cin_true, cin, cout,
                     flip ? 1 : 0, out, total);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_conv(hipStream_t s, const uint16_t *x, int n, int cin, int cout, int dout, int pad,
                const uint4 *wsw, const float *bias, bool relu, uint16_t *y) {
  MSBG_CHECK(n >= 1 && n <= 65535, MSBG_EINVAL, "conv: batch %d outside [1, 65535]", n);
#define MSBG_CONV_CASE(CI, CO, D, P) \
  if (cin == CI && cout == CO && dout == D && pad == P) return conv_t<CI, CO, D, P>(s, x, n, wsw, bias, relu, y);
  // forward layers
  if (cin == 1 && cout == 16 && dout == 34 && pad == 0)
    return conv_c1_t<16, 34, 0>(s, x, n, wsw, bias, relu, y);
  MSBG_CONV_CASE(16, 16, 32, 0)
  MSBG_CONV_CASE(16, 32, 16, 1)
  MSBG_CONV_CASE(32, 32, 16, 1)
  MSBG_CONV_CASE(32, 64, 8, 1)
  MSBG_CONV_CASE(64, 64, 8, 1)
  MSBG_CONV_CASE(96, 32, 16, 1)
  MSBG_CONV_CASE(48, 16, 32, 1)
  MSBG_CONV_CASE(16, 16, 32, 1)
  // backward-data convolutions
  MSBG_CONV_CASE(16, 48, 32, 1)
  MSBG_CONV_CASE(32, 96, 16, 1)
  MSBG_CONV_CASE(64, 32, 8, 1)
  MSBG_CONV_CASE(32, 16, 16, 1)
  MSBG_CONV_CASE(16, 16, 34, 2)
#undef MSBG_CONV_CASE
  set_error("conv: no kernel for cin %d cout %d dout %d pad %d", cin, cout, dout, pad);
  return MSBG_EUNSUPP;
}

int launch_conv_cat(hipStream_t s, const uint16_t *a, int ca, const uint16_t *b, int cb, int n,
                    int cout, int dout, const uint4 *wsw, const float *bias, bool relu,
                    uint16_t *y) {
  MSBG_CHECK(n >= 1 && n <= 65535, MSBG_EINVAL, "conv: batch %d outside [1, 65535]", n);
  if (ca == 64 && cb == 32 && cout == 32 && dout == 16)
    return conv_t<96, 32, 16, 1, 64>(s, a, n, wsw, bias, relu, y, b);
  if (ca == 32 && cb == 16 && cout == 16 && dout == 32)
    return conv_t<48, 16, 32, 1, 32>(s, a, n, wsw, bias, relu, y, b);
  set_error("conv: no concatenating kernel for %d+%d -> %d at %d^3", ca, cb, cout, dout);
  return MSBG_EUNSUPP;
}

int launch_conv_head(hipStream_t s, const uint16_t *x, int n, int cin, int cout, int dout,
                     int pad, const uint4 *wsw, const float *bias, bool relu, uint16_t *y,
                     const float *head, float *yh) {
  MSBG_CHECK(n >= 1 && n <= 65535, MSBG_EINVAL, "conv: batch %d outside [1, 65535]", n);
  MSBG_CHECK(head && yh, MSBG_EINVAL, "conv head: null head or output");
  if (cin == 16 && cout == 16 && dout == kOut && pad == 1) {
    using K = ConvShape<16, 16, kOut, 1>;
    dim3 grid((kOut / K::TZ) * (kOut / K::TY), (unsigned)n);
    hipLaunchKernelGGL((k_conv<16, 16, kOut, 1, 0, true>), grid, dim3(256), 0, s, x, nullptr, wsw,
                       bias, relu ? 1 : 0, y, head, yh);
    MSBG_HIP(hipGetLastError());
    return 0;
  }
  set_error("conv head: no kernel for cin %d cout %d dout %d pad %d", cin, cout, dout, pad);
  return MSBG_EUNSUPP;
}

int wgrad_chunks(int cin, int cout, int dout, int n) {
  (void)cin;
  (void)cout;
  const int64_t steps = ((int64_t)dout * dout * dout * (n / 8) + 3) / 4;
  const int64_t c = steps / 64;  // at least 64 steps per wave
  return (int)(c < 1 ? 1 : (c > 96 ? 96 : c));
}

int launch_wgrad(hipStream_t s, const uint16_t *dz_t, const uint16_t *x_tp, int n, int cin16,
                 int cout, int dout, float *partial) {
  MSBG_CHECK(lg_groups(n) >= 0, MSBG_EINVAL, "wgrad: batch %d is not 8 * 2^k", n);
#define MSBG_WG_CASE(CI, CO, D) \
  if (cin16 == CI && cout == CO && dout == D) return wgrad_t<CI, CO, D>(s, dz_t, x_tp, n, partial);
  MSBG_WG_CASE(16, 16, 34)
  MSBG_WG_CASE(16, 16, 32)
  MSBG_WG_CASE(16, 32, 16)
  MSBG_WG_CASE(32, 32, 16)
  MSBG_WG_CASE(32, 64, 8)
  MSBG_WG_CASE(64, 64, 8)
  MSBG_WG_CASE(96, 32, 16)
  MSBG_WG_CASE(48, 16, 32)
#undef MSBG_WG_CASE
  set_error("wgrad: no kernel for cin %d cout %d dout %d", cin16, cout, dout);
  return MSBG_EUNSUPP;
}

int launch_wgrad_reduce(hipStream_t s, const float *partial, int chunks, int cin16, int cin_true,
                        int cout, float *dw) {
  const int64_t total = (int64_t)cout * cin_true * 27;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for(total)), dim3(256), 0, s, partial, chunks,
                     cin16, cin_true, cout, dw);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_xpose_pad(hipStream_t s, const uint16_t *x, int n, int din, int c, int pad, int cpad,
                     uint16_t *x_tp) {
  const int lg = lg_groups(n);
  MSBG_CHECK(lg >= 0 && cpad >= c && din <= kXR, MSBG_EINVAL,
             "xpose: batch %d is not 8 * 2^k, or rows of %d voxels", n, din);
  const int dp = din + 2 * pad;
  const dim3 grid((unsigned)(dp * dp), (unsigned)((cpad + kXC - 1) / kXC),
                  (unsigned)((n + kXS - 1) / kXS));
  hipLaunchKernelGGL(k_xpose_pad, grid, dim3(256), 0, s, x, n, din, c, pad, cpad, x_tp);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_relu_bwd_xpose(hipStream_t s, uint16_t *dy, const uint16_t *y, int n, int64_t v, int c,
                          bool relu, uint16_t *dz_t) {
  const int lg = lg_groups(n);
  int d = 1;
  while ((int64_t)d * d * d < v) d++;
  MSBG_CHECK(lg >= 0 && (int64_t)d * d * d == v && d <= kXR, MSBG_EINVAL,
             "relu_bwd: batch %d is not 8 * 2^k, or volume %lld is not a cube of <= %d",
             n, (long long)v, kXR);
  const dim3 grid((unsigned)(d * d), (unsigned)((c + kXC - 1) / kXC),
                  (unsigned)((n + kXS - 1) / kXS));
  hipLaunchKernelGGL(k_relu_bwd_xpose, grid, dim3(256), 0, s, dy, y, n, d, c, relu ? 1 : 0, dz_t);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int rowsum_scratch(int rows) { return rows * kRowChunks; }

int launch_rowsum(hipStream_t s, const uint16_t *a, int rows, int64_t len, float *scratch,
                  float *out) {
  hipLaunchKernelGGL(k_rowsum, dim3((unsigned)rows, kRowChunks), dim3(256), 0, s, a, len, scratch);
  MSBG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_rowsum_final, dim3(1), dim3(128), 0, s, scratch, rows, out);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_pool2(hipStream_t s, const uint16_t *x, int n, int d, int c, uint16_t *y) {
  const int64_t total = (int64_t)n * (d / 2) * (d / 2) * (d / 2) * (c / 8);
  hipLaunchKernelGGL(k_pool2, dim3(blocks_for(total)), dim3(256), 0, s, x, n, d, c, y);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_pool2_bwd(hipStream_t s, const uint16_t *dy, int n, int d, int c, bool accumulate,
                     uint16_t *dx) {
  const int64_t total = (int64_t)n * d * d * d * (c / 8);
  hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks_for(total)), dim3(256), 0, s, dy, n, d, c,
                     accumulate ? 1 : 0, dx);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_upcat(hipStream_t s, const uint16_t *a, int ca, const uint16_t *b, int cb, int n, int d,
                 uint16_t *out) {
  const int64_t total = (int64_t)n * d * d * d * ((ca + cb) / 8);
  hipLaunchKernelGGL(k_upcat, dim3(blocks_for(total)), dim3(256), 0, s, a, ca, b, cb, n, d, out);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_upcat_bwd(hipStream_t s, const uint16_t *dout, int ca, int cb, int n, int d,
                     uint16_t *da, uint16_t *db) {
  const int h = d / 2;
  const int64_t ta = (int64_t)n * h * h * h * (ca / 8), tb = (int64_t)n * d * d * d * (cb / 8);
  hipLaunchKernelGGL(k_upcat_bwd_a, dim3(blocks_for(ta)), dim3(256), 0, s, dout, ca, cb, n, d, da);
  MSBG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_upcat_bwd_b, dim3(blocks_for(tb)), dim3(256), 0, s, dout, ca, cb, n, d, db);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_head(hipStream_t s, const uint16_t *d, int n, const float *head, float *y) {
  const int64_t total = (int64_t)n * kOut * kOut * kOut;
  hipLaunchKernelGGL(k_head, dim3(blocks_for(total)), dim3(256), 0, s, d, n, head, y);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int head_bwd_blocks(int n) { return (int)blocks_for((int64_t)n * kOut * kOut * kOut); }

int launch_head_bwd(hipStream_t s, const float *dy, const uint16_t *d, int n, const float *head,
                    uint16_t *dd, float *partial) {
  hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)head_bwd_blocks(n)), dim3(256), 0, s, dy, d, n,
                     head, dd, partial);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_sum_partials(hipStream_t s, const float *partial, int rows, int width, float *out) {
  hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)width), dim3(256), 0, s, partial, rows, width,
                     out);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int mse_blocks(int64_t count) { return (int)blocks_for(count); }

int launch_mse(hipStream_t s, const float *y, const float *t, int64_t count, float *dy,
               double *partial) {
  hipLaunchKernelGGL(k_mse, dim3((unsigned)mse_blocks(count)), dim3(256), 0, s, y, t, count, dy,
                     partial);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_sum_f64(hipStream_t s, const double *partial, int n, double *out) {
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, s, partial, n, out);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_loss_acc(hipStream_t s, const double *sum, double scale, double *acc) {
  hipLaunchKernelGGL(k_loss_acc, dim3(1), dim3(64), 0, s, sum, scale, acc);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_gather_rows(hipStream_t s, void *dst, const void *src, const int32_t *rows, int n,
                       int64_t rowBytes) {
  if (n <= 0) return 0;
  if (rowBytes % 16) {
    set_error("gather rows: row of %lld bytes is not a multiple of 16", (long long)rowBytes);
    return MSBG_EINVAL;
  }
  const int64_t row16 = rowBytes / 16;
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)std::min<int64_t>((row16 + 255) / 256, 64), n),
                     dim3(256), 0, s, (uint4 *)dst, (const uint4 *)src, rows, row16);
  MSBG_HIP(hipGetLastError());
  return 0;
}

int launch_adam(hipStream_t s, float *p, const float *g, float *m, float *v, int64_t n, float lr,
                float b1, float b2, float eps, float bc1, float bc2) {
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps,
                     bc1, bc2);
  MSBG_HIP(hipGetLastError());
  return 0;
}

}  // namespace cnn
}  // namespace msbg
