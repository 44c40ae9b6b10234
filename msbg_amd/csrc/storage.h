// Storage formats of a channel's pages: raw element type, decode to the
// arithmetic type, encode with round-to-nearest-even. Shared by host and
// device code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/msbg_gpu.h"

namespace msbg {

__host__ __device__ inline uint32_t f2u(float f) {
  union {
    float f;
    uint32_t u;
  } c;
  c.f = f;
  return c.u;
}
__host__ __device__ inline float u2f(uint32_t u) {
  union {
    float f;
    uint32_t u;
  } c;
  c.u = u;
  return c.f;
}

// IEEE binary16 from float, round to nearest even (overflow -> inf).
__host__ __device__ inline uint16_t f16_bits_from_float(float f) {
  uint32_t x = f2u(f);
  uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
  if (ax < 0x38800000u) {
    uint32_t e = ax >> 23;
    if (e < 102) return (uint16_t)sign;
    uint32_t m = (ax & 0x7fffffu) | 0x800000u, sh = 126u - e;
    uint32_t q = m >> sh, rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t h = (((ax >> 23) - 112u) << 10) | ((ax & 0x7fffffu) >> 13), rem = ax & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)(sign | h);
}

__host__ __device__ inline float f16_bits_to_float(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  if (e == 0) {
    float v = (float)m * (1.0f / 16777216.0f);
    return u2f(f2u(v) | sign);
  }
  if (e == 31) return u2f(sign | 0x7f800000u | (m << 13));
  return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

__host__ __device__ inline uint16_t bf16_bits_from_float(float f) {
  uint32_t x = f2u(f);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}

__host__ __device__ inline float bf16_bits_to_float(uint16_t h) { return u2f((uint32_t)h << 16); }

template <int ST>
struct Stor;

// UNORM16: renderDensToFloat_ u*(1.f/65535) (sbg.h:233-244); store
// round_to_int(v*65535) = round half to even (sbg.h:343-351).
template <>
struct Stor<MSBG_STORAGE_U16> {
  using raw = uint16_t;
  using ct = float;
  __device__ static inline float dec(raw u) { return (float)u * (1.0f / 65535.0f); }
  __device__ static inline raw enc(float f) { return (raw)(int)__builtin_rintf(f * 65535.0f); }
};

template <>
struct Stor<MSBG_STORAGE_F32> {
  using raw = float;
  using ct = float;
  __device__ static inline float dec(raw u) { return u; }
  __device__ static inline raw enc(float f) { return f; }
};

template <>
struct Stor<MSBG_STORAGE_F16> {
  using raw = uint16_t;
  using ct = float;
  __device__ static inline float dec(raw u) {
    return (float)__builtin_bit_cast(_Float16, u);
  }
  __device__ static inline raw enc(float f) {
    return __builtin_bit_cast(uint16_t, (_Float16)f);  // v_cvt_f16_f32, RNE
  }
};

template <>
struct Stor<MSBG_STORAGE_BF16> {
  using raw = uint16_t;
  using ct = float;
  __device__ static inline float dec(raw u) { return u2f((uint32_t)u << 16); }
  __device__ static inline raw enc(float f) {
    uint32_t x = f2u(f);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (raw)(x >> 16);
  }
};

// UNORM8: renderDensToFloat_<uint8_t> u*(1.f/255) (sbg.h:228-231); the
// conversion store is round_to_int(v*255) = round half to even
// (renderDensFromFloat_storeSimd8<uint8_t>, sbg.h:319-323); the sweep's store
// rounds stochastically (enc_sr).
template <>
struct Stor<MSBG_STORAGE_U8> {
  using raw = uint8_t;
  using ct = float;
  __device__ static inline float dec(raw u) { return (float)u * (1.0f / 255.0f); }
  __device__ static inline raw enc(float f) { return (raw)(int)__builtin_rintf(f * 255.0f); }
};

// Stochastic rounding of the sweep's UNORM8 store
// (renderDensFromFloat_storeSimd8_SR<uint8_t>, sbg.h:324-340): with
// v = 255*f, the result is floor(v) + 1 when the uniform number u in [0,1)
// is below v - floor(v), else floor(v). The reference draws u from a
// per-thread Lehmer stream (FMA_FastRand, fastmath.h:68-73) whose order
// depends on its thread schedule; here u comes from a counter-based hash of
// (sweep step, page, voxel), with the reference's own construction of a
// float in [1,2) from the top 23 bits minus 1, so that one run is
// reproducible and the oracle draws the same numbers.
__host__ __device__ inline uint32_t sr_mix(uint32_t h) {  // murmur3's fmix32
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
__host__ __device__ inline float sr_uniform(uint32_t step, uint32_t page, uint32_t vid) {
  const uint32_t h = sr_mix(sr_mix(sr_mix(step ^ 0x6a09e667u) + page) + vid);
  return u2f((h >> 9) | 0x3f800000u) - 1.0f;
}
__device__ inline uint8_t u8_from_float_sr(float f, float u) {
#pragma clang fp contract(off)  // v is rounded before the fraction, as in the reference
  const float v = 255.0f * f, fl = floorf(v), fr = v - fl;
  return (uint8_t)(int)(u < fr ? fl + 1.0f : fl);
}

template <>
struct Stor<MSBG_STORAGE_F64> {
  using raw = double;
  using ct = double;
  __device__ static inline double dec(raw u) { return u; }
  __device__ static inline raw enc(double f) { return f; }
};

// raw storage value (as double) -> the storage's raw element, on the host
template <int ST>
inline typename Stor<ST>::raw raw_value(double raw) {
  if constexpr (ST == MSBG_STORAGE_U16) return (uint16_t)raw;
  else if constexpr (ST == MSBG_STORAGE_U8) return (uint8_t)raw;
  else if constexpr (ST == MSBG_STORAGE_F32) return (float)raw;
  else if constexpr (ST == MSBG_STORAGE_F16) return f16_bits_from_float((float)raw);
  else if constexpr (ST == MSBG_STORAGE_BF16) return bf16_bits_from_float((float)raw);
  else return raw;
}

// raw storage value (as double) -> float, for const block values on the host
inline float raw_to_float_host(int storage, double raw) {
  if (storage == MSBG_STORAGE_U16) return (float)(uint16_t)raw * (1.0f / 65535.0f);
  if (storage == MSBG_STORAGE_U8) return (float)(uint8_t)raw * (1.0f / 255.0f);
  return (float)raw;
}

}  // namespace msbg
