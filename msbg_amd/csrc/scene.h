// Seeded scene objects shared by the host and device point generators:
// PRNGs, gradient noise and the per-object draw (centre, semi-axes,
// rotation, point stream seed, noise window).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

namespace msbg {
namespace scene {

inline uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}


// xoshiro256** (Blackman & Vigna)
struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed) {
    uint64_t x = seed;
    for (auto &v : s) v = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  // Box-Muller; returns one normal deviate, caches the other
  bool have = false;
  double cached = 0;
  double normal() {
    if (have) {
      have = false;
      return cached;
    }
    double u1 = 1.0 - uniform(), u2 = uniform();
    double r = std::sqrt(-2.0 * std::log(u1)), a = 2.0 * M_PI * u2;
    cached = r * std::sin(a);
    have = true;
    return r * std::cos(a);
  }
};

// Gradient noise after Perlin, "Improving Noise" (SIGGRAPH 2002): quintic
// fade, 12 edge gradients picked by a hashed lattice corner. The
// permutation is a seeded shuffle, so each scene seed has its own field.
__host__ __device__ inline double perlin_fade(double t) {
  return t * t * t * (t * (t * 6 - 15) + 10);
}
__host__ __device__ inline double perlin_lerp(double t, double a, double b) {
  return a + t * (b - a);
}
__host__ __device__ inline double perlin_grad(int h, double x, double y, double z) {
  h &= 15;
  double u = h < 8 ? x : y, v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
  return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
}
// in [-1, 1]; p is the doubled 512-entry permutation
__host__ __device__ inline double perlin_noise(const int *p, double x, double y, double z) {
  const double fx = floor(x), fy = floor(y), fz = floor(z);
  const int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
  x -= fx;
  y -= fy;
  z -= fz;
  const double u = perlin_fade(x), v = perlin_fade(y), w = perlin_fade(z);
  const int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z, B = p[X + 1] + Y, BA = p[B] + Z,
            BB = p[B + 1] + Z;
  return perlin_lerp(
      w,
      perlin_lerp(v, perlin_lerp(u, perlin_grad(p[AA], x, y, z), perlin_grad(p[BA], x - 1, y, z)),
                  perlin_lerp(u, perlin_grad(p[AB], x, y - 1, z),
                              perlin_grad(p[BB], x - 1, y - 1, z))),
      perlin_lerp(v,
                  perlin_lerp(u, perlin_grad(p[AA + 1], x, y, z - 1),
                              perlin_grad(p[BA + 1], x - 1, y, z - 1)),
                  perlin_lerp(u, perlin_grad(p[AB + 1], x, y - 1, z - 1),
                              perlin_grad(p[BB + 1], x - 1, y - 1, z - 1))));
}

struct Perlin {
  int p[512];
  explicit Perlin(uint64_t seed) {
    Xoshiro r(seed);
    for (int i = 0; i < 256; i++) p[i] = i;
    for (int i = 255; i > 0; i--) {
      int j = (int)(r.next() % (uint64_t)(i + 1));
      int t = p[i];
      p[i] = p[j];
      p[j] = t;
    }
    for (int i = 0; i < 256; i++) p[256 + i] = p[i];
  }
  double operator()(double x, double y, double z) const { return perlin_noise(p, x, y, z); }
};

// One object of a scene: centre, semi-axes, rotation (row-major), the seed
// of its point stream and its window into the noise field.
struct Obj {
  double c[3], ax[3], R[9];
  uint64_t seed;
  double off[3];
};

inline std::vector<Obj> draw_objects(uint64_t seed, int nObj, double cmin, double cmax,
                                     double rmin, double rmax, double ellipsoidFrac) {
  std::vector<Obj> objs((size_t)(nObj > 0 ? nObj : 0));
  Xoshiro sr(seed);
  for (int o = 0; o < nObj; o++) {
    Obj &ob = objs[(size_t)o];
    for (int k = 0; k < 3; k++) ob.c[k] = cmin + (cmax - cmin) * sr.uniform();
    bool ell = sr.uniform() < ellipsoidFrac;
    double r0 = rmin + (rmax - rmin) * sr.uniform();
    for (int k = 0; k < 3; k++) ob.ax[k] = ell ? rmin + (rmax - rmin) * sr.uniform() : r0;
    // rotation from a normalised Gaussian quaternion
    double q[4];
    double qn = 0;
    for (double &v : q) {
      v = sr.normal();
      qn += v * v;
    }
    qn = std::sqrt(qn);
    double w = q[0] / qn, x = q[1] / qn, y = q[2] / qn, z = q[3] / qn;
    if (!ell) w = 1, x = y = z = 0;
    double *R = ob.R;
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z);     R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y);     R[7] = 2 * (y * z + w * x);     R[8] = 1 - 2 * (x * x + y * y);
    uint64_t t = seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(o + 1));
    ob.seed = splitmix64(t);
    uint64_t ns = ob.seed ^ 0xA0761D6478BD642Full;
    for (double &v : ob.off) v = (double)(splitmix64(ns) >> 11) * 0x1.0p-53 * 256.0;
  }
  return objs;
}

// the permutation of a scene's noise field
inline Perlin scene_noise(uint64_t seed) { return Perlin(seed ^ 0x6A09E667F3BCC909ull); }

// Surface point of object `ob` in direction n (any non-zero vector), noise
// displaced along the ray, plus the jitter j (already scaled by sigma).
__host__ __device__ inline void object_point(const Obj &ob, const int *perm, const double n[3],
                                             const double j[3], double amp, double freq,
                                             float out[3]) {
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double u[3] = {ob.ax[0] * n[0] / nn, ob.ax[1] * n[1] / nn, ob.ax[2] * n[2] / nn};
  if (amp > 0) {  // displace along the ray: radius * (1 + amp * noise)
    const double f = 1.0 + amp * perlin_noise(perm, freq * u[0] + ob.off[0],
                                              freq * u[1] + ob.off[1], freq * u[2] + ob.off[2]);
    for (double &v : u) v *= f;
  }
  for (int k = 0; k < 3; k++) {
    double p = ob.c[k] + ob.R[3 * k] * u[0] + ob.R[3 * k + 1] * u[1] + ob.R[3 * k + 2] * u[2];
    out[k] = (float)(p + j[k]);
  }
}

}  // namespace scene
}  // namespace msbg
