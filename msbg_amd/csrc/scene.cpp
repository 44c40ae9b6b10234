// Seeded synthetic particle scenes: jittered surface points of spheres and
// randomly rotated ellipsoids, optionally with Perlin-noise displaced
// surfaces. Each object has its own PRNG stream derived from (seed, object
// index), so the output does not depend on the number of host threads.
// Positions are computed in double and rounded once to float.

#include <cmath>
#include <cstdint>

#include "internal.h"

namespace {

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
  static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
  static double lerp(double t, double a, double b) { return a + t * (b - a); }
  static double grad(int h, double x, double y, double z) {
    h &= 15;
    double u = h < 8 ? x : y, v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
  }
  // in [-1, 1]
  double operator()(double x, double y, double z) const {
    const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
    x -= fx;
    y -= fy;
    z -= fz;
    const double u = fade(x), v = fade(y), w = fade(z);
    const int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z, B = p[X + 1] + Y, BA = p[B] + Z,
              BB = p[B + 1] + Z;
    return lerp(w,
                lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                     lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(p[AB + 1], x, y - 1, z - 1),
                          grad(p[BB + 1], x - 1, y - 1, z - 1))));
  }
};

}  // namespace

extern "C" int msbg_scene_spheres(uint64_t seed, int nObj, float cmin, float cmax, float rmin,
                                  float rmax, int64_t nPts, float sigma, float ellipsoidFrac,
                                  float *xyz) {
  return msbg_scene_objects(seed, nObj, cmin, cmax, rmin, rmax, nPts, sigma, ellipsoidFrac, 0.0f,
                            0.0f, xyz);
}

extern "C" int msbg_scene_objects(uint64_t seed, int nObj, float cmin, float cmax, float rmin,
                                  float rmax, int64_t nPts, float sigma, float ellipsoidFrac,
                                  float perlinAmp, float perlinFreq, float *xyz) {
  MSBG_CHECK(nObj >= 0 && nPts >= 0 && xyz && cmax >= cmin && rmax >= rmin && sigma >= 0 &&
                 perlinAmp >= 0 && perlinAmp < 1 && perlinFreq >= 0,
             MSBG_EINVAL, "msbg_scene_objects: bad args");
  struct Obj {
    double c[3], ax[3], R[9];
    uint64_t seed;
  };
  const Perlin noise(seed ^ 0x6A09E667F3BCC909ull);
  Obj *objs = new Obj[nObj > 0 ? nObj : 1];
  {
    Xoshiro sr(seed);
    for (int o = 0; o < nObj; o++) {
      Obj &ob = objs[o];
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
    }
  }
  msbg::parallel_for(nObj, [&](int64_t o0, int64_t o1) {
  for (int64_t o = o0; o < o1; o++) {
    const Obj &ob = objs[o];
    Xoshiro r(ob.seed);
    // this object's window into the noise field (its own stream)
    uint64_t ns = ob.seed ^ 0xA0761D6478BD642Full;
    double off[3];
    for (double &v : off) v = (double)(splitmix64(ns) >> 11) * 0x1.0p-53 * 256.0;
    float *out = xyz + (size_t)3 * nPts * o;
    for (int64_t i = 0; i < nPts; i++) {
      double n[3], nn = 0;
      do {
        nn = 0;
        for (double &v : n) {
          v = r.normal();
          nn += v * v;
        }
      } while (nn < 1e-24);
      nn = std::sqrt(nn);
      double u[3] = {ob.ax[0] * n[0] / nn, ob.ax[1] * n[1] / nn, ob.ax[2] * n[2] / nn};
      if (perlinAmp > 0) {  // displace along the ray: radius * (1 + amp * noise)
        const double f = 1.0 + perlinAmp * noise(perlinFreq * u[0] + off[0],
                                                  perlinFreq * u[1] + off[1],
                                                  perlinFreq * u[2] + off[2]);
        for (double &v : u) v *= f;
      }
      for (int k = 0; k < 3; k++) {
        double p = ob.c[k] + ob.R[3 * k] * u[0] + ob.R[3 * k + 1] * u[1] + ob.R[3 * k + 2] * u[2];
        p += sigma * r.normal();
        out[3 * i + k] = (float)p;
      }
    }
  }
  });
  delete[] objs;
  return 0;
}
