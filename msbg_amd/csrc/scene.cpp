// Seeded synthetic particle scenes on the host: jittered surface points of
// spheres and randomly rotated ellipsoids, optionally with Perlin-noise
// displaced surfaces. Each object has its own PRNG stream derived from
// (seed, object index), so the output does not depend on the number of host
// threads. Positions are computed in double and rounded once to float.

#include <cmath>
#include <cstdint>

#include "internal.h"
#include "scene.h"

using namespace msbg::scene;

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
  const Perlin noise = scene_noise(seed);
  const std::vector<Obj> objs = draw_objects(seed, nObj, cmin, cmax, rmin, rmax, ellipsoidFrac);
  msbg::parallel_for(nObj, [&](int64_t o0, int64_t o1) {
  for (int64_t o = o0; o < o1; o++) {
    const Obj &ob = objs[(size_t)o];
    Xoshiro r(ob.seed);
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
      double j[3];
      for (double &v : j) v = sigma * r.normal();
      object_point(ob, noise.p, n, j, perlinAmp, perlinFreq, out + 3 * i);
    }
  }
  });
  return 0;
}
