// The demo's hot path (msbg_demo.cpp:339-727) written against the C++
// facade, the way a reference caller would: create, activate, refinement
// map, allocate, splat, finalize, applyChannelPdeFast, then render_scene's
// ray-march render (msbg_demo.cpp:129-185). Writes the final UNORM16 pages
// of the active blocks to argv[1] (raw, block order of the active list)
// and, given argv[2], a 320x180 frame of BMP_MKRGB words to it, for
// comparison with the Python binding. Given argv[3], it also copies the
// field into a float channel, runs applyChannelPdeCnf's Taubin-smoothed
// laplacian flow on it and writes those pages there, and checks that a
// storage type other than the channel's is rejected. Given argv[4], it
// converts the field to UNORM8 and runs applyChannelPdeFast<uint8_t> (the
// stochastically rounded store) on it, writing those pages there.
#include <cstdio>
#include <vector>

#include "../../include/msbg/msbg.hpp"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: facade_demo OUT\n");
    return 2;
  }
  const int S = 128, bsx = 16;
  std::vector<float> xyz(3 * 6 * 4000);
  MSBG::check(msbg_scene_spheres(99, 6, 30, 98, 10, 25, 4000, 0.5f, 0.3f, xyz.data()));
  try {
    MSBG::MultiresSparseGrid *msbg = MSBG::MultiresSparseGrid::create(
        "TEST", S, S, S, bsx, -1, 0, -1, MSBG::OPT_SINGLE_LEVEL);
    float rParticle = 2, nbDist = 2, rScan = rParticle + nbDist;
    msbg->uploadParticles(xyz.data(), (int64_t)xyz.size() / 3);
    std::vector<uint32_t> blockActive = msbg->activateBlocks(rScan);
    std::vector<int> blockLevels(msbg->nBlocks());
    for (int i = 0; i < msbg->nBlocks(); i++)
      blockLevels[i] = blockActive[i] ? 0 : msbg->getNumLevels() - 1;
    msbg->regularizeRefinementMap(blockLevels.data());
    msbg->setRefinementMap(blockLevels.data());
    std::vector<int> activeBlocks = msbg->allocateLevel0Blocks<uint16_t>(MSBG::CH_UINT16_1);
    msbg->splatParticles(MSBG::CH_UINT16_1, rParticle, nbDist);
    msbg->finalizeDensity(MSBG::CH_UINT16_1, rParticle, nbDist);
    msbg->applyChannelPdeFast<uint16_t>(MSBG::CH_UINT16_1, MSBG::CH_NULL, MSBG::CH_NULL,
                                        &activeBlocks,
                                        -(MSBG::PDE_MEAN_CURVATURE + MSBG::OPT_8_COLOR_SCHEME),
                                        1, 0, 0, 0, 5, 1.0f, 0.1f);
    FILE *f = std::fopen(argv[1], "wb");
    std::vector<uint16_t> blk(bsx * bsx * bsx);
    for (int bid : activeBlocks) {
      msbg->getBlockData<uint16_t>(MSBG::CH_UINT16_1, bid, blk.data());
      std::fwrite(blk.data(), 2, blk.size(), f);
    }
    std::fclose(f);
    if (argc > 2) {
      float camPos[3] = {.5f, .8f, .7f}, camLookAt[3] = {.5, .5, .5}, camLight[3] = {.5, 5, 5};
      SBG::SparseGrid<uint16_t> *sg = msbg->getUint16Channel(MSBG::CH_UINT16_1, 0);
      RDR::RaymarchRenderer renderer(320, 180, sg);
      renderer.setCamera({camPos[0], camPos[1], camPos[2]},
                         {camLookAt[0], camLookAt[1], camLookAt[2]}, 1.0f);
      renderer.setSunLight({camLight[0], camLight[1], camLight[2]});
      renderer.setSurfaceColor({0.8f, 0.6f, 0.4f});
      std::vector<uint32_t> img = renderer.render();
      FILE *fi = std::fopen(argv[2], "wb");
      std::fwrite(img.data(), 4, img.size(), fi);
      std::fclose(fi);
    }
    if (argc > 3) {
      // applyChannelPdeCnf: laplacian flow (laplTyp 1) in 8 colours with
      // Taubin steps, the reference's configured smoothing pass
      msbg->copyChannel<float>(MSBG::CH_UINT16_1, MSBG::CH_FLOAT_1);
      MSBG::WTURB_LapSmConf cnf{};
      cnf.typ = -(1 + (int)MSBG::OPT_8_COLOR_SCHEME);
      cnf.iter = 4;
      cnf.dt = 0.05f;
      cnf.taubin = 0.5f;
      msbg->applyChannelPdeCnf(MSBG::CH_FLOAT_1, &cnf);
      FILE *ff = std::fopen(argv[3], "wb");
      std::vector<float> fb(bsx * bsx * bsx);
      for (int bid : activeBlocks) {
        msbg->getBlockData<float>(MSBG::CH_FLOAT_1, bid, fb.data());
        std::fwrite(fb.data(), 4, fb.size(), ff);
      }
      std::fclose(ff);
      try {  // float is not the UNORM16 channel's type
        msbg->applyChannelPdeFast<float>(MSBG::CH_UINT16_1, MSBG::CH_NULL, MSBG::CH_NULL,
                                         nullptr,
                                         -(MSBG::PDE_MEAN_CURVATURE + MSBG::OPT_8_COLOR_SCHEME),
                                         1, 0, 0, 0, 2, 1.0f, 0.1f);
        std::fprintf(stderr, "storage mismatch not rejected\n");
        return 1;
      } catch (const MSBG::Error &e) {
        if (e.code != MSBG_EINVAL) return 1;
      }
    }
    if (argc > 4) {  // applyChannelPdeFast<uint8_t> (msbg4.cpp:1008-1030)
      msbg->copyChannel<uint8_t>(MSBG::CH_UINT16_1, MSBG::CH_UINT8_1);
      msbg->applyChannelPdeFast<uint8_t>(MSBG::CH_UINT8_1, MSBG::CH_NULL, MSBG::CH_NULL,
                                         &activeBlocks,
                                         -(MSBG::PDE_MEAN_CURVATURE + MSBG::OPT_8_COLOR_SCHEME),
                                         1, 0, 0, 0, 4, 1.0f, 0.1f);
      FILE *f8 = std::fopen(argv[4], "wb");
      std::vector<uint8_t> b8(bsx * bsx * bsx);
      for (int bid : activeBlocks) {
        msbg->getBlockData<uint8_t>(MSBG::CH_UINT8_1, bid, b8.data());
        std::fwrite(b8.data(), 1, b8.size(), f8);
      }
      std::fclose(f8);
    }
    std::printf("facade_demo: %zu active blocks\n", activeBlocks.size());
    MSBG::MultiresSparseGrid::destroy(msbg);
  } catch (const MSBG::Error &e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
