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
// stochastically rounded store) on it, writing those pages there. Given
// argv[5], it builds a small pressure hierarchy (a tank with a solid box,
// two block levels), runs two V-cycles, one relax() and the residual
// multiply, and writes the levels, cell types, X and R there.
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
    if (argc > 5) {  // relax / multiplyLaplacianMatrixOpt (msbg3.cpp:1756, 2244)
      const int nbx = 4, nby = 4, nbz = 2, b0 = 16, sx = nbx * b0, sy = nby * b0, sz = nbz * b0;
      std::vector<int8_t> lev(nbx * nby * nbz);
      for (int b = 0; b < (int)lev.size(); b++) lev[b] = (b % nbx) < 2 ? 0 : 1;
      std::vector<uint8_t> typ((size_t)sx * sy * sz, MSBG_PS_FLUID);
      for (int z = 0; z < sz; z++)
        for (int y = 0; y < sy; y++)
          for (int x = 0; x < sx; x++) {
            uint8_t &t = typ[x + (size_t)sx * (y + (size_t)sy * z)];
            if (z >= 24) t = MSBG_PS_VOID;
            if (x >= 20 && x < 28 && y >= 20 && y < 44 && z >= 4 && z < 14) t = MSBG_PS_SOLID;
          }
      const int dbc[3][2] = {{MSBG::DBC_OPEN, MSBG::DBC_OPEN},
                             {MSBG::DBC_SOLID, MSBG::DBC_SOLID},
                             {MSBG::DBC_SOLID, MSBG::DBC_OPEN}};
      MSBG::PressureGrid ps(nbx, nby, nbz, b0, 1.0 / sx, dbc, lev.data(), typ.data(), 1);
      std::vector<float> B((size_t)ps.channelSize(0));
      for (size_t i = 0; i < B.size(); i++) B[i] = 1e-4f * (float)((int)(i % 13) - 6);
      ps.setChannel(MSBG::PS_CH_B, 0, B);
      ps.vCycle(2, 2, 20);
      ps.vCycle(2, 2, 20);
      ps.relax(0, 0, 0, MSBG::PS_CH_X, MSBG::PS_CH_B, MSBG::PS_CH_T, 1);
      ps.multiplyLaplacianMatrixOpt(MSBG::OPT_CALC_RESIDUAL, 0, MSBG::PS_CH_X, MSBG::PS_CH_B, 0.0f,
                                    MSBG::PS_CH_R, nullptr);
      std::vector<float> X = ps.getChannel(MSBG::PS_CH_X, 0), R = ps.getChannel(MSBG::PS_CH_R, 0);
      FILE *fp = std::fopen(argv[5], "wb");
      std::fwrite(lev.data(), 1, lev.size(), fp);
      std::fwrite(typ.data(), 1, typ.size(), fp);
      std::fwrite(X.data(), 4, X.size(), fp);
      std::fwrite(R.data(), 4, R.size(), fp);
      std::fclose(fp);
    }
    std::printf("facade_demo: %zu active blocks\n", activeBlocks.size());
    MSBG::MultiresSparseGrid::destroy(msbg);
  } catch (const MSBG::Error &e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
