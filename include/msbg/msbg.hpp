// msbg.hpp -- header-only C++ facade over the C ABI of libmsbg_gpu.so with
// the reference's MultiresSparseGrid names for the sparse-block stencil path.
//
// A caller of the reference's
//   MSBG::MultiresSparseGrid::create(...), regularizeRefinementMap(...),
//   setRefinementMap(...), applyChannelPdeFast<T>(...), prepareGhostBlocks(...)
// (msbg.h:834-3907; paths relative to the reference's src/ directory) swaps
// the include and keeps the call sites. Errors throw msbg::Error carrying the
// C ABI's code and message (the reference's rcThis/raiseRc pattern,
// globdef.h:216, becomes an exception here).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../msbg_gpu.h"

namespace MSBG {

enum : unsigned {
  OPT_INTERPOLATE = MSBG_OPT_INTERPOLATE,
  OPT_BC_NEUMANN = MSBG_OPT_BC_NEUMANN,
  OPT_SINGLE_LEVEL = MSBG_OPT_SINGLE_LEVEL,
  OPT_BC_COARSE_LEVEL = MSBG_OPT_BC_COARSE_LEVEL,
  OPT_8_COLOR_SCHEME = MSBG_OPT_8_COLOR_SCHEME,
};
enum { PDE_MEAN_CURVATURE = MSBG_PDE_MEAN_CURVATURE };
// channel ids (msbg.h:274-343)
enum {
  CH_NULL = MSBG_CH_NULL,
  CH_FLOAT_1 = MSBG_CH_FLOAT_1,
  CH_FLOAT_2 = 6,
  CH_FLOAT_3 = 7,
  CH_UINT16_1 = MSBG_CH_UINT16_1,
  CH_UINT8_1 = 38,
};

// WTURB_LapSmConf (wturbglob.h:217-233): the smoothing pass that
// applyChannelPdeCnf runs
typedef struct {
  int typ, use_pres_grad_mag;
  int iter;
  float thr, thr_tol1, thr_tol2;
  float dt;
  float velmag_thr;
  float lambda;
  float unused;
  float taubin;
  float lambda_sdf;
  float redist;
} WTURB_LapSmConf;

struct BlockInfo {  // msbg.h:673-678
  uint16_t level, flags;
};

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
  int code;
};

inline void check(int rc) {
  if (rc) throw Error(rc, msbg_last_error());
}

template <typename T>
struct StorageOf;
template <>
struct StorageOf<uint16_t> {
  static constexpr int value = MSBG_STORAGE_U16;
};
template <>
struct StorageOf<uint8_t> {
  static constexpr int value = MSBG_STORAGE_U8;
};
template <>
struct StorageOf<float> {
  static constexpr int value = MSBG_STORAGE_F32;
};
template <>
struct StorageOf<double> {
  static constexpr int value = MSBG_STORAGE_F64;
};
// The 2-byte float storages have no type of their own in the reference
// (its channels are float, uint16_t or uint8_t): these tags name them.
struct half16 {
  uint16_t bits;
};
struct bfloat16 {
  uint16_t bits;
};
template <>
struct StorageOf<half16> {
  static constexpr int value = MSBG_STORAGE_F16;
};
template <>
struct StorageOf<bfloat16> {
  static constexpr int value = MSBG_STORAGE_BF16;
};

// getBlockColor8 (msbg.h:642-646)
static inline int getBlockColor8(int bx, int by, int bz) {
  return (bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2);
}

class MultiresSparseGrid;

}  // namespace MSBG

namespace SBG {

// One channel level of a MultiresSparseGrid, in the role of the reference's
// SBG::SparseGrid<T> (sbg.h) as the renderer and interpolators use it. Owned
// by its MultiresSparseGrid.
template <typename Data_T>
class SparseGrid {
 public:
  SparseGrid(msbg_grid *g, int chan, int level, int sx, int sy, int sz)
      : g_(g), chan_(chan), level_(level), sx_(sx), sy_(sy), sz_(sz) {}
  int sx() const { return sx_; }
  int sy() const { return sy_; }
  int sz() const { return sz_; }
  int64_t sxyzMax() const { return sx_ > sy_ ? (sx_ > sz_ ? sx_ : sz_) : (sy_ > sz_ ? sy_ : sz_); }
  // interpolateUint16ToFloatFast / interpolateFloatFast (sbg.h:4623-5031) at
  // n positions xyz[3*i..] (voxel coordinates), batched: one launch
  void interpolateFast(const float *xyz, int64_t n, float *out) const {
    MSBG::check(msbg_interpolate(g_, chan_, level_, MSBG_IP_LINEAR, xyz, n, out, nullptr));
  }
  // interpolateWithDerivs<0>(IP_BSPLINE_CUBIC, ...) (sbg.h:3903-4135), batched
  void interpolateBSplineWithDerivs(const float *xyz, int64_t n, float *val, float *grad) const {
    MSBG::check(msbg_interpolate(g_, chan_, level_, MSBG_IP_BSPLINE3, xyz, n, val, grad));
  }
  msbg_grid *grid() const { return g_; }
  int channel() const { return chan_; }
  int level() const { return level_; }

 private:
  msbg_grid *g_;
  int chan_, level_, sx_, sy_, sz_;
};

}  // namespace SBG

namespace MSBG {

class MultiresSparseGrid {
 public:
  // MultiresSparseGrid::create (msbg.h:834-850). dTransRes, maxSizeMB and
  // initialSizeMB do not change this path; the solver configs are unused.
  static MultiresSparseGrid *create(const char *name, int sx, int sy, int sz, int blockSize0 = 16,
                                    int /*dTransRes*/ = -1, int /*maxSizeMB*/ = 0,
                                    int /*initialSizeMB*/ = -1, unsigned options = 0,
                                    void * /*miscConf*/ = nullptr, void * /*pConf*/ = nullptr,
                                    int nLevels = -1) {
    msbg_grid *g = nullptr;
    check(msbg_create(name, sx, sy, sz, blockSize0, options, nLevels, &g));
    return new MultiresSparseGrid(g);
  }
  // MultiresSparseGrid::destroy (msbg.h:851-852)
  static void destroy(MultiresSparseGrid *&m) {
    delete m;
    m = nullptr;
  }
  ~MultiresSparseGrid() { msbg_destroy(g_); }
  MultiresSparseGrid(const MultiresSparseGrid &) = delete;
  MultiresSparseGrid &operator=(const MultiresSparseGrid &) = delete;

  int nBlocks() const { return (int)nblocks_; }
  int getNumLevels() const { return dims_[7]; }
  int nbx() const { return dims_[4]; }
  int nby() const { return dims_[5]; }
  int nbz() const { return dims_[6]; }
  int bsx0() const { return dims_[3]; }
  msbg_grid *handle() const { return g_; }

  void regularizeRefinementMap(int *blockmapIn, int *blockmapOut = nullptr,
                               unsigned * /*blockFlags*/ = nullptr, int /*rbzWidthBlk*/ = 0) {
    check(msbg_regularize_refinement_map(g_, blockmapIn, blockmapOut ? blockmapOut : blockmapIn));
  }
  void setRefinementMap(int *blockLevels, unsigned *blockFlags = nullptr, int /*levelMgMax*/ = -1,
                        MultiresSparseGrid * /*msbgObstacles*/ = nullptr,
                        bool /*initCellFlags*/ = true, bool /*resetChannels*/ = true) {
    check(msbg_set_refinement_map(g_, blockLevels, reinterpret_cast<const uint32_t *>(blockFlags)));
    refreshBlockInfo();
  }
  BlockInfo *getBlockInfo(int bid, int /*levelMg*/ = 0) { return &info_[bid]; }

  void prepareGhostBlocks(int chan) { check(msbg_prepare_ghost_blocks(g_, chan)); }

  // applyChannelPdeFast<T> (msbg.h:2150-2172, msbg4.cpp:60-957) with the
  // reference's parameter list. Data_T must be the channel's storage
  // (StorageOf); laplTyp, lambda, taubin and anisDiffAlpha go to
  // msbg_apply_channel_pde, which rejects the operators and modes outside
  // this path. The velocity-dependent, SDF and redistancing terms
  // (velMagThr, lambdaSDF, redist, veldep_on) are outside it too: non-zero
  // values throw MSBG_EUNSUPP instead of being dropped. residual (an
  // addition) receives one FP64 RMS update per sweep.
  template <typename Data_T>
  void applyChannelPdeFast(int chPhi, int chTmp, int chTmp2, std::vector<int> *blockList,
                           int laplTyp, int doConstrZeroOne, float densThr, float densThrTol1,
                           float densThrTol2, int numIter, float gridSpacing = 1,
                           float timeStep = 0.025f, float velMagThr = 0,
                           MultiresSparseGrid * /*msbgVel*/ = nullptr, float lambda = 0,
                           float lambdaSDF = 0, float taubin = 0, float redist = 0,
                           float anisDiffAlpha = 0, int veldep_on = 0,
                           float * /*veldep_fit*/ = nullptr, int /*visOpt*/ = 0,
                           double *residual = nullptr) {
    int st = 0;
    check(msbg_channel_storage(g_, chPhi, &st));
    if (st != StorageOf<Data_T>::value)
      throw Error(MSBG_EINVAL, "applyChannelPdeFast<T>: channel " + std::to_string(chPhi) +
                                   " has storage " + std::to_string(st) + ", T names storage " +
                                   std::to_string(StorageOf<Data_T>::value));
    if (velMagThr != 0 || lambdaSDF != 0 || redist != 0 || veldep_on != 0)
      throw Error(MSBG_EUNSUPP,
                  "applyChannelPdeFast: velMagThr, lambdaSDF, redist and veldep_on are outside "
                  "this path (must be 0)");
    check(msbg_apply_channel_pde(g_, chPhi, chTmp, chTmp2, blockList ? blockList->data() : nullptr,
                                 blockList ? (int64_t)blockList->size() : 0, laplTyp,
                                 doConstrZeroOne, densThr, densThrTol1, densThrTol2, numIter,
                                 gridSpacing, timeStep, lambda, taubin, anisDiffAlpha, residual));
  }

  // applyChannelPdeCnf (msbg.h:2174-2195): the configured smoothing pass on
  // a float channel with CH_FLOAT_2 / CH_FLOAT_3 as its temporaries
  void applyChannelPdeCnf(int chan, WTURB_LapSmConf *cnf) {
    applyChannelPdeFast<float>(chan, CH_FLOAT_2, CH_FLOAT_3, nullptr, cnf->typ, 1, cnf->thr,
                               cnf->thr_tol1, cnf->thr_tol2, cnf->iter, 1.0f, cnf->dt,
                               cnf->velmag_thr, this, cnf->lambda, cnf->lambda_sdf, cnf->taubin,
                               cnf->redist, 0, 0);
  }

  // The same 8-colour MCF sweep on the blocks of one coarser level, on that
  // level's own grid (no counterpart: the reference sweeps level 0 only,
  // msbg4.cpp:467-470). haloOpt is OPT_BC_NEUMANN, OPT_BC_COARSE_LEVEL or
  // OPT_BC_COARSE_LEVEL | OPT_INTERPOLATE; residual (optional) gets one
  // FP64 RMS update per sweep.
  void applyChannelPdeLevel(int chan, int level, unsigned haloOpt, std::vector<int> *blockList,
                            int numIter, float timeStep, int doConstrZeroOne,
                            double *residual = nullptr) {
    check(msbg_apply_channel_pde_level(g_, chan, level, haloOpt,
                                       blockList ? blockList->data() : nullptr,
                                       blockList ? (int64_t)blockList->size() : 0, numIter,
                                       timeStep, doConstrZeroOne, residual));
  }

  // ---- demo phases (msbg_demo.cpp:339-680) ----
  int64_t uploadParticles(const float *xyz, int64_t n) {
    int64_t kept = 0;
    check(msbg_points_upload(g_, xyz, n, &kept));
    return kept;
  }
  std::vector<uint32_t> activateBlocks(float rScan) {
    std::vector<uint32_t> c((size_t)nblocks_);
    check(msbg_activate_from_points(g_, rScan, c.data(), nullptr));
    return c;
  }
  template <typename Data_T>
  std::vector<int> allocateLevel0Blocks(int chan) {
    std::vector<int> active((size_t)nblocks_);
    int64_t n = 0;
    check(msbg_channel_alloc_demo(g_, chan, StorageOf<Data_T>::value, active.data(),
                                  (int64_t)active.size(), &n));
    active.resize((size_t)n);
    refreshBlockInfo();
    return active;
  }
  void splatParticles(int chan, float rParticle, float nbDist) {
    check(msbg_splat_points(g_, chan, rParticle, nbDist));
  }
  void finalizeDensity(int chan, float rParticle, float nbDist) {
    check(msbg_finalize_density(g_, chan, rParticle, nbDist));
  }
  // SparseGrid<T>::copyDataTo (sbg.h:550-557) into channel chDst of storage
  // Dst_T (values decoded exactly, re-encoded with round-to-nearest-even)
  template <typename Dst_T>
  void copyChannel(int chSrc, int chDst, int level = 0) {
    check(msbg_channel_copy(g_, chSrc, chDst, StorageOf<Dst_T>::value, level));
  }
  template <typename Data_T>
  void getBlockData(int chan, int bid, Data_T *out, int level = 0) {
    check(msbg_download_block(g_, chan, level, bid, out));
  }
  // getUint16Channel / getFloatChannel (msbg.h): the channel's level grid
  SBG::SparseGrid<uint16_t> *getUint16Channel(int chan, int level) {
    return channelGrid<uint16_t>(u16_, chan, level);
  }
  SBG::SparseGrid<float> *getFloatChannel(int chan, int level) {
    return channelGrid<float>(f32_, chan, level);
  }

 private:
  explicit MultiresSparseGrid(msbg_grid *g) : g_(g) {
    check(msbg_get_dims(g_, dims_, &nblocks_));
    info_.resize((size_t)nblocks_);
    refreshBlockInfo();
  }
  void refreshBlockInfo() {
    std::vector<uint16_t> l((size_t)nblocks_), f((size_t)nblocks_);
    check(msbg_get_block_info(g_, l.data(), f.data()));
    for (size_t i = 0; i < info_.size(); i++) info_[i] = BlockInfo{l[i], f[i]};
  }
  template <typename T>
  SBG::SparseGrid<T> *channelGrid(std::map<int64_t, std::unique_ptr<SBG::SparseGrid<T>>> &m,
                                  int chan, int level) {
    auto &p = m[(int64_t)chan * 16 + level];
    if (!p)
      p.reset(new SBG::SparseGrid<T>(g_, chan, level, dims_[0] >> level, dims_[1] >> level,
                                     dims_[2] >> level));
    return p.get();
  }
  msbg_grid *g_;
  int dims_[8] = {0};
  int64_t nblocks_ = 0;
  std::vector<BlockInfo> info_;
  std::map<int64_t, std::unique_ptr<SBG::SparseGrid<uint16_t>>> u16_;
  std::map<int64_t, std::unique_ptr<SBG::SparseGrid<float>>> f32_;
};

// The pressure solver's hierarchy and its relax / multiply
// (MultiresSparseGrid::relax, multiplyLaplacianMatrixOpt, msbg3.cpp:1756,
// 2244), over msbg_ps_*. Channels are PS_CH_X / _B / _R / _T.
enum { PS_CH_X = MSBG_PS_CH_X, PS_CH_B = MSBG_PS_CH_B, PS_CH_R = MSBG_PS_CH_R,
       PS_CH_T = MSBG_PS_CH_T, PS_CH_D = MSBG_PS_CH_D };
enum { DBC_OPEN = MSBG_DBC_OPEN, DBC_SOLID = MSBG_DBC_SOLID };
enum : unsigned {
  OPT_BOUNDZONE_ONLY = MSBG_OPT_BOUNDZONE_ONLY,
  OPT_CALC_RESIDUAL = MSBG_OPT_CALC_RESIDUAL,
  OPT_REVERSE_ORDER = MSBG_OPT_REVERSE_ORDER,
};

class PressureGrid {
 public:
  // blockLevels[nbx*nby*nbz]; cellType (MSBG_PS_FLUID/SOLID/VOID) per finest
  // cell, x fastest. omega and the sub-iteration count default to the
  // reference's _mgSmOmega 6/7 and _mgSmBlockIter 5 (msbg.cpp:819, 837).
  PressureGrid(int nbx, int nby, int nbz, int bsx0, double dx0, const int domainBc[3][2],
               const int8_t *blockLevels, const uint8_t *cellType, int denseLevels = 0,
               float omega = 6.f / 7.f, int blockIter = 5) {
    msbg_ps_desc d{};
    d.nbx = nbx, d.nby = nby, d.nbz = nbz, d.bsx0 = bsx0, d.dx0 = dx0;
    for (int a = 0; a < 3; a++)
      for (int s = 0; s < 2; s++) d.domainBc[a][s] = domainBc[a][s];
    d.omega = omega, d.blockIter = blockIter, d.denseLevels = denseLevels;
    check(msbg_ps_create(&d, blockLevels, cellType, &ps_));
    check(msbg_ps_get_levels(ps_, &nmg_, &nlev_));
  }
  ~PressureGrid() { msbg_ps_destroy(ps_); }
  PressureGrid(const PressureGrid &) = delete;
  PressureGrid &operator=(const PressureGrid &) = delete;

  int getNumMgLevels() const { return nmg_; }
  int64_t channelSize(int levelMg) const {
    int64_t n = 0;
    check(msbg_ps_level_size(ps_, levelMg, &n));
    return n;
  }
  void setChannel(int chan, int levelMg, const std::vector<float> &v) {
    if ((int64_t)v.size() != channelSize(levelMg)) throw Error(MSBG_EINVAL, "channel size");
    check(msbg_ps_upload(ps_, chan, levelMg, v.data()));
  }
  std::vector<float> getChannel(int chan, int levelMg) {
    std::vector<float> v((size_t)channelSize(levelMg));
    check(msbg_ps_download(ps_, chan, levelMg, v.data()));
    return v;
  }
  void resetChannel(int chan, int levelMg) { check(msbg_ps_fill(ps_, chan, levelMg, 0.0f)); }

  void relax(int boundaryZoneOnly, int reverseOrder, int levelMg, int chanX, int chanB,
             int chanTmp, int nIter, unsigned options = 0) {
    (void)options;
    check(msbg_ps_relax(ps_, boundaryZoneOnly, reverseOrder, levelMg, chanX, chanB, chanTmp,
                        nIter));
  }
  void multiplyLaplacianMatrixOpt(unsigned options, int levelMg, int chanX, int chanB, float kap,
                                  int chanDstY, long double *pDotprod,
                                  std::vector<int> *blockList = nullptr) {
    if (blockList) throw Error(MSBG_EUNSUPP, "multiplyLaplacianMatrixOpt: block list");
    double dot = 0;
    check(msbg_ps_multiply(ps_, options, levelMg, chanX, chanB, kap, chanDstY,
                           pDotprod ? &dot : nullptr));
    if (pDotprod) *pDotprod = dot;
  }
  void restrictChannel(int levelMg, int chanSrc, int chanDst) {
    check(msbg_ps_restrict(ps_, levelMg, chanSrc, chanDst));
  }
  void prolongateChannel(int levelMg, int chanSrc, int chanDst) {
    check(msbg_ps_prolongate(ps_, levelMg, chanSrc, chanDst));
  }
  void vCycle(int nu1, int nu2, int coarseIter) { check(msbg_ps_vcycle(ps_, nu1, nu2, coarseIter)); }

 private:
  msbg_ps *ps_ = nullptr;
  int nmg_ = 0, nlev_ = 0;
};

}  // namespace MSBG

namespace RDR {

typedef double RealType;

struct Vec3 {  // render.h:23-40
  RealType x = 0, y = 0, z = 0;
  constexpr Vec3() = default;
  constexpr Vec3(RealType xx, RealType yy, RealType zz) : x(xx), y(yy), z(zz) {}
};

// RDR::RaymarchRenderer (render.h:55-160, render.cpp:19-100) over the
// library's renderer. render() returns the image as BMP_MKRGB words, row 0
// at the top, where the reference returns a BmpBitmap holding the same
// words. Camera, sun and colour are passed on as float, the precision the
// demo gives them (msbg_demo.cpp:122-166).
class RaymarchRenderer {
 public:
  RaymarchRenderer(int w, int h, SBG::SparseGrid<uint16_t> *sgDensity)
      : width_(w), height_(h), sg_(sgDensity) {
    setCamera({0, 0, -5}, {0, 0, 0}, 1.0f);
    setSunLight({10, 10, -10});
    setSurfaceColor({1, 1, 1});
  }
  void setCamera(const Vec3 &position, const Vec3 &lookAt, float focalLength) {
    set3(cam_, position);
    set3(look_, lookAt);
    focal_ = focalLength;
  }
  void setSunLight(const Vec3 &sunPos) { set3(sun_, sunPos); }
  void setSurfaceColor(const Vec3 &color) { set3(color_, color); }
  std::vector<uint32_t> render() const {
    std::vector<uint32_t> rgb((size_t)width_ * height_);
    MSBG::check(msbg_render(sg_->grid(), sg_->channel(), width_, height_, cam_, look_, focal_,
                            sun_, color_, rgb.data(), nullptr));
    return rgb;
  }
  int width() const { return width_; }
  int height() const { return height_; }

 private:
  static void set3(float *d, const Vec3 &v) {
    d[0] = (float)v.x;
    d[1] = (float)v.y;
    d[2] = (float)v.z;
  }
  int width_, height_;
  SBG::SparseGrid<uint16_t> *sg_;
  float cam_[3], look_[3], sun_[3], color_[3], focal_ = 1.0f;
};

}  // namespace RDR
