/*
 * msbg_oracle.h -- CPU restatement of MSBG's sparse-block stencil hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker, not the product: only the
 * test suite, the smoke check and the benchmark's CPU-baseline leg may load
 * it. The product path (libmsbg_gpu.so) never calls into it.
 *
 * Every function restates the behaviour of a reference function; the
 * reference file:line it follows is cited at each declaration (paths are
 * relative to the reference's src/ directory).
 *
 * Storage codes (shared with the product's C-ABI):
 *   1 = UNORM16 (uint16, value = u/65535, the reference's RenderDensity)
 *   2 = FP32, 3 = FP16 (IEEE binary16), 4 = BF16, 5 = FP64
 *
 * Block map codes (per block id of one resolution level):
 *   >= 0  page index of an owned value block
 *   -1    NULL block (no data at this level)
 *   -2    const EMPTY block (reads as the level's empty value everywhere)
 *   -3    const FULL block (reads as the level's full value everywhere)
 *
 * Page layout: pages[page * bsx^3 + vid], vid = (vz*bsx + vy)*bsx + vx
 * (sbg.h:1870-1889, getVoxelInBlockIndex).
 */
#ifndef MSBG_ORACLE_H
#define MSBG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_U16 = 1, OR_F32 = 2, OR_F16 = 3, OR_BF16 = 4, OR_F64 = 5, OR_U8 = 6 };
enum { OR_NULL = -1, OR_EMPTY = -2, OR_FULL = -3 };

/* halo options, values of msbg.h:110,113,126 */
enum { OR_OPT_INTERPOLATE = 1 << 7, OR_OPT_BC_NEUMANN = 1 << 10,
       OR_OPT_BC_COARSE_LEVEL = 1 << 26 };

/* block flags, values of msbg.h:561-573 */
enum { OR_BLK_DOM_BORDER = 1 << 1, OR_BLK_COARSE_FINE = 1 << 2,
       OR_BLK_FINE_COARSE = 1 << 3, OR_BLK_EXISTS = 1 << 12,
       OR_BLK_UNIFORM_EMPTY = 1 << 13 };

/* One resolution level of one channel. */
typedef struct {
  int sx, sy, sz;        /* voxel resolution of this level */
  int bsx;               /* block edge at this level (power of two) */
  int nbx, nby, nbz;     /* block grid (identical on every level) */
  const int32_t *page;   /* [nbx*nby*nbz] block map codes */
  void *pages;           /* page pool of this level */
  int storage;           /* OR_U16 .. OR_U8 */
  double empty_raw;      /* raw storage value of the EMPTY block */
  double full_raw;       /* raw storage value of the FULL block */
  uint32_t sr_step;      /* UNORM8: sweep step of the next call's first sweep */
} or_level;

/* ---- scalar conversions (sbg.h:225-366, vectorf256.h:1064-1068) ---- */
float or_u16_to_float(uint16_t u);                 /* u * (1.f/65535) */
float or_u8_to_float(uint8_t u);                   /* u * (1.f/255) */
/* The UNORM8 sweep store's stochastic rounding (sbg.h:324-340): the uniform
 * number of (sweep step, page, voxel), and floor(255 f) + (u < frac). The
 * numbers are the library's counter-based stream (the reference draws from
 * per-thread Lehmer streams, fastmath.h:68-73). */
float or_sr_uniform(uint32_t step, uint32_t page, uint32_t vid);
uint8_t or_u8_from_float_sr(float f, float u);
uint16_t or_u16_from_float_roundf(float f);        /* roundf(f*65535), scalar path */
uint16_t or_u16_from_float_rne(float f);           /* round-half-even, SIMD store path */
uint16_t or_f16_from_float(float f);               /* IEEE binary16, RNE */
float or_f16_to_float(uint16_t h);
uint16_t or_bf16_from_float(float f);              /* bfloat16, RNE */
float or_bf16_to_float(uint16_t h);

/* ---- A1: domain filter (sbg.cpp:289-316, msbg_demo.cpp:356-360) ---- */
void or_domain_bounds(int sx, int sy, int sz, float dmin[3], float dmax[3]);
/* keep[i] = 1 iff point i passes isInDomainRange and trunc(pos) inRange */
int64_t or_domain_filter(int sx, int sy, int sz, const float *xyz, int64_t n,
                         uint8_t *keep);

/* ---- A2/A3: block activation and home blocks (msbg_demo.cpp:368-389) ----
 * count[bid] += 1 for every block in the trunc box of each kept point;
 * home[i] = home block id of point i (-1 if the point is not kept). */
void or_activate(int bsx, int nbx, int nby, int nbz, const float *xyz,
                 const uint8_t *keep, int64_t n, float rScan, uint32_t *count,
                 int32_t *home);

/* ---- A4: demo level rule + regularizeRefinementMap (msbg_demo.cpp:418-428,
 * msbg.cpp:1202-1423). in and out may alias. */
void or_levels_from_activity(const uint32_t *count, int64_t nblocks,
                             int nLevels, int *levels);
int or_regularize(int nbx, int nby, int nbz, int nLevels, const int *in,
                  int *out);

/* ---- A5: setRefinementMap block flags (msbg.cpp:1615-1745), levelMg 0.
 * inFlags (may be NULL) is the caller's blockFlags[] input. */
void or_block_flags(int nbx, int nby, int nbz, int nLevels, const int *levels,
                    const uint32_t *inFlags, uint16_t *level_out,
                    uint16_t *flags_out);

/* ---- A7: colour (msbg.h:642-646) and per-colour lists (msbg_demo.cpp:527-539)
 * lists: out[] holds the blocks of colour 0, then 1, ..., in list order;
 * off[c]..off[c+1] delimits colour c. */
int or_color8(int bx, int by, int bz);
void or_color_lists(int nbx, int nby, const int32_t *blocks, int n,
                    int32_t *out, int32_t off[9]);

/* ---- A8: 8-colour min-distance splat into UNORM16 pages
 * (msbg_demo.cpp:541-636). Points are visited per colour, per home block in
 * the order of `blocks`, in list order inside a block -- the result does not
 * depend on that order (min), except through the non-value-block abandon. */
void or_splat_u16(const or_level *L, const float *xyz, const int32_t *home,
                  int64_t n, float rScan);

/* ---- A9: finalize distances to density (msbg_demo.cpp:655-680) ---- */
uint16_t or_finalize_value(uint16_t u, float rParticle, float nbDist);
float or_finalize_density_f(uint16_t u, float rParticle, float nbDist);
void or_finalize_u16(uint16_t *pages, int64_t nvox, float rParticle,
                     float nbDist);

/* ---- A11/A13: halo fill of block bid into a float (or double) buffer of
 * (bsx+2*dmax)^3 (halo.cpp:267-479, halo.h:267-355,422-501,
 * sbg.h:3002-3095). Lo may be NULL unless OPT_BC_COARSE_LEVEL is set. */
/* SparseGrid<Data_T>::interpolateGhost(p, 0) (sbg.h:3002-3095) of level Lo
 * at the position (px,py,pz) in Lo's voxel coordinates, in Lo's raw domain
 * (UNORM16: the truncated integer of each interpolation stage). */
double or_interpolate_ghost(const or_level *Lo, float px, float py, float pz);
void or_halo_fill(const or_level *L, const or_level *Lo, int bid, int dmax,
                  unsigned opt, float *out);
void or_halo_fill_f64(const or_level *L, int bid, int dmax, unsigned opt,
                      double *out);

/* ---- A10/A12: 8-colour in-place mean-curvature-flow sweeps
 * (msbg4.cpp:60-957). `blocks` is the caller's block list, `levels` the
 * per-block resolution level (only level-0 blocks are updated,
 * msbg4.cpp:467-470). numIter is rounded up to even (msbg4.cpp:150-152).
 * Arithmetic is FP32 for every storage except FP64, which runs the whole
 * sweep in double. residual (optional, [nSweeps]) receives the FP64 RMS of
 * the per-sweep update over the swept voxels. Returns the sweep count. */
int or_mcf_sweeps(const or_level *L, const int32_t *blocks, int n,
                  const uint16_t *levels, int numIter, float dt,
                  int doConstrZeroOne, double *residual, int nthreads);
/* Coarse-level sweeps: an extension of the sweep above that the reference
 * does not run (it updates level 0 only, msbg4.cpp:467-470). L is level
 * sweepLevel of the channel (bsx_l blocks, its pages including the ghost
 * pages prepareGhostBlocks writes for finer neighbours); the blocks of
 * `blocks` whose level is sweepLevel are swept in 8 colour passes, each
 * from the halo fillHaloBlock_ builds with `opt`: OR_OPT_BC_NEUMANN (the
 * level-0 rule), or OR_OPT_BC_COARSE_LEVEL [| OR_OPT_INTERPOLATE], where
 * NULL neighbours read level sweepLevel+1 (Lo) and out-of-domain voxels are
 * clipped first (halo.h:280-352). FP64 supports NEUMANN only. Returns the
 * sweep count, -1 for an unsupported combination. */
int or_mcf_sweeps_level(const or_level *L, const or_level *Lo, const int32_t *blocks, int n,
                        const uint16_t *levels, int sweepLevel, unsigned opt, int numIter,
                        float dt, int doConstrZeroOne, double *residual, int nthreads);
/* The per-voxel MCF update from a halo buffer (msbg4.cpp:559-637,813-825). */
float or_mcf_voxel(const float *h, int hsx, int hid, float dt,
                   int doConstrZeroOne);
double or_mcf_voxel_f64(const double *h, int hsx, int hid, double dt,
                        int doConstrZeroOne);

/* ---- the other stencils and driver modes of applyChannelPdeFast
 * (msbg4.cpp:60-957). op is laplTyp with the mode flags removed:
 *   1  laplacian flow  D = dt*L*|grad|, Taubin on odd/even sweeps  (:682-712)
 *   2  bi-laplacian    D = clamp(-dt*H*|grad|, +-0.1), halo 2      (:637-680)
 *   3  diffusion       phi += dt*(mean6 - phi)                      (:713-735)
 *   4  mean curvature  (the hot path above)
 *   14 Perona-Malik    phi += dt*sum G(g_i) g_i                      (:736-790)
 *   15 Perona-Malik, Jacobi form                                      (:736-790)
 *   17 diffusion, Jacobi form                                         (:791-805)
 * Ops 1, 3 and 17 read neighbours at fixed offsets 18 and 18*18 in the
 * reference, the halo strides of 16^3 blocks; the restatement reads the true
 * neighbours, which is the same thing at bsx 16 only.
 * mode: colour (8 passes by parity bits, in place), red-black (2 passes;
 * pass p updates blocks with (bx+by+bz)&1 != p, in place) or Jacobi (through
 * a temporary channel whose value blocks that are not swept are NULL, as
 * the reference's reset temporary is; an odd numIter makes the last sweep a
 * copy). Hard constraints (densThr or densThrTol1 > 1e-12) clamp to
 * [max(0, phi0 - tol1), min(1, phi0 + tol2)] around the field at entry,
 * else doConstrZeroOne clamps to [0, 1]. FP32 arithmetic; FP64 channels
 * support op 4 only. Returns the sweep count, or -1 for an unsupported
 * combination. */
enum { OR_MODE_COLOR8 = 0, OR_MODE_REDBLACK = 1, OR_MODE_JACOBI = 2 };
typedef struct {
  int op, mode;
  float dt;
  int doConstrZeroOne;
  float densThr, densThrTol1, densThrTol2;
  float taubin;  /* op 1: odd sweeps add D, even ones subtract taubin*D */
  float alpha;   /* ops 14/15: edge scale; < 0 selects the exponential */
} or_pde_params;
int or_pde_sweeps(const or_level *L, const int32_t *blocks, int n, const uint16_t *levels,
                  int numIter, const or_pde_params *P, double *residual, int nthreads);
/* one voxel of op (1, 2, 3, 14, 15, 17) from a halo buffer of width hw */
float or_pde_voxel(int op, const float *h, int hsx, int hid, float dt, float taubin,
                   float alpha, int iter);
/* FMA_FAST_EXP (fastmath.h:583-628), including its repeated PX2 term */
float or_fast_exp(float x);

/* ---- A14: 4x4x4 ghost downsample of one fine halo block into a coarse
 * block of bsxHi/2 edge (msbg.cpp:298-343, 2902-2974). */
void or_ghost_downsample(const float *halo, int bsxHi, float *lo);
/* prepareGhostBlocks (msbg.cpp:3059-3170) for FP32 channels on levels
 * 0..nLevels-2, given the fine-coarse block lists per level. lvl[l] are the
 * channel's levels; ghost pages are written into lvl[l+1]'s page map,
 * which must already own a page for each fine-coarse block of level l
 * that has BLK_EXISTS, and hold OR_EMPTY for the others. */
void or_prepare_ghost_blocks(or_level *lvl, int nLevels, const int32_t *fc,
                             const int32_t *fc_off, const uint16_t *flags);

/* ---- field consumers of a level (msbg.h:469-484, render.h, render.cpp).
 * Positions are voxel coordinates of the level (cell centres at i + 0.5).
 * Results are in the interpolant's units: raw UNORM16 integers for UNORM16
 * levels, decoded values otherwise (the caller scales by 1/65535). The
 * contraction of multiply-adds into FMAs, which the reference leaves to
 * its compiler (-ffast-math), is fixed as stated in the source. */
/* interpolateUint16ToFloatFast / interpolateFloatFast (sbg.h:4623-4800,
 * 4920-5031), with the border fallback of interpolateGen0 (sbg.h:3170-3245) */
float or_interp_linear(const or_level *L, const float p[3]);
float or_interp_linear_border(const or_level *L, const float p[3]);
/* MtCubicBSplineWeights1D (mtool.h:344-363) */
void or_bspline3_weights(float t, float w[4], float d[4]);
/* interpolateWithDerivs0, IP_BSPLINE_CUBIC, options 0 (sbg.h:3903-4135) */
void or_interp_bspline3(const or_level *L, const float p[3], float *val, float grad[3]);
/* RaymarchRenderer (render.h:23-160, render.cpp:19-100): camera, sun and
 * surface colour as the demo passes them (msbg_demo.cpp:122-166), widened
 * to double. rgb[(y-y0)*W + x] = BMP_MKRGB(r,g,b) (bitmap.h:545-546) for
 * rows y0..y1-1 of a W x H image. */
typedef struct {
  double pos[3], look[3];
  double focal;
  double sun[3], color[3];
} or_camera;
void or_camera_basis(const double pos[3], const double look[3], double fwd[3],
                     double right[3], double up[3]);
void or_render(const or_level *L, const or_camera *C, int W, int H, int y0, int y1,
               uint32_t *rgb, int nthreads);
/* batch form: kind 0 = or_interp_linear (val), 1 = or_interp_bspline3
 * (val, grad[3*i+k]; grad may be NULL) */
void or_interp_points(const or_level *L, int kind, const float *pts, int64_t n, float *val,
                      float *grad);

/* ---- misc ---- */
uint64_t or_version(void);

#ifdef __cplusplus
}
#endif
#endif
