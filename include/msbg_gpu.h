/*
 * msbg_gpu.h -- C ABI of the MI355X-native MSBG sparse-block stencil path.
 *
 * The library (libmsbg_gpu.so) owns all device memory of a grid: the block
 * tables, the particle buffer and the page pools of every channel. Callers
 * pass plain host pointers and sizes. Every entry point returns 0 on success
 * or a negative MSBG_E* code; msbg_last_error() describes the last failure of
 * the calling thread.
 *
 * Each declaration names the reference interface it replaces. Reference
 * paths are relative to the reference's src/ directory. The reference is a
 * C++ class library (MultiresSparseGrid, SparseGrid<T>, HaloBlockSet); this
 * ABI flattens the calls the sweep/halo/splat path makes into C functions
 * on an opaque grid handle.
 */
#ifndef MSBG_GPU_H
#define MSBG_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct msbg_grid msbg_grid;

/* ---- error codes (the reference's rcThis/MI_E* pattern, globdef.h:216) ---- */
enum {
  MSBG_OK = 0,
  MSBG_EINVAL = -1,   /* invalid argument (MI_EINVARG) */
  MSBG_ENOMEM = -2,   /* device or host allocation failed (MI_ENOMEM) */
  MSBG_EDEVICE = -3,  /* HIP runtime error or no device */
  MSBG_ESTATE = -4,   /* call order violated (e.g. sweep before allocation) */
  MSBG_EUNSUPP = -5   /* option outside the implemented path */
};

/* ---- storage formats of a channel's pages ---- */
enum {
  MSBG_STORAGE_U16 = 1,  /* UNORM16: value u/65535 (RenderDensity, msbgcell.h:68) */
  MSBG_STORAGE_F32 = 2,
  MSBG_STORAGE_F16 = 3,  /* IEEE binary16, FP32 arithmetic */
  MSBG_STORAGE_BF16 = 4, /* bfloat16, FP32 arithmetic */
  MSBG_STORAGE_F64 = 5,  /* FP64 storage and arithmetic (oracle run) */
  MSBG_STORAGE_U8 = 6    /* UNORM8: value u/255 (the uint8_t RenderDensity of
                            applyChannelPdeFast<uint8_t>, msbg4.cpp:1008-1030);
                            the sweeps store it with stochastic rounding */
};

/* ---- option bits and ids with the reference's values (msbg.h:101-391) ---- */
#define MSBG_OPT_INTERPOLATE (1u << 7)
#define MSBG_OPT_BC_NEUMANN (1u << 10)
#define MSBG_OPT_SINGLE_LEVEL (1u << 14)
#define MSBG_OPT_BC_COARSE_LEVEL (1u << 26)
#define MSBG_OPT_8_COLOR_SCHEME (1u << 29)
#define MSBG_PDE_MEAN_CURVATURE 4
#define MSBG_CH_NULL 0
#define MSBG_CH_FLOAT_1 1
#define MSBG_CH_UINT16_1 30

/* block flags (msbg.h:557-575) */
#define MSBG_BLK_DOM_BORDER (1u << 1)
#define MSBG_BLK_COARSE_FINE (1u << 2)
#define MSBG_BLK_FINE_COARSE (1u << 3)
#define MSBG_BLK_EXISTS (1u << 12)
#define MSBG_BLK_UNIFORM_EMPTY (1u << 13)

/* block map codes returned by msbg_channel_get_pagemap / halo tables */
#define MSBG_BLOCK_NULL (-1)  /* no block at this level (_blockmap[bid]==NULL) */
#define MSBG_BLOCK_EMPTY (-2) /* const EMPTY block (sbg.h:2042-2047) */
#define MSBG_BLOCK_FULL (-3)  /* const FULL block */
#define MSBG_BLOCK_OUT (-4)   /* outside the domain (halo tables only) */

/* block kinds accepted by msbg_channel_define */
#define MSBG_KIND_NULL 0
#define MSBG_KIND_EMPTY 1
#define MSBG_KIND_FULL 2
#define MSBG_KIND_VALUE 3

/* arithmetic mode of the sweep (msbg_set_arith) */
#define MSBG_ARITH_EXACT 0 /* reference op order, no FMA contraction, IEEE divide */
#define MSBG_ARITH_FAST 1  /* FMA-contracted, reciprocal divide */

/* ======================================================================== */
/* grid lifetime                                                            */
/* ======================================================================== */

/* MultiresSparseGrid::create(name, sx, sy, sz, blockSize0, dTransRes,
 * maxSizeMB, initialSizeMB, options, ..., nLevels) (msbg.h:834-850,
 * msbg.cpp:117-1106). Block-aligned domain: sx,sy,sz must be multiples of
 * blockSize0 (sbg.cpp:155-163); blockSize0 in {8,16,32,64}; nLevels in
 * [1,4] (MSBG_MAXRESLEVELS, msbgcell.h:147), -1 = 3 (msbg.cpp:137).
 * options: MSBG_OPT_SINGLE_LEVEL or 0. */
int msbg_create(const char *name, int sx, int sy, int sz, int blockSize0,
                unsigned options, int nLevels, msbg_grid **out);

/* MultiresSparseGrid::destroy (msbg.h:851-852). */
int msbg_destroy(msbg_grid *g);

/* Thread-local description of the last error (TRCERR log lines). */
const char *msbg_last_error(void);

/* dims[0..7] = sx, sy, sz, bsx0, nbx, nby, nbz, nLevels; nBlocks() in *nblocks
 * (msbg.h:3425-3429, sbg.h nbx()/nby()/nbz()/bsx()). */
int msbg_get_dims(const msbg_grid *g, int dims[8], int64_t *nblocks);

/* Select the sweep arithmetic (MSBG_ARITH_*). Default MSBG_ARITH_FAST. */
int msbg_set_arith(msbg_grid *g, int mode);

/* Device synchronisation of the grid's stream. */
int msbg_synchronize(msbg_grid *g);

/* Device memory for the entry points that take DEVICE pointers (msbg_unet_*,
 * msbg_cnn_*), so that a caller needs no other GPU runtime to hold their
 * tensors: msbg_dev_alloc returns `bytes` of device memory on the current
 * device (NULL for 0 bytes), msbg_dev_free releases it, and the copies are
 * synchronous host <-> device transfers. No counterpart in the reference,
 * whose buffers are host memory. */
int msbg_dev_alloc(int64_t bytes, void **out);
int msbg_dev_free(void *p);
int msbg_dev_upload(void *dst, const void *src, int64_t bytes);
int msbg_dev_download(void *dst, const void *src, int64_t bytes);

/* ======================================================================== */
/* particles, activation, refinement map (msbg_demo.cpp:339-441)            */
/* ======================================================================== */

/* Copy n points (xyz float triples, host memory) into the grid's device
 * particle buffer, dropping those outside the domain with the reference's
 * filter: isInDomainRange(pos) (msbg.h:2900-2908, bounds sbg.cpp:262-316)
 * and inRange(truncate_to_int(pos)) (msbg_demo.cpp:356-360).
 * *nkept receives the number of particles kept (nActParticles). */
int msbg_points_upload(msbg_grid *g, const float *xyz, int64_t n,
                       int64_t *nkept);

/* Download the kept particles (in upload order) into xyz[3*nkept]. */
int msbg_points_download(msbg_grid *g, float *xyz, int64_t cap);

/* Block activation (msbg_demo.cpp:368-382): for each kept particle,
 * blockActive[b] += 1 for every block b in
 * [max(trunc(pos/bsx - rScan/bsx),0), min(trunc(pos/bsx + rScan/bsx),nb-1)]^3.
 * counts (optional, host [nBlocks]) receives blockActive[]; *nactive the
 * number of blocks with a nonzero count. */
int msbg_activate_from_points(msbg_grid *g, float rScan, uint32_t *counts,
                              int64_t *nactive);

/* Demo level rule on the GPU (msbg_demo.cpp:418-428), generalised to
 * distance bands: level = 0 for blocks activated with rScan[0], else 1 for
 * blocks activated with rScan[1] (if nBands > 1), ..., else nLevels-1.
 * Runs one activation pass per band. Writes blockLevels (host, optional). */
int msbg_levels_from_points(msbg_grid *g, const float *rScan, int nBands,
                            int *blockLevels);

/* MultiresSparseGrid::regularizeRefinementMap(blockmapIn, blockmapOut)
 * (msbg.h:897-901, msbg.cpp:1202-1423): passes ll = 1..nLevels-2 on the
 * GPU. blockmapOut may equal blockmapIn; with nLevels == 2 it is left
 * untouched, like the reference. Host arrays of nBlocks ints. */
int msbg_regularize_refinement_map(msbg_grid *g, const int *blockmapIn,
                                   int *blockmapOut);

/* MultiresSparseGrid::setRefinementMap(blockLevels, blockFlags, ...)
 * (msbg.h:903-908, msbg.cpp:1428-1745) at levelMg 0: sets BlockInfo.level
 * and flags (BLK_DOM_BORDER, BLK_FINE_COARSE, BLK_COARSE_FINE,
 * BLK_UNIFORM_EMPTY; BLK_EXISTS from blockFlags) on the GPU and builds the
 * per-level fine-coarse block lists (unless OPT_SINGLE_LEVEL).
 * blockLevels == NULL uses the levels left by msbg_levels_from_points.
 * blockFlags may be NULL. */
int msbg_set_refinement_map(msbg_grid *g, const int *blockLevels,
                            const uint32_t *blockFlags);

/* getBlockInfo(bid)->level / ->flags for all blocks (msbg.h:945-950,
 * 673-678). Either pointer may be NULL. */
int msbg_get_block_info(msbg_grid *g, uint16_t *levels, uint16_t *flags);

/* Set or clear BLK_EXISTS for every block (the demo's allocation loop,
 * msbg_demo.cpp:482-497, does FLAG_RESET then FLAG_SET for level-0 blocks).
 * exists == NULL: set it exactly for level-0 blocks. */
int msbg_set_exists(msbg_grid *g, const uint8_t *exists);

/* _blocksFineCoarsePerLevel[level] (msbg.cpp:1720-1727), ascending bid.
 * *n receives the count; out (optional) receives up to cap block ids. */
int msbg_get_fine_coarse_blocks(msbg_grid *g, int level, int32_t *out,
                                int64_t cap, int64_t *n);

/* ======================================================================== */
/* channels (getUint16Channel / getFloatChannel, msbg.h:3048-3077)          */
/* ======================================================================== */

/* Define the block map of channel `chan` at resolution `level`: the
 * equivalent of reset() + prepareDataAccess() + setEmptyValue/FullValue +
 * per block getBlockDataPtr(bid,1) / setEmptyBlock / setFullBlock
 * (sbg.h:1937-1940, 2042-2047, 2165-2168). kind[bid] is MSBG_KIND_*;
 * value blocks get pages in ascending bid order, filled with fill_raw.
 * Raw values are in the storage's own domain (UNORM16: 0..65535).
 * All levels of one channel share one storage format. */
int msbg_channel_define(msbg_grid *g, int chan, int storage, int level,
                        const uint8_t *kind, double empty_raw,
                        double full_raw, double fill_raw);

/* The demo's allocation step (msbg_demo.cpp:451-511) for level 0:
 * empty value 0, full value 1, level-0 blocks become value blocks filled
 * with the full value, all other blocks the const EMPTY block; BLK_EXISTS
 * is set for level-0 blocks and cleared elsewhere. The ascending active
 * block list goes to activeBlocks (optional, cap entries); *nactive gets
 * its length. */
int msbg_channel_alloc_demo(msbg_grid *g, int chan, int storage,
                            int32_t *activeBlocks, int64_t cap,
                            int64_t *nactive);

/* Storage format (MSBG_STORAGE_*) of a defined channel. No counterpart in the
 * reference, where it is the channel's template type Data_T
 * (getChannel<Data_T>, msbg4.cpp:236-240). */
int msbg_channel_storage(msbg_grid *g, int chan, int *storage);

/* Page map of (chan, level): codes MSBG_BLOCK_* or page index per block. */
int msbg_channel_get_pagemap(msbg_grid *g, int chan, int level,
                             int32_t *pagemap, int64_t *npages);

/* Bulk page transfer of (chan, level): npages * bsx_level^3 values of the
 * channel's storage type, page-major, vid = (vz*bsx + vy)*bsx + vx. */
int msbg_channel_download(msbg_grid *g, int chan, int level, void *pages,
                          int64_t bytes);
int msbg_channel_upload(msbg_grid *g, int chan, int level, const void *pages,
                        int64_t bytes);

/* SparseGrid::getBlockDataPtr(bid) read/write of one block (sbg.h:912-977). */
int msbg_download_block(msbg_grid *g, int chan, int level, int bid,
                        void *data);
int msbg_upload_block(msbg_grid *g, int chan, int level, int bid,
                      const void *data);

/* SparseGrid<T>::copyDataTo (sbg.h:550-557) with a storage change: chanDst
 * gets chanSrc's block map at `level` and its values converted (decoded
 * exactly, re-encoded with round-to-nearest-even) into dstStorage. */
int msbg_channel_copy(msbg_grid *g, int chanSrc, int chanDst, int dstStorage,
                      int level);

/* Max |a-b| and RMS(a-b) over all value voxels of two channels with the
 * same block map at `level`, accumulated in FP64 on the device. No
 * counterpart in the reference; used to report storage-format error against
 * an FP64 run. */
int msbg_channel_diff(msbg_grid *g, int chanA, int chanB, int level,
                      double *maxAbs, double *rms);

/* Halo index table of (chan, level): for each value page p (in page order)
 * 27 entries k = (dz+1)*9 + (dy+1)*3 + (dx+1) holding the neighbour's page
 * or a MSBG_BLOCK_* code (OUT outside the domain). out: npages*27 ints. */
int msbg_get_halo_table(msbg_grid *g, int chan, int level, int32_t *out,
                        int64_t cap);

/* ======================================================================== */
/* hot path                                                                 */
/* ======================================================================== */

/* 8-colour min-distance splat of the kept particles into the UNORM16
 * distance field of `chan` (msbg_demo.cpp:541-636): for every voxel of the
 * footprint [ceil(p-rScan-.5), floor(p+rScan-.5)] with d^2 <= rScan^2,
 * value = min(value, roundf(d^2/rScan^2 * 65535)); a particle is abandoned
 * at its first in-sphere voxel (z,y,x order) whose block is not a value
 * block. rScan = rParticle + nbDist. For FP16/BF16 storage the distance
 * code is held in the page bits until msbg_finalize_density; for FP32/FP64
 * in the low 16 bits of each element. */
int msbg_splat_points(msbg_grid *g, int chan, float rParticle, float nbDist);

/* Finalize (msbg_demo.cpp:655-680): q = u/65535; q < 0.999 ->
 * 1 - linstep(-rParticle, rParticle+nbDist, sqrt(q*rScan^2) - rParticle),
 * else 0; stored as roundf(f*65535) for UNORM16, or converted with
 * round-to-nearest-even for the other storages. */
int msbg_finalize_density(msbg_grid *g, int chan, float rParticle,
                          float nbDist);

/* HaloBlockSet::fillHaloBlock_<T>(chan, bid, levelMg=0, tid, opt)
 * (halo.h:382-388, halo.cpp:267-479) for one block at its own level:
 * writes (bsx_L+2*dmax)^3 floats to out (host). opt: MSBG_OPT_BC_NEUMANN,
 * MSBG_OPT_BC_COARSE_LEVEL [| MSBG_OPT_INTERPOLATE], or 0 (empty value). */
int msbg_halo_fill(msbg_grid *g, int chan, int bid, int dmax, unsigned opt,
                   float *out);

/* MultiresSparseGrid::prepareGhostBlocks(chan) (msbg.h:878,
 * msbg.cpp:3059-3170): for each BLK_FINE_COARSE block of levels
 * 0..nLevels-2 with BLK_EXISTS, fill its halo with OPT_BC_COARSE_LEVEL and
 * downsample 4x4x4 ([1,3,3,1]/8 separable, msbg.cpp:298-343, 2902-2974)
 * into the level+1 block slot (allocated as needed); blocks without
 * BLK_EXISTS get the const EMPTY block at level+1. */
int msbg_prepare_ghost_blocks(msbg_grid *g, int chan);

/* MultiresSparseGrid::applyChannelPdeFast<T>(chPhi, chTmp, chTmp2,
 * blockList, laplTyp, doConstrZeroOne, densThr, densThrTol1, densThrTol2,
 * numIter, gridSpacing, timeStep, ..., lambda, ..., taubinStrength, ...,
 * anisDiffAlpha, ...) (msbg.h:2150-2172, msbg4.cpp:60-957).
 * Mode from laplTyp: < 0 with OPT_8_COLOR_SCHEME = 8 colour passes in
 * place; < 0 without it = red-black, 2 passes in place; >= 0 = Jacobi
 * through the temporary channel chTmp (removed afterwards, as the reference
 * resets it). Operators (|laplTyp| without the flags): 4 mean curvature,
 * 1 laplacian flow (Taubin with taubinStrength), 2 bi-laplacian (halo 2),
 * 3 diffusion, 14/15
 * Perona-Malik (anisDiffAlpha != 0; < 0 selects the exponential), 17 Jacobi
 * diffusion; 1, 3 and 17 need 16^3 blocks (they read fixed 16^3-halo
 * offsets in the reference). numIter is rounded up to even; only level-0
 * blocks of blockList are swept (all level-0 value blocks when NULL);
 * Neumann halo. densThr or densThrTol1 > 1e-12 turn on hard constraints
 * around the field at entry, snapshot into chTmp2 (removed afterwards);
 * else doConstrZeroOne clamps to [0,1]. lambda is forced to 0 by the
 * reference (no soft constraints). gridSpacing must be 1. The 8-colour
 * MCF without constraints follows msbg_set_arith (fast or exact); every
 * other combination runs the exact arithmetic. residual (optional, host
 * [sweeps]) receives the FP64 RMS of each sweep's update over the swept
 * voxels. Unsupported values return MSBG_EUNSUPP (laplTyp 10, 11,
 * numIter < 0, timeStep < 0, FP64 channels with other operators). */
int msbg_apply_channel_pde(msbg_grid *g, int chPhi, int chTmp, int chTmp2,
                           const int32_t *blockList, int64_t nList, int laplTyp,
                           int doConstrZeroOne, float densThr,
                           float densThrTol1, float densThrTol2, int numIter,
                           float gridSpacing, float timeStep, float lambda,
                           float taubinStrength, float anisDiffAlpha,
                           double *residual);
/* msbg_apply_channel_pde with lambda, taubinStrength and anisDiffAlpha 0. */
int msbg_apply_channel_pde_fast(msbg_grid *g, int chPhi, int chTmp,
                                int chTmp2, const int32_t *blockList,
                                int64_t nList, int laplTyp,
                                int doConstrZeroOne, float densThr,
                                float densThrTol1, float densThrTol2,
                                int numIter, float gridSpacing,
                                float timeStep, double *residual);

/* Coarse-level sweeps. No counterpart in the reference, which sweeps level
 * 0 only (msbg4.cpp:467-470): the 8-colour MCF sweep of
 * msbg_apply_channel_pde_fast (laplTyp -(4 + OPT_8_COLOR_SCHEME)) applied to
 * the blocks of resolution level `level` on that level's own pages (block
 * edge bsx0 >> level, which must be 8, 16 or 32). blockList as there (NULL:
 * every value block of the level; blocks of other levels are skipped).
 * haloOpt is the halo rule of fillHaloBlock_ (halo.h:267-355):
 * MSBG_OPT_BC_NEUMANN (NULL neighbours and the outside of the domain give
 * the block's own clamped voxel, the level-0 rule), or
 * MSBG_OPT_BC_COARSE_LEVEL [| MSBG_OPT_INTERPOLATE] (NULL neighbours read
 * level+1, which must be defined: its voxel at x/2, or interpolateGhost at
 * (x+.5)/2; halo voxels outside the domain are clipped into it first).
 * Finer neighbours are read from the level's ghost pages
 * (msbg_prepare_ghost_blocks) and are not swept. numIter is rounded up to
 * even; residual (optional, host [sweeps]) gets the FP64 RMS update per
 * sweep. FP64 channels support MSBG_OPT_BC_NEUMANN only. Arithmetic follows
 * msbg_set_arith. */
int msbg_apply_channel_pde_level(msbg_grid *g, int chan, int level,
                                 unsigned haloOpt, const int32_t *blockList,
                                 int64_t nList, int numIter, float timeStep,
                                 int doConstrZeroOne, double *residual);

/* Device-time of the last msbg_apply_channel_pde_fast or
 * msbg_apply_channel_pde_level call (HIP events on the grid's stream around
 * the sweep kernels): total ms and per-launch average ms of the colour-pass
 * kernel. */
int msbg_last_sweep_timing(msbg_grid *g, double *total_ms,
                           double *avg_pass_ms, int64_t *npasses);

/* Read+write bandwidth of a plain streaming copy of `bytes` between two
 * fresh device buffers, averaged over `reps` launches (GB/s, counting the
 * bytes read plus the bytes written). No counterpart in the reference: it
 * is the measured bandwidth reference the sweep is reported against, and a
 * kernel of known traffic for scaling profiler byte counters. */
int msbg_copy_bandwidth(int64_t bytes, int reps, double *gbps);

/* ======================================================================== */
/* field consumers: interpolation and the ray-march renderer                */
/* ======================================================================== */

enum {
  MSBG_IP_LINEAR = 0,   /* cell-centred trilinear (IP_LINEAR) */
  MSBG_IP_BSPLINE3 = 1  /* cubic B-spline with gradient (IP_BSPLINE_CUBIC) */
};

/* Interpolate (chan, level) at n positions xyz[3*i..] in that level's voxel
 * coordinates (cell centres at i + 0.5), options 0:
 *  MSBG_IP_LINEAR: SparseGrid<uint16_t>::interpolateUint16ToFloatFast /
 *    SparseGrid<float>::interpolateFloatFast (sbg.h:4920-5031, 4623-4800),
 *    with interpolateGen0<IP_LINEAR>'s clamped border branch
 *    (sbg.h:3170-3245) where the stencil leaves the domain;
 *  MSBG_IP_BSPLINE3: interpolateWithDerivs0(IP_BSPLINE_CUBIC)
 *    (sbg.h:3903-4135): value and gradient (grad[3*i+k], may be NULL).
 * Values are in the interpolant's units: raw UNORM16 integers (0..65535)
 * for UNORM16 channels, as the reference returns them before
 * RENDER_DENS_TO_FLOAT_UI16; decoded values for the other storages. NULL
 * blocks read as the empty value (the reference's fast paths would
 * dereference NULL). xyz, val, grad: host arrays. The reference leaves FMA
 * contraction to its compiler (-ffast-math); this library fixes it as its
 * kernel source states, so results agree with the reference to rounding. */
int msbg_interpolate(msbg_grid *g, int chan, int level, int kind,
                     const float *xyz, int64_t n, float *val, float *grad);

/* RDR::RaymarchRenderer(width, height, level-0 grid of chan) with
 * setCamera(camPos, lookAt, zoom), setSunLight(sunPos),
 * setSurfaceColor(color), then render() (render.h:58-125,
 * render.cpp:19-100; the demo's render_scene, msbg_demo.cpp:129-185):
 * iso-0.5 ray march in steps of one voxel (1/max(sx,sy,sz) of the unit
 * cube), linear hit refinement, cubic B-spline gradient normal, Lambert
 * shading. rgb[y*width + x] = BMP_MKRGB(r,g,b) = 0x02000000 | b<<16 | g<<8
 * | r (bitmap.h:545-546), row 0 at the top. UNORM16 densities are scaled
 * by 1/65535; the other storages are used as they are. rgb is a host
 * array; *ms (may be NULL) receives the device time of the render kernel.
 * The kernel advances through EMPTY blocks without sampling every step and
 * stops rays that have left the domain for good; the pixels are those of
 * the step-by-step march. */
int msbg_render(msbg_grid *g, int chan, int width, int height,
                const float camPos[3], const float lookAt[3], float zoom,
                const float sunPos[3], const float color[3], uint32_t *rgb,
                double *ms);
/* msbg_render through a counting instance of the same kernel: also
 * returns counts = {march positions advanced, density samples taken,
 * samples that interpolated (value or FULL blocks), rays that hit}. The
 * kernel advances through EMPTY blocks without sampling and stops rays that
 * have left the domain for good; neither changes a pixel. No counterpart in
 * the reference: it gives the per-pixel work behind the render rate. */
int msbg_render_counts(msbg_grid *g, int chan, int width, int height,
                       const float camPos[3], const float lookAt[3], float zoom,
                       const float sunPos[3], const float color[3],
                       uint32_t *rgb, int64_t counts[4]);

/* ======================================================================== */
/* batched sweeps of independent grids                                      */
/* ======================================================================== */

/* A batch runs applyChannelPdeFast's 8-colour MCF sweep (msbg4.cpp:60-957,
 * same per-grid semantics as msbg_apply_channel_pde_fast over every level-0
 * value block) on many independent grids at once. Grids are grouped by
 * kernel (block size x storage x arithmetic). A group runs either as one
 * launch per colour pass covering every grid's blocks of that colour, or as
 * one persistent launch that takes (sweep, colour, grid, block) items from a
 * work queue and waits per grid, not per launch, for the grid's previous
 * colour pass (block sizes 16 and 32). The whole run is captured into one
 * HIP graph. No counterpart in the reference, which sweeps one grid per
 * call. Member grids must outlive the batch and stay idle while it runs; a
 * change to a member's blocks or pages is picked up by the next call. */
typedef struct msbg_batch msbg_batch;
int msbg_batch_create(msbg_batch **out);
int msbg_batch_destroy(msbg_batch *b);
int msbg_batch_add(msbg_batch *b, msbg_grid *g, int chan, float timeStep,
                   int doConstrZeroOne);
/* Time each voxel-column variant of each kernel group on the group's own
 * blocks (`passes` launches of every colour), then the persistent launch
 * with each column shape that fits one block per workgroup (`passes`
 * sweeps), and keep the fastest mapping; the pages are restored after.
 * Member grids adopt the per-pass winner for their own sweeps. */
int msbg_batch_autotune(msbg_batch *b, int passes);
/* Kernel groups of the batch and their variant (0 = 2x2, 1 = 4x2, 2 = 4x4
 * voxel columns). */
int msbg_batch_get_tiles(msbg_batch *b, int *bsx, int *storage, int *tile,
                         int cap, int *n);
/* ... and per group whether the tuned run uses the persistent launch (1) or
 * one launch per colour pass (0). */
int msbg_batch_get_modes(msbg_batch *b, int *persistent, int cap, int *n);
#define MSBG_BATCH_EAGER 0      /* one launch per colour pass and group, no graph */
#define MSBG_BATCH_GRAPH 1      /* one graph launch; each group as tuned (per pass by default) */
#define MSBG_BATCH_PERSISTENT 2 /* one graph launch; every group of block size 16 or 32 as
                                   one persistent launch */
/* numIter rounded up to even sweeps of every member grid, in mode useGraph
 * (MSBG_BATCH_*). Graphs are captured on first use per sweep count and mode.
 * *ms receives the device time of the run. A persistent run whose dependency
 * waits ran past their bound fails with MSBG_EDEVICE (its results are then
 * not to be used). */
int msbg_batch_run(msbg_batch *b, int numIter, int useGraph, double *ms);
/* Wall-clock bound of one dependency wait of a persistent run, in
 * microseconds (default 2 000 000). A test hook: 0 makes every wait give up
 * at once, which exercises the failure report without a long wait. */
int msbg_batch_set_wait_limit(msbg_batch *b, unsigned microseconds);
/* Voxel-column variant used by this grid's own sweeps (-1 = default). */
int msbg_set_tile(msbg_grid *g, int tile);

/* ======================================================================== */
/* synthetic scenes (host)                                                  */
/* ======================================================================== */

/* Seeded sphere/ellipsoid surface-point scene: nObj objects, centres
 * uniform in [cmin,cmax]^3, radii (or semi-axes) uniform in [rmin,rmax],
 * nPts jittered surface points each (Gaussian sigma). ellipsoidFrac of the
 * objects are ellipsoids with a random rotation. Deterministic for a given
 * seed, independent of thread count. xyz: 3*nObj*nPts floats (host). */
int msbg_scene_spheres(uint64_t seed, int nObj, float cmin, float cmax,
                       float rmin, float rmax, int64_t nPts, float sigma,
                       float ellipsoidFrac, float *xyz);
/* The same scene with each surface displaced along the ray from its centre
 * by radius * perlinAmp * noise(perlinFreq * point) (gradient noise, [-1,1];
 * 0 <= perlinAmp < 1). perlinAmp = 0 gives msbg_scene_spheres' output. */
int msbg_scene_objects(uint64_t seed, int nObj, float cmin, float cmax,
                       float rmin, float rmax, int64_t nPts, float sigma,
                       float ellipsoidFrac, float perlinAmp, float perlinFreq,
                       float *xyz);

/* The scene of msbg_scene_objects (same objects; the per-point random
 * numbers come from a counter-based generator, so the points differ from
 * the host generator's) generated on the device into the grid's particle
 * buffer, filtered like msbg_points_upload. append != 0 keeps the points
 * already held. *nkept receives the grid's particle count. */
int msbg_points_generate(msbg_grid *g, uint64_t seed, int nObj, float cmin,
                         float cmax, float rmin, float rmax, int64_t nPts,
                         float sigma, float ellipsoidFrac, float perlinAmp,
                         float perlinFreq, int append, int64_t *nkept);

/* ======================================================================== */
/* learned warm start (no counterpart in the reference)                     */
/* ======================================================================== */
/* A 3-stage 3D U-Net (16/32/64 channels, 3^3 convolutions, BF16 weights and
 * activations, FP32 accumulation on the matrix cores) that maps a 32^3 block
 * with a 2-voxel halo (36^3 window) to the change of the block over K plain
 * MCF sweeps.
 * Tensors passed to msbg_unet_* are DEVICE pointers: windows x are BF16 bits
 * [n][36][36][36], predictions and targets FP32 [n][32][32][32]. Training
 * and backward batches must be 8 * 2^k. Parameters (FP32, host arrays): per
 * convolution w[cout][cin][27] (tap = (dz*3+dy)*3+dx) then b[cout], ten
 * convolutions, then the 1x1 head hw[16], hb. */
typedef struct msbg_unet msbg_unet;
/* He-normal weights from `seed`, zero biases, zero head (predicts 0). */
int msbg_unet_create(uint64_t seed, msbg_unet **out);
int msbg_unet_destroy(msbg_unet *u);
int msbg_unet_num_params(msbg_unet *u, int64_t *n);
int msbg_unet_get_params(msbg_unet *u, float *host, int64_t n);
/* also resets the optimizer state */
int msbg_unet_set_params(msbg_unet *u, const float *host, int64_t n);
/* gradients of the last msbg_unet_backward / msbg_unet_train_step */
int msbg_unet_get_grads(msbg_unet *u, float *host, int64_t n);
int msbg_unet_forward(msbg_unet *u, const uint16_t *x, int n, float *y);
/* forward, then the gradients of sum(dy * y) */
int msbg_unet_backward(msbg_unet *u, const uint16_t *x, const float *dy,
                       int n);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on mean((y - target)^2);
 * *loss receives the loss before the step. With loss == NULL the step is
 * queued without waiting and its loss goes to a running sum that
 * msbg_unet_take_loss returns. */
int msbg_unet_train_step(msbg_unet *u, const uint16_t *x,
                         const float *target, int n, float lr, double *loss);
/* The same step on rows of a resident sample set: the windows X[rows[i]]
 * ([N][36^3] BF16, device) and targets T[rows[i]] ([N][32^3] FP32, device),
 * rows a host array of n sample indices, gathered on the device. */
int msbg_unet_train_step_rows(msbg_unet *u, const uint16_t *X, const float *T,
                              const int32_t *rows, int n, float lr,
                              double *loss);
/* forward of the rows X[rows[i]] into y [n][32^3] (device) */
int msbg_unet_forward_rows(msbg_unet *u, const uint16_t *X,
                           const int32_t *rows, int n, float *y);
/* the summed mean losses of the steps queued with loss == NULL since the
 * last call, and their count; waits for them and resets the sum */
int msbg_unet_take_loss(msbg_unet *u, double *sum, int64_t *steps);
/* 36^3 windows of level-0 value blocks (host block ids) of a 32^3-block
 * channel, under the sweep's halo rule, into x (device). */
int msbg_unet_gather(msbg_grid *g, int chan, const int32_t *blocks, int n,
                     uint16_t *x);
/* the blocks' values as FP32 [n][32^3] into y (device) */
int msbg_unet_extract(msbg_grid *g, int chan, const int32_t *blocks, int n,
                      float *y);
/* add dy (device, FP32) to the blocks' values, clamped to [0,1] */
int msbg_unet_scatter(msbg_grid *g, int chan, const int32_t *blocks, int n,
                      const float *y);
/* Add the network's predicted change to every level-0 value block of `chan`
 * (all windows are read before any block is written), `batch` blocks per
 * forward pass; *ms receives the device time. */
int msbg_warm_start(msbg_grid *g, int chan, msbg_unet *u, int batch,
                    double *ms);
/* One convolution and its two backward products on device tensors
 * (channels-last BF16; w, b, dw, db FP32), for layer tests. Shapes must be
 * ones the network uses. */
int msbg_cnn_conv3d(const uint16_t *x, int n, int cin, int cout, int dout,
                    int pad, const float *w, const float *b, int relu,
                    uint16_t *y);
/* msbg_cnn_conv3d without any allocation: the same kernels, with the
 * swizzled weights in a fixed per-device scratch array of 256 KiB (every
 * layer shape of the network fits; weight sets that do not return
 * MSBG_EINVAL). No counterpart in the reference: it is for callers that
 * cannot allocate device memory, such as code under HIP graph capture or a
 * tight serving loop. All calls on one GPU share that GPU's array: calls
 * that overlap on one GPU race. */
int msbg_cnn_conv3d_noalloc(const uint16_t *x, int n, int cin, int cout,
                            int dout, int pad, const float *w, const float *b,
                            int relu, uint16_t *y);
int msbg_cnn_conv3d_bwd_data(const uint16_t *dy, int n, int cin, int cout,
                             int din, int pad, const float *w, uint16_t *dx);
int msbg_cnn_conv3d_bwd_weights(const uint16_t *x, const uint16_t *dy, int n,
                                int cin, int cout, int dout, int pad,
                                float *dw, float *db);

/* ======================================================================== */
/* adaptive pressure relax (msbg3.cpp)                                      */
/* ======================================================================== */

/* A pressure hierarchy on a block grid of nbx*nby*nbz blocks of bsx0^3
 * cells, each block at its own resolution level (edge bsx0 >> level). The
 * reference keeps the solver's channels inside MultiresSparseGrid
 * (CH_CELL_FLAGS, CH_DIAGONAL, CH_FACE_AREA, msbg.h:291-294) and fills them
 * in code that is not part of its sources; this library computes them from
 * the finest grid's cell types:
 *   coarse cell: fluid if any child is fluid, else void if any is void, else
 *   solid; face area 0 on faces touching a solid cell (outside the domain:
 *   void where the border is open, solid where it is closed), 1 otherwise,
 *   a coarse face the mean of its four children; CELL_PARTIAL_SOLID marks a
 *   fluid cell with a face area below 1; the diagonal is 1 / the sum of the
 *   coefficients the relax applies (0 on non-fluid cells).
 * Multigrid level m holds block b at level max(level_b, m) (msbg.cpp:3374);
 * block levels run down to 2^3 blocks, then denseLevels dense grids of
 * nbx>>k x nby>>k x nbz>>k cells follow.
 *
 * Channel layout of level m: block levels hold the pages of all blocks in
 * block-id order (x fastest), vid = (vz*bsx + vy)*bsx + vx; a dense level is
 * one (sx+2)*(sy+2)*(sz+2) array with a zero ghost shell. */
typedef struct msbg_ps msbg_ps;

enum { MSBG_PS_FLUID = 0, MSBG_PS_SOLID = 1, MSBG_PS_VOID = 2 };
/* domain boundary types (DBC_OPEN, DBC_SOLID, msbg.h:191-192) */
enum { MSBG_DBC_OPEN = 0, MSBG_DBC_SOLID = 1 };
/* channels of the hierarchy: unknown X, right-hand side B, residual R,
 * Jacobi temporary T, diagonal D (read only) */
enum { MSBG_PS_CH_X = 0, MSBG_PS_CH_B = 1, MSBG_PS_CH_R = 2, MSBG_PS_CH_T = 3,
       MSBG_PS_CH_D = 4 };
/* options with the reference's values (msbg.h:107,115,121) */
#define MSBG_OPT_BOUNDZONE_ONLY (1u << 4)
#define MSBG_OPT_CALC_RESIDUAL (1u << 12)
#define MSBG_OPT_REVERSE_ORDER (1u << 18)
/* cell and block flags (msbg.h:565-615) */
#define MSBG_CELL_SOLID (1u << 0)
#define MSBG_CELL_PARTIAL_SOLID (1u << 6)
#define MSBG_CELL_VOID (1u << 12)
#define MSBG_BLK_NO_FLUID (1u << 5)
#define MSBG_BLK_CUTS_SOLID (1u << 15)

typedef struct {
  int nbx, nby, nbz; /* block grid */
  int bsx0;          /* finest block edge: 4, 8 or 16 */
  double dx0;        /* finest cell size (_dx0); level l has h = 2^l dx0 */
  int domainBc[3][2];/* MSBG_DBC_* per axis, lower/upper (_domainBc) */
  float omega;       /* relaxation factor (_mgSmOmega, 6/7, msbg.cpp:819) */
  int blockIter;     /* sub-iterations per block (_mgSmBlockIter, 5, msbg.cpp:837) */
  int denseLevels;   /* dense multigrid levels below the block levels */
} msbg_ps_desc;

/* Builds the hierarchy. blockLevels[nb] are the blocks' resolution levels
 * (face neighbours may differ by one level at most, msbg3.cpp:690);
 * cellType[(nbx*bsx0)*(nby*bsx0)*(nbz*bsx0)] holds MSBG_PS_* per finest
 * cell, x fastest. All channels start at zero. */
int msbg_ps_create(const msbg_ps_desc *desc, const int8_t *blockLevels,
                   const uint8_t *cellType, msbg_ps **out);
int msbg_ps_destroy(msbg_ps *ps);
int msbg_ps_get_levels(const msbg_ps *ps, int *nLevelsMg, int *nLevelsBlock);
/* floats in one channel of multigrid level levelMg */
int msbg_ps_level_size(const msbg_ps *ps, int levelMg, int64_t *n);
int msbg_ps_upload(msbg_ps *ps, int chan, int levelMg, const float *host);
int msbg_ps_download(msbg_ps *ps, int chan, int levelMg, float *host);
/* cell flags (one per channel float) and, on block levels, block flags */
int msbg_ps_download_flags(msbg_ps *ps, int levelMg, uint16_t *cellFlags,
                           uint16_t *blockFlags);
int msbg_ps_fill(msbg_ps *ps, int chan, int levelMg, float v);

/* MultiresSparseGrid::relax(boundaryZoneOnly, reverseOrder, levelMg, chanX,
 * chanB, chanTmp, nIter) (msbg3.cpp:2244-2369). Block levels: nIter passes
 * of red-black block Gauss-Seidel (relaxBlockList, msbg3.cpp:2170-2239),
 * every block relaxed by IprocessBlockLaplacian's blockIter sub-iterations
 * (msbg3.cpp:418-1600). Dense levels: nIter damped-Jacobi sweeps through
 * chanTmp (ImultiplyLaplacianMatrixDenseLevel, msbg3.cpp:1917-2165).
 * boundaryZoneOnly != 0 returns MSBG_EUNSUPP. */
int msbg_ps_relax(msbg_ps *ps, int boundaryZoneOnly, int reverseOrder,
                  int levelMg, int chanX, int chanB, int chanTmp, int nIter);
/* MultiresSparseGrid::multiplyLaplacianMatrixOpt(options, levelMg, chanX,
 * chanB, kap, chanDstY, pDotprod) (msbg3.cpp:1756-1912): Y = A X, or with
 * MSBG_OPT_CALC_RESIDUAL Y = B - A X; without it *dotprod (if given)
 * receives sum(Y * X) on a block level. kap is unused, as in the
 * reference's body. Vectors of cells without fluid are not written. */
int msbg_ps_multiply(msbg_ps *ps, unsigned options, int levelMg, int chanX,
                     int chanB, float kap, int chanDstY, double *dotprod);
/* Constant restriction from levelMg to levelMg+1 (a coarse cell gets the
 * sum of its eight children) and constant prolongation from levelMg+1
 * added into levelMg (MG_CONST_RESTRICTION, MG_CONST_PROLONGATION with
 * alpha 1, msbgcell.h:99-101; the reference's operators are not part of
 * its sources, these are this library's). */
int msbg_ps_restrict(msbg_ps *ps, int levelMg, int chanSrc, int chanDst);
int msbg_ps_prolongate(msbg_ps *ps, int levelMg, int chanSrc, int chanDst);
/* One V-cycle on X/B of level 0 (the reference's multigrid driver is not
 * part of its sources): nu1 relax passes, residual, restriction, recursion
 * with zero start, prolongation, nu2 passes in reverse order; the coarsest
 * level gets coarseIter passes. */
int msbg_ps_vcycle(msbg_ps *ps, int nu1, int nu2, int coarseIter);
int msbg_ps_synchronize(msbg_ps *ps);
/* Device time of the last msbg_ps_relax / _multiply / _vcycle call (HIP
 * events on the hierarchy's stream around all of its launches); waits for
 * it to finish. The reference's relax() reports its time the same way per
 * call (msbg3.cpp:2349-2358). */
int msbg_ps_last_timing(msbg_ps *ps, double *ms);
/* Cells one relax iteration updates at levelMg: the cells of every block
 * with fluid (relax()'s nTotalCells, msbg3.cpp:2228), or all cells of a
 * dense level. */
int msbg_ps_relaxed_cells(const msbg_ps *ps, int levelMg, int64_t *cells);

#ifdef __cplusplus
}
#endif
#endif
