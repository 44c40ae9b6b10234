// Internal types of libmsbg_gpu: grid, channels, device buffers, launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/msbg_gpu.h"

namespace msbg {

constexpr int kMaxLevels = 4;  // MSBG_MAXRESLEVELS (msbgcell.h:147)

void set_error(const char *fmt, ...);

#define MSBG_HIP(call)                                                      \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      ::msbg::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,          \
                        hipGetErrorString(e_));                             \
      return MSBG_EDEVICE;                                                  \
    }                                                                       \
  } while (0)

#define MSBG_CHECK(cond, code, ...)        \
  do {                                     \
    if (!(cond)) {                         \
      ::msbg::set_error(__VA_ARGS__);      \
      return (code);                       \
    }                                      \
  } while (0)

#define MSBG_TRY(expr)      \
  do {                      \
    int rc_ = (expr);       \
    if (rc_) return rc_;    \
  } while (0)

int storage_bytes(int storage);

// float bounds of the domain of isInDomainRange (sbg.cpp:262-316)
void domain_bounds(int sx, int sy, int sz, float dmin[3], float dmax[3]);

// Run body(i) for i in [0, n) on up to hardware_concurrency() host threads,
// in contiguous chunks.
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &body);

// Device buffer with explicit ownership.
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;  // owns its allocation: moved, never copied
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      free();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  int alloc(size_t count) {
    free();
    if (count == 0) return 0;
    if (hipMalloc((void **)&p, count * sizeof(T)) != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc of %zu bytes failed", count * sizeof(T));
      return MSBG_ENOMEM;
    }
    n = count;
    return 0;
  }
  int ensure(size_t count) { return count <= n ? 0 : alloc(count); }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { free(); }
};

// Everything the sweep kernel needs to know about one grid's channel at
// level 0 (device-side descriptor; one per grid of a batched launch).
struct SweepGrid {
  uint8_t *pages;  // pages and side-cars the halo and the old values come from
  uint8_t *side;
  const uint8_t *consts;
  const int32_t *nbr;
  int64_t npages;
  float dt, lo, hi;
  // outside-the-domain halo cells: 0 = the block's own clamped voxel
  // (OPT_BC_NEUMANN), 1 = clipped to the domain first (OPT_BC_COARSE_LEVEL)
  int outClip;
  // where the new values go: the same pages in place, another channel with
  // the same page numbering in the Jacobi mode
  uint8_t *dst_pages;
  uint8_t *dst_side;
  const uint8_t *phi0;  // hard constraints: the field at entry, same numbering
  float tol1, tol2;     // ... clamp to [max(0, phi0-tol1), min(1, phi0+tol2)]
  float taubin;         // laplacian flow: even sweeps subtract taubin * D
  float alpha;          // Perona-Malik edge scale
};

// One channel level as the interpolators and the renderer read it.
struct RenderVol {
  const int32_t *page;  // block map codes
  const void *pages;    // page pool
  int sx, sy, sz;       // voxel extents of the level
  int lb;               // log2 of the level's block edge
  int nbx, nby;
  float empty, full;    // interpolant inputs of the const blocks
  float dmin[3], dmax[3];  // float domain bounds (sbg.cpp:262-316)
};

// RaymarchRenderer's camera and material, widened to double (render.h:66-77)
struct RenderCam {
  double pos[3], look[3];
  double focal;
  double sun[3], color[3];
};

// One resolution level of one channel.
struct ChanLevel {
  bool defined = false;
  int bsx = 0;                       // block edge at this level
  std::vector<int32_t> h_page;       // block map codes (host mirror)
  DevBuf<int32_t> d_page;            // block map codes (device)
  DevBuf<uint8_t> d_pages;           // page pool
  int64_t npages = 0;
  std::vector<int32_t> h_page_bid;   // page -> block id
  DevBuf<int32_t> d_page_bid;
  DevBuf<int32_t> d_nbr;             // halo index table [npages][27]
  bool nbr_valid = false;
  // side-car copies of each page's two x-faces [page][2][z][y] followed by
  // its four z-edges [page][4][z]: the sweep reads neighbour x-faces from
  // here contiguously instead of with a bsx-strided gather from the page
  DevBuf<uint8_t> d_side;
  DevBuf<uint8_t> d_consts;          // 4 x empty value, 4 x full value (raw)
  bool side_valid = false;
  // renewed by every write to the pages that leaves the side-cars stale (the
  // sweeps that keep them current do not renew it): a coarse-level sweep
  // compares it with the value after its own last call
  uint64_t wstamp = 0;
  double empty_raw = 0, full_raw = 0;
};

// the pages of `cl` were written by something that did not update its
// side-cars; stamps come from one process-wide counter, so a level that is
// defined afresh (wstamp 0) never repeats a stamp another set-up recorded
inline void pages_written(ChanLevel &cl) {
  static std::atomic<uint64_t> clock{0};
  cl.side_valid = false;
  cl.wstamp = ++clock;
}

struct Channel {
  int id = 0;
  int storage = 0;
  ChanLevel lv[kMaxLevels];
};

// Cached set-up of a coarse-level sweep of one (channel, level, halo rule):
// the working pool (the level's pages followed by ghost pages), its map and
// halo table, the ghost list and the colour-sorted items. Valid while the
// grid's topology version and the block list are those of `key`.
struct LevelWork {
  uint64_t key = 0;
  // what the key was computed from: the topology version and the block list
  // (has_list false: every value block of the level), compared in full
  // before the set-up is reused, so that a hash collision cannot hand one
  // list's set-up to another
  uint64_t topo = 0;
  bool has_list = false;
  std::vector<int32_t> list;
  ChanLevel wl;
  DevBuf<int32_t> d_ghost;
  DevBuf<uint32_t> d_gmask;  // per ghost page: bit (dz+1)*9+(dy+1)*3+dx+1 = swept block at +d
  int64_t nghost = 0;
  DevBuf<int2> d_items;
  int64_t off[9] = {0};
  DevBuf<SweepGrid> d_desc;
  // the level's wstamp when this set-up's last call ended: while it still
  // holds (levels >= 1, whose every other writer renews it), wl's side-cars
  // of the level's own pages are current and only the ghost pages' are
  // rebuilt
  uint64_t side_stamp = ~0ull;
};

}  // namespace msbg

struct msbg_grid {
  std::string name;
  int sx = 0, sy = 0, sz = 0, bsx0 = 0, lg0 = 0;
  int nbx = 0, nby = 0, nbz = 0;
  int64_t nblocks = 0;
  int nlevels = 0;
  unsigned options = 0;
  int arith = MSBG_ARITH_FAST;
  hipStream_t stream = nullptr;

  // block info (levelMg 0)
  std::vector<uint16_t> h_level, h_flags;
  msbg::DevBuf<uint16_t> d_level, d_flags;
  bool have_levels = false;
  std::vector<int32_t> fc_list[msbg::kMaxLevels];  // fine-coarse blocks per level

  // particles
  msbg::DevBuf<float> d_xyz;
  int64_t npoints = 0;
  // the particles counting-sorted into 16^3 cells (splat, activation);
  // rebuilt after the point set changes
  msbg::DevBuf<float> bin_xyz;
  msbg::DevBuf<uint32_t> bin_off, bin_cnt;
  bool bin_valid = false;
  msbg::DevBuf<uint32_t> d_count;
  msbg::DevBuf<int32_t> d_tmp_levels;

  std::map<int, msbg::Channel> channels;

  // sweep scratch
  msbg::DevBuf<int2> d_sweep_list;      // colour-sorted {grid 0, page} items
  msbg::DevBuf<msbg::SweepGrid> d_desc; // this grid's sweep descriptor
  int tile[8] = {-1, -1, -1, -1, -1, -1, -1, -1};  // voxel-column variant per bsx log2 (-1 = default)
  std::map<uint64_t, msbg::LevelWork> level_work;  // coarse-level sweep set-ups
  msbg::DevBuf<double> d_partial;       // per-block residual partials
  msbg::DevBuf<double> d_resid;         // per-sweep residuals
  uint64_t topo_version = 1;             // bumped when block maps or levels change
  // sweeps run on this grid by msbg_apply_channel_pde*: the step the UNORM8
  // store's stochastic rounding draws from (always even between calls)
  uint64_t sweep_step = 0;
  uint64_t sweep_key = 0;                // key of the cached colour lists
  int64_t sweep_off[9] = {0};
  hipEvent_t ev[2] = {nullptr, nullptr};
  double last_total_ms = 0, last_avg_pass_ms = 0;
  int64_t last_npasses = 0;
};

namespace msbg {

// ---- launchers implemented in the .hip files ----
// activation: counts[bid] += 1 over each particle's trunc box
int launch_activate(msbg_grid *g, float rScan, uint32_t *d_count);
// builds g->bin_* from the current particles unless they are up to date;
// returns MSBG_EUNSUPP when the point set is too large to bin
int ensure_binned(msbg_grid *g);
// the reference's domain filter over d_raw[n][3] (device), kept points into
// g->d_xyz in their order
int filter_points(msbg_grid *g, const float *d_raw, int64_t n, int64_t *nkept);
// level bands: level = first band whose count is nonzero, else nLevels-1
int launch_level_band(msbg_grid *g, const uint32_t *d_count, int band,
                      int nbands, int32_t *d_levels);
int launch_regularize(msbg_grid *g, const int32_t *d_in, int32_t *d_out,
                      int ll);
int launch_flags(msbg_grid *g, const int32_t *d_levels,
                 const uint32_t *d_in_flags);
int launch_build_nbr(msbg_grid *g, ChanLevel &cl, int level);
int launch_fill_pages(msbg_grid *g, ChanLevel &cl, int storage,
                      double fill_raw, int64_t first_page, int64_t npages);
int launch_splat(msbg_grid *g, Channel &ch, float rScan);
int launch_finalize(msbg_grid *g, Channel &ch, float rParticle, float nbDist);
int launch_halo_fill(msbg_grid *g, Channel &ch, int level, int bid, int dmax,
                     unsigned opt, float *d_out);
int launch_ghost(msbg_grid *g, Channel &ch, int level, const int32_t *d_bids,
                 int64_t n);
// sweep kernel of one colour pass over work items {grid, page}
int sweep_prepare(msbg_grid *g, Channel &ch);  // side-car faces/edges + consts
// the same for any channel level (its own block size) or a working copy;
// when the side-cars are stale, pages [fromPage, npages) are refreshed
int sweep_prepare_level(msbg_grid *g, ChanLevel &cl, int storage, int64_t fromPage = 0);
// ghost pages of a coarse-level sweep (level `level`, values from level+1);
// with d_side (side-cars of a pool of nTotal pages) the INTERPOLATE fill also
// writes the side-car entries of what it fills, and *sideDone says it did
int launch_fill_ghost_pages(msbg_grid *g, Channel &ch, int level, unsigned opt,
                            const int32_t *d_bids, const uint32_t *d_mask, int64_t n,
                            uint8_t *d_pages, int64_t firstPage, uint8_t *d_side,
                            int64_t nTotal, bool *sideDone);
SweepGrid sweep_desc(Channel &ch, float dt, int doConstr01);
// c01: every grid of d_grids clamps to [0,1] (doConstrZeroOne), which
// selects the fast kernel whose last fma clamps; the other kernels clamp to
// each grid's own [lo,hi]
// iter: the sweep step (per grid, counted across calls) that the UNORM8
// store's stochastic rounding draws from
int launch_sweep_items(hipStream_t s, int bsx, int storage, bool exact, bool c01, int tile,
                       const SweepGrid *d_grids, const int2 *d_items, int64_t n,
                       double *d_partial, int iter);
// The persistent batched sweep of one group of a batch: one launch runs
// every sweep and colour of the group's grids from a work queue, waiting per
// grid for its previous colour pass (k_sweep_queue).
constexpr int kDoneStride = 16;  // queue counters one 64-byte line apart
constexpr unsigned kQueueRun = 8;  // queue items a workgroup takes at a time
struct BatchQueue {
  const int2 *items;   // the group's items (entry, page), sorted by colour
  const int *pre;      // per grid entry, 9 values: its items of colours < c; [8] = all
  unsigned *done;      // per grid entry (stride kDoneStride): its items finished in this run
  unsigned *head;      // next queue index of this group
  unsigned *fail;      // set when a dependency wait ran past waitTicks
  int off[9];          // colour offsets into items
  int sweeps;
  uint64_t waitTicks;  // wall-clock ticks a dependency wait may last
};
// one wave and one block per workgroup (4x4 voxel columns at 32^3, 2x2 at
// 16^3; not FP64); nwg = 0: as many workgroups as fit on the device at
// once, but no more than the largest colour pass holds blocks (workgroups
// beyond it would only wait)
int launch_sweep_queue(hipStream_t s, int bsx, int storage, bool exact, bool c01,
                       const SweepGrid *d_grids, const BatchQueue &q, int nwg = 0);
int launch_queue_reset(hipStream_t s, unsigned *d_state, int n);
// wall-clock ticks (s_memrealtime) in `us` microseconds on the current device
uint64_t wall_ticks(uint64_t us);
// the other applyChannelPdeFast operators (1, 3, 14, 15, 17) and MCF with
// hard constraints, exact arithmetic, on the same item lists
int launch_pde_items(hipStream_t s, int bsx, int storage, int op, bool hard,
                     const SweepGrid *d_grids, const int2 *d_items, int64_t n, double *d_partial,
                     int iter);
// batched interpolation of a channel level (kind MSBG_IP_*) and the renderer
int launch_interp_points(hipStream_t s, const RenderVol &V, int storage, int kind,
                         const float *d_pts, int64_t n, float *d_val, float *d_grad);
int launch_render(hipStream_t s, const RenderVol &V, int storage, const RenderCam &C, int W,
                  int H, uint32_t *d_rgb, unsigned long long *d_counts = nullptr);
// bi-laplacian (halo 2) of the listed blocks into `scratch` (n pages), then
// committed to dst_pages and their side-cars
int launch_bilap_items(hipStream_t s, int bsx, int storage, bool hard, const SweepGrid *d_grid,
                       const int2 *d_items, int64_t n, double *d_partial, uint8_t *scratch);
// copy the listed pages and side-cars from pages to dst_pages of *d_grid
int launch_copy_items(hipStream_t s, int bsx, int storage, const SweepGrid *d_grid,
                      const int2 *d_items, int64_t n);
int default_tile(int bsx, int storage, int64_t nItems);
constexpr int kNumTiles = 3;  // voxel-column variants: 0 = 2x2, 1 = 4x2, 2 = 4x4
int ensure_nbr(msbg_grid *g, ChanLevel &cl, int level);
int launch_reduce_residual(msbg_grid *g, const double *d_partial, int64_t n,
                           double *d_out);
int launch_convert(msbg_grid *g, const ChanLevel &src, int sst, ChanLevel &dst,
                   int dst_st);
int launch_diff(msbg_grid *g, const ChanLevel &a, int ast, const ChanLevel &b, int bst,
                double *maxAbs, double *sumSq);

}  // namespace msbg
