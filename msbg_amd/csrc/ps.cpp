// Adaptive pressure relax: hierarchy set-up and host drivers (msbg_ps_*).
//
// The reference's relax path (msbg3.cpp) consumes cell flags, a diagonal and
// three face-area channels on every multigrid level, and its multigrid driver
// restricts and prolongates between levels. The code that computes those
// coefficients and the driver are not part of the reference sources; this
// file defines them (rules below) so that the relax has a complete hierarchy
// to run on:
//   - cell types per level: a coarse cell is fluid if any child is fluid,
//     else void if any child is void, else solid;
//   - face areas: 0 on a face touching a solid cell at the finest level
//     (outside the domain: void where the border is open, solid where it is
//     closed), 1 otherwise; a coarse face is the mean of its four child faces;
//   - cell flags: CELL_SOLID, CELL_VOID, CELL_PARTIAL_SOLID for a fluid cell
//     with a face area below 1; block flags BLK_NO_FLUID, BLK_CUTS_SOLID;
//   - diagonal D = 1 / (sum of the coefficients the relax applies to the six
//     neighbours), so that every row of the operator sums to zero apart from
//     Dirichlet neighbours; 0 for non-fluid cells.
// Multigrid level m holds block b at level max(level_b, m) (msbg.cpp:3374);
// below the last block level (2^3 blocks) come the requested dense levels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "internal.h"
#include "ps.h"

using namespace msbg;

namespace {

constexpr int kMaxMg = 14;

struct HostLevel {  // global dense data of one resolution level
  int sx = 0, sy = 0, sz = 0;
  std::vector<uint8_t> type;
  std::vector<float> fa[3];  // u: (sx+1)*sy*sz, v: sx*(sy+1)*sz, w: sx*sy*(sz+1)
  std::vector<uint16_t> flags;
};

struct MgLevel {
  bool dense = false;
  int sx = 0, sy = 0, sz = 0;
  int64_t n = 0;
  std::vector<int8_t> hL;
  std::vector<int64_t> hoff;
  std::vector<uint16_t> hflags, hbflags;
  // block lists per (coordinate-sum parity, level, BLK_CUTS_SOLID)
  std::vector<int32_t> hlist;
  std::vector<PsItem> hitems;  // the listed blocks' records, same order
  int listOff[2][kPsMaxBlockLevels][2] = {}, listLen[2][kPsMaxBlockLevels][2] = {};
  DevBuf<int8_t> L;
  DevBuf<int64_t> off;
  DevBuf<float> ch[4], D, fa[3];
  DevBuf<uint16_t> flags, bflags;
  DevBuf<PsItem> list;
};

}  // namespace

struct msbg_ps {
  msbg_ps_desc desc{};
  int nb = 0, nlev = 0, nmg = 0;
  std::vector<int8_t> level;
  HostLevel g[kMaxMg];
  std::vector<float> pfa[kPsMaxBlockLevels][3];
  std::vector<int32_t> hnbr;
  MgLevel mg[kMaxMg];
  DevBuf<float> dpfa[kPsMaxBlockLevels][3];
  DevBuf<double> dot;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};  // start / end of the last timed call
  bool timed = false;
  ~msbg_ps() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int vox(int bsx, int x, int y, int z) { return x + bsx * (y + bsx * z); }
int simdw(int bsx) { return bsx >= 8 ? 8 : bsx; }

int64_t fa_index(const HostLevel &l, int d, int x, int y, int z) {
  const int ex = l.sx + (d == 0), ey = l.sy + (d == 1);
  return x + (int64_t)ex * (y + (int64_t)ey * z);
}

int type_at(const msbg_ps &ps, const HostLevel &l, int x, int y, int z) {
  int d = -1, s = 0;
  if (x < 0) d = 0, s = 0;
  else if (x >= l.sx) d = 0, s = 1;
  else if (y < 0) d = 1, s = 0;
  else if (y >= l.sy) d = 1, s = 1;
  else if (z < 0) d = 2, s = 0;
  else if (z >= l.sz) d = 2, s = 1;
  if (d >= 0) return ps.desc.domainBc[d][s] == MSBG_DBC_OPEN ? MSBG_PS_VOID : MSBG_PS_SOLID;
  return l.type[x + (int64_t)l.sx * (y + (int64_t)l.sy * z)];
}

void build_levels(msbg_ps &ps) {
  for (int l = 1; l < ps.nmg; l++) {
    HostLevel &c = ps.g[l], &f = ps.g[l - 1];
    c.type.resize((size_t)c.sx * c.sy * c.sz);
    for (int z = 0; z < c.sz; z++)
      for (int y = 0; y < c.sy; y++)
        for (int x = 0; x < c.sx; x++) {
          bool anyFluid = false, anyVoid = false;
          for (int k = 0; k < 8; k++) {
            const int t = f.type[(2 * x + (k & 1)) +
                                 (int64_t)f.sx * ((2 * y + ((k >> 1) & 1)) + (int64_t)f.sy * (2 * z + (k >> 2)))];
            anyFluid |= t == MSBG_PS_FLUID;
            anyVoid |= t == MSBG_PS_VOID;
          }
          c.type[x + (int64_t)c.sx * (y + (int64_t)c.sy * z)] =
              anyFluid ? MSBG_PS_FLUID : anyVoid ? MSBG_PS_VOID : MSBG_PS_SOLID;
        }
  }
  for (int l = 0; l < ps.nmg; l++) {
    HostLevel &L = ps.g[l];
    for (int d = 0; d < 3; d++) {
      const int ex = L.sx + (d == 0), ey = L.sy + (d == 1), ez = L.sz + (d == 2);
      std::vector<float> &fa = L.fa[d];
      fa.resize((size_t)ex * ey * ez);
      for (int z = 0; z < ez; z++)
        for (int y = 0; y < ey; y++)
          for (int x = 0; x < ex; x++) {
            float v;
            if (l == 0) {
              const int lo = type_at(ps, L, x - (d == 0), y - (d == 1), z - (d == 2));
              const int hi = type_at(ps, L, x, y, z);
              v = (lo == MSBG_PS_SOLID || hi == MSBG_PS_SOLID) ? 0.0f : 1.0f;
            } else {
              const HostLevel &F = ps.g[l - 1];
              float s = 0.0f;
              for (int k = 0; k < 4; k++) {
                const int p = k & 1, q = k >> 1;
                int cx = 2 * x, cy = 2 * y, cz = 2 * z;
                if (d == 0) cy += p, cz += q;
                else if (d == 1) cx += p, cz += q;
                else cx += p, cy += q;
                s = s + F.fa[d][fa_index(F, d, cx, cy, cz)];
              }
              v = s * 0.25f;
            }
            fa[x + (int64_t)ex * (y + (int64_t)ey * z)] = v;
          }
    }
    L.flags.assign((size_t)L.sx * L.sy * L.sz, 0);
    for (int z = 0; z < L.sz; z++)
      for (int y = 0; y < L.sy; y++)
        for (int x = 0; x < L.sx; x++) {
          const int64_t i = x + (int64_t)L.sx * (y + (int64_t)L.sy * z);
          const int t = L.type[i];
          uint16_t f = 0;
          if (t == MSBG_PS_SOLID) {
            f = kPsCellSolid;
          } else if (t == MSBG_PS_VOID) {
            f = kPsCellVoid;
          } else {
            float m = 1.0f;
            for (int d = 0; d < 3; d++)
              for (int s = 0; s < 2; s++)
                m = std::min(m, L.fa[d][fa_index(L, d, x + s * (d == 0), y + s * (d == 1), z + s * (d == 2))]);
            if (m < 1.0f) f = kPsCellPartialSolid;
          }
          L.flags[i] = f;
        }
  }
}

void block_xyz(const msbg_ps &ps, int b, int c[3]) {
  c[0] = b % ps.desc.nbx;
  c[1] = (b / ps.desc.nbx) % ps.desc.nby;
  c[2] = b / (ps.desc.nbx * ps.desc.nby);
}

bool fluid(uint16_t f) { return (f & (kPsCellSolid | kPsCellVoid)) == 0; }

// the coefficient sum of a mixed-resolution face (procNeighMixedRes2's a's)
double mixed_coeff(const msbg_ps &ps, const MgLevel &g, int b, int b2, int d, int s, int vx, int vy,
                   int vz) {
  const int L = g.hL[b], L2 = g.hL[b2], bsx = ps.desc.bsx0 >> L;
  const bool cuts = g.hbflags[b] & kPsBlkCutsSolid;
  double cs = 0;
  if (L2 < L) {
    const int bh = 2 * bsx;
    const float *fa = ps.pfa[L2][d].data() + (int64_t)(s ? b2 : b) * bh * bh * bh;
    for (int k = 0; k < 4; k++) {
      const int p = k & 1, q = k >> 1;
      int fx, fy, fz;
      if (d == 0) fx = 0, fy = 2 * vy + p, fz = 2 * vz + q;
      else if (d == 1) fx = 2 * vx + p, fy = 0, fz = 2 * vz + q;
      else fx = 2 * vx + p, fy = 2 * vy + q, fz = 0;
      float a = (float)(1.0 * (1. / (4 * (3. / 4.))));
      if (cuts) a = a * fa[vox(bh, fx, fy, fz)];
      cs += (double)a;
    }
  } else {
    int fx = vx, fy = vy, fz = vz;
    if (s) (d == 0 ? fx : d == 1 ? fy : fz) = 0;
    const float *fa = ps.pfa[L][d].data() + (int64_t)(s ? b2 : b) * bsx * bsx * bsx;
    float a = (float)(1.0 * 2 * (1. / (4 * (3. / 4.))));
    if (cuts) a = a * fa[vox(bsx, fx, fy, fz)];
    cs += (double)a;
  }
  return cs;
}

// weight the relax applies to face (d,s) of cell c of block b, and its kind
// (0 same level, 1 outside, 2 mixed resolution)
float face_weight(const msbg_ps &ps, const MgLevel &g, int b, int d, int s, const int c0[3], int *kind) {
  const int L = g.hL[b], bsx = ps.desc.bsx0 >> L;
  int c[3] = {c0[0], c0[1], c0[2]};
  const float *own = ps.pfa[L][d].data() + (int64_t)b * bsx * bsx * bsx;
  *kind = 0;
  if (!s && c[d] > 0) return own[vox(bsx, c[0], c[1], c[2])];
  if (s && c[d] < bsx - 1) {
    c[d]++;
    return own[vox(bsx, c[0], c[1], c[2])];
  }
  const int b2 = ps.hnbr[6 * b + 2 * d + s];
  if (b2 < 0) {
    *kind = 1;
    if (!s) return own[vox(bsx, c[0], c[1], c[2])];
    return ps.desc.domainBc[d][1] == MSBG_DBC_OPEN ? 1.0f : 0.0f;
  }
  if (g.hL[b2] != L) {
    *kind = 2;
    return 1.0f;
  }
  if (!s) return own[vox(bsx, c[0], c[1], c[2])];
  c[d] = 0;
  return ps.pfa[L][d][(int64_t)b2 * bsx * bsx * bsx + vox(bsx, c[0], c[1], c[2])];
}

int build_block_level(msbg_ps &ps, int m) {
  MgLevel &g = ps.mg[m];
  const int nb = ps.nb, bsx0 = ps.desc.bsx0;
  g.hL.resize(nb);
  g.hoff.resize(nb);
  int64_t n = 0;
  for (int b = 0; b < nb; b++) {
    const int L = std::max<int>(ps.level[b], m), bsx = bsx0 >> L;
    g.hL[b] = (int8_t)L;
    g.hoff[b] = n;
    n += (int64_t)bsx * bsx * bsx;
  }
  g.n = n;
  g.hflags.assign(n, 0);
  g.hbflags.assign(nb, 0);
  std::vector<float> D(n, 0.0f);
  for (int b = 0; b < nb; b++) {
    const int L = g.hL[b], bsx = bsx0 >> L;
    int bc[3];
    block_xyz(ps, b, bc);
    const HostLevel &G = ps.g[L];
    bool anyFluid = false, cuts = false;
    for (int z = 0; z < bsx; z++)
      for (int y = 0; y < bsx; y++)
        for (int x = 0; x < bsx; x++) {
          const uint16_t f =
              G.flags[(bc[0] * bsx + x) + (int64_t)G.sx * ((bc[1] * bsx + y) + (int64_t)G.sy * (bc[2] * bsx + z))];
          g.hflags[g.hoff[b] + vox(bsx, x, y, z)] = f;
          anyFluid |= fluid(f);
          cuts |= (f & (kPsCellSolid | kPsCellPartialSolid)) != 0;
        }
    g.hbflags[b] = (uint16_t)((anyFluid ? 0 : kPsBlkNoFluid) | (cuts ? kPsBlkCutsSolid : 0));
  }
  for (int b = 0; b < nb; b++) {
    const int L = g.hL[b], bsx = bsx0 >> L, sw = simdw(bsx);
    const uint16_t *fl = g.hflags.data() + g.hoff[b];
    const bool cuts = g.hbflags[b] & kPsBlkCutsSolid;
    for (int z = 0; z < bsx; z++)
      for (int y = 0; y < bsx; y++)
        for (int x = 0; x < bsx; x++) {
          const int v = vox(bsx, x, y, z);
          if (!fluid(fl[v])) continue;
          bool gated = false;
          for (int k = (v / sw) * sw; k < (v / sw + 1) * sw; k++)
            gated |= (fl[k] & (kPsCellSolid | kPsCellPartialSolid)) != 0;
          gated = gated && cuts;
          const int c[3] = {x, y, z};
          double sum = 0;
          for (int d = 0; d < 3; d++)
            for (int s = 0; s < 2; s++) {
              int kind;
              const float w = face_weight(ps, g, b, d, s, c, &kind);
              const bool border = s ? c[d] == bsx - 1 : c[d] == 0;
              if (border && kind == 2)
                sum += mixed_coeff(ps, g, b, ps.hnbr[6 * b + 2 * d + s], d, s, x, y, z);
              else
                sum += gated ? (double)w : 1.0;
            }
          D[g.hoff[b] + v] = sum > 0 ? (float)(1.0 / sum) : 0.0f;
        }
  }
  // block lists per (coordinate-sum parity, level, cuts a solid) of the
  // blocks with fluid; the relax takes one parity per pass, the multiply both
  g.hlist.clear();
  for (int par = 0; par < 2; par++)
    for (int L = 0; L < kPsMaxBlockLevels; L++)
      for (int cu = 0; cu < 2; cu++) {
        g.listOff[par][L][cu] = (int)g.hlist.size();
        for (int b = 0; b < nb; b++) {
          if (g.hL[b] != L || (g.hbflags[b] & kPsBlkNoFluid)) continue;
          if (((g.hbflags[b] & kPsBlkCutsSolid) != 0) != (cu != 0)) continue;
          int bc[3];
          block_xyz(ps, b, bc);
          if (((bc[0] + bc[1] + bc[2]) & 1) == par) g.hlist.push_back(b);
        }
        g.listLen[par][L][cu] = (int)g.hlist.size() - g.listOff[par][L][cu];
      }
  g.hitems.resize(g.hlist.size());
  for (size_t k = 0; k < g.hlist.size(); k++) {
    const int b = g.hlist[k];
    PsItem &it = g.hitems[k];
    std::memset(&it, 0, sizeof(it));
    it.b = b;
    it.off = g.hoff[b];
    it.L = g.hL[b];
    for (int f = 0; f < 6; f++) {
      const int b2 = ps.hnbr[6 * b + f];
      it.nbr[f] = b2;
      it.noff[f] = b2 >= 0 ? g.hoff[b2] : 0;
      it.nL[f] = b2 >= 0 ? g.hL[b2] : 0;
    }
  }
  MSBG_TRY(g.D.alloc(n));
  MSBG_HIP(hipMemcpy(g.D.p, D.data(), n * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

int build_dense_level(msbg_ps &ps, int m) {
  MgLevel &g = ps.mg[m];
  const HostLevel &G = ps.g[m];
  g.dense = true;
  g.sx = G.sx, g.sy = G.sy, g.sz = G.sz;
  const int px = g.sx + 2, py = g.sy + 2;
  g.n = (int64_t)px * py * (g.sz + 2);
  auto P = [&](int x, int y, int z) { return (x + 1) + (int64_t)px * ((y + 1) + (int64_t)py * (z + 1)); };
  std::vector<float> fa[3];
  for (int d = 0; d < 3; d++) {
    fa[d].assign(g.n, 0.0f);
    const int ex = g.sx + (d == 0), ey = g.sy + (d == 1), ez = g.sz + (d == 2);
    for (int z = 0; z < ez; z++)
      for (int y = 0; y < ey; y++)
        for (int x = 0; x < ex; x++) fa[d][P(x, y, z)] = G.fa[d][x + (int64_t)ex * (y + (int64_t)ey * z)];
  }
  g.hflags.assign(g.n, kPsCellSolid);  // the ghost shell reads as solid
  for (int z = 0; z < g.sz; z++)
    for (int y = 0; y < g.sy; y++)
      for (int x = 0; x < g.sx; x++)
        g.hflags[P(x, y, z)] = G.flags[x + (int64_t)g.sx * (y + (int64_t)g.sy * z)];
  std::vector<float> D(g.n, 0.0f);
  const int64_t st[3] = {1, px, (int64_t)px * py};
  for (int z = 0; z < g.sz; z++)
    for (int y = 0; y < g.sy; y++)
      for (int x = 0; x < g.sx; x++) {
        const int64_t p = P(x, y, z);
        if (!fluid(g.hflags[p])) continue;
        bool gated = false;
        for (int k = x & ~3; k < (x & ~3) + 4 && k < g.sx; k++)
          gated |= (g.hflags[p - x + k] & (kPsCellSolid | kPsCellPartialSolid)) != 0;
        double sum = 0;
        for (int d = 0; d < 3; d++) {
          sum += gated ? (double)fa[d][p] : 1.0;
          sum += gated ? (double)fa[d][p + st[d]] : 1.0;
        }
        D[p] = sum > 0 ? (float)(1.0 / sum) : 0.0f;
      }
  MSBG_TRY(g.D.alloc(g.n));
  MSBG_HIP(hipMemcpy(g.D.p, D.data(), g.n * sizeof(float), hipMemcpyHostToDevice));
  for (int d = 0; d < 3; d++) {
    MSBG_TRY(g.fa[d].alloc(g.n));
    MSBG_HIP(hipMemcpy(g.fa[d].p, fa[d].data(), g.n * sizeof(float), hipMemcpyHostToDevice));
  }
  return 0;
}

int upload_level(msbg_ps &ps, int m) {
  MgLevel &g = ps.mg[m];
  for (int c = 0; c < 4; c++) {
    MSBG_TRY(g.ch[c].alloc(g.n));
    MSBG_TRY(launch_ps_fill(ps.stream, g.ch[c].p, g.n, 0.0f));
  }
  MSBG_TRY(g.flags.alloc(g.n));
  MSBG_HIP(hipMemcpy(g.flags.p, g.hflags.data(), g.n * sizeof(uint16_t), hipMemcpyHostToDevice));
  if (!g.dense) {
    MSBG_TRY(g.L.alloc(ps.nb));
    MSBG_TRY(g.off.alloc(ps.nb));
    MSBG_TRY(g.bflags.alloc(ps.nb));
    MSBG_HIP(hipMemcpy(g.L.p, g.hL.data(), ps.nb, hipMemcpyHostToDevice));
    MSBG_HIP(hipMemcpy(g.off.p, g.hoff.data(), ps.nb * sizeof(int64_t), hipMemcpyHostToDevice));
    MSBG_HIP(hipMemcpy(g.bflags.p, g.hbflags.data(), ps.nb * sizeof(uint16_t), hipMemcpyHostToDevice));
    if (!g.hitems.empty()) {
      MSBG_TRY(g.list.alloc(g.hitems.size()));
      MSBG_HIP(hipMemcpy(g.list.p, g.hitems.data(), g.hitems.size() * sizeof(PsItem), hipMemcpyHostToDevice));
    }
  }
  return 0;
}

PsBlockArgs block_args(msbg_ps &ps, int m, int cx, int cb, int cdst) {
  MgLevel &g = ps.mg[m];
  PsBlockArgs a{};
  a.X = g.ch[cx].p;
  a.B = cb >= 0 ? g.ch[cb].p : nullptr;
  a.Y = g.ch[cdst].p;
  a.D = g.D.p;
  a.flags = g.flags.p;
  a.bflags = g.bflags.p;
  for (int l = 0; l < kPsMaxBlockLevels; l++)
    for (int d = 0; d < 3; d++) a.fa[l][d] = ps.dpfa[l][d].p;
  a.dx0 = ps.desc.dx0;
  a.omega = ps.desc.omega;
  for (int d = 0; d < 3; d++) a.openRight[d] = ps.desc.domainBc[d][1] == MSBG_DBC_OPEN;
  a.biter = ps.desc.blockIter;
  return a;
}

PsDenseArgs dense_args(msbg_ps &ps, int m, const float *X, const float *B, float *Y) {
  MgLevel &g = ps.mg[m];
  PsDenseArgs a{};
  a.sx = g.sx, a.sy = g.sy, a.sz = g.sz;
  a.X = X, a.B = B, a.Y = Y;
  a.D = g.D.p;
  a.flags = g.flags.p;
  for (int d = 0; d < 3; d++) a.fa[d] = g.fa[d].p;
  a.h = (float)((double)(1 << m) * ps.desc.dx0);
  const float omega = ps.desc.omega;
  a.hromega = omega * 1.0f / a.h;
  a.om1 = 1.0f - omega;
  return a;
}

PsXferArgs xfer_args(msbg_ps &ps, int m) {
  const MgLevel &f = ps.mg[m], &c = ps.mg[m + 1];
  PsXferArgs a{};
  a.kind = !c.dense ? 0 : !f.dense ? 1 : 2;
  a.nbx = ps.desc.nbx, a.nby = ps.desc.nby, a.nbz = ps.desc.nbz, a.bsx0 = ps.desc.bsx0;
  a.Lf = f.L.p, a.Lc = c.L.p, a.offF = f.off.p, a.offC = c.off.p;
  a.fsx = f.sx, a.fsy = f.sy, a.fsz = f.sz;
  a.csx = c.sx, a.csy = c.sy, a.csz = c.sz;
  a.flagsF = f.flags.p, a.flagsC = c.flags.p;
  return a;
}

bool chan_ok(int c) { return c >= MSBG_PS_CH_X && c <= MSBG_PS_CH_T; }

int relax_impl(msbg_ps *ps, int reverseOrder, int m, int chX, int chB, int chTmp, int nIter) {
  MgLevel &g = ps->mg[m];
  if (g.dense) {  // msbg3.cpp:1917-2165: Jacobi, ping-pong through chanTmp
    int rd = chTmp, wr = chX;
    for (int it = 0; it < nIter; it++) {
      std::swap(rd, wr);
      PsDenseArgs a = dense_args(*ps, m, g.ch[rd].p, g.ch[chB].p, g.ch[wr].p);
      MSBG_TRY(launch_ps_dense(ps->stream, PS_MODE_RELAX, a));
    }
    if (wr != chX) MSBG_TRY(launch_ps_copy(ps->stream, g.ch[wr].p, g.ch[chX].p, g.n));
    return 0;
  }
  PsBlockArgs a = block_args(*ps, m, chX, chB, chX);
  a.rev = reverseOrder ? 1 : 0;
  for (int it = 0; it < nIter; it++)
    for (int pass = 0; pass < 2; pass++) {
      // relaxBlockList (msbg3.cpp:2187-2208): pass 0 takes the blocks whose
      // coordinate sum has parity 1 (parity 0 with OPT_REVERSE_ORDER)
      const int par = reverseOrder ? pass : 1 - pass;
      for (int L = 0; L < ps->nlev; L++)
        for (int cu = 0; cu < 2; cu++) {
          const int n = g.listLen[par][L][cu];
          if (!n) continue;
          MSBG_TRY(launch_ps_block(ps->stream, ps->desc.bsx0 >> L, PS_MODE_RELAX, cu != 0, a,
                                   g.list.p + g.listOff[par][L][cu], n, nullptr));
        }
    }
  return 0;
}

int multiply_impl(msbg_ps *ps, unsigned options, int m, int chX, int chB, int chDst, double *dot) {
  MgLevel &g = ps->mg[m];
  const bool resid = options & MSBG_OPT_CALC_RESIDUAL;
  if (g.dense) {
    PsDenseArgs a = dense_args(*ps, m, g.ch[chX].p, g.ch[chB].p, g.ch[chDst].p);
    return launch_ps_dense(ps->stream, resid ? PS_MODE_RESIDUAL : PS_MODE_MULTIPLY, a);
  }
  PsBlockArgs a = block_args(*ps, m, chX, chB, chDst);
  // the multiply kernel writes one partial dot product per workgroup
  int total = 0;
  for (int par = 0; par < 2; par++)
    for (int L = 0; L < ps->nlev; L++)
      for (int cu = 0; cu < 2; cu++) total += ps_block_workgroups(ps->desc.bsx0 >> L, g.listLen[par][L][cu]);
  if (!resid) MSBG_TRY(ps->dot.ensure(std::max(total, 1)));
  int wgOff = 0;
  for (int L = 0; L < ps->nlev; L++)
    for (int par = 0; par < 2; par++)
      for (int cu = 0; cu < 2; cu++) {
        const int n = g.listLen[par][L][cu];
        if (!n) continue;
        a.dot = resid ? nullptr : ps->dot.p + wgOff;
        int wg = 0;
        MSBG_TRY(launch_ps_block(ps->stream, ps->desc.bsx0 >> L, resid ? PS_MODE_RESIDUAL : PS_MODE_MULTIPLY,
                                 cu != 0, a, g.list.p + g.listOff[par][L][cu], n, &wg));
        wgOff += wg;
      }
  if (dot) {
    double s = 0;
    if (!resid && wgOff) {
      std::vector<double> h(wgOff);
      MSBG_HIP(hipMemcpyAsync(h.data(), ps->dot.p, wgOff * sizeof(double), hipMemcpyDeviceToHost, ps->stream));
      MSBG_HIP(hipStreamSynchronize(ps->stream));
      for (double v : h) s += v;
    }
    *dot = s;
  }
  return 0;
}

int vcycle_level(msbg_ps *ps, int m, int nu1, int nu2, int coarseIter) {
  if (m == ps->nmg - 1)
    return relax_impl(ps, 0, m, MSBG_PS_CH_X, MSBG_PS_CH_B, MSBG_PS_CH_T, coarseIter);
  MSBG_TRY(relax_impl(ps, 0, m, MSBG_PS_CH_X, MSBG_PS_CH_B, MSBG_PS_CH_T, nu1));
  MSBG_TRY(launch_ps_fill(ps->stream, ps->mg[m].ch[MSBG_PS_CH_R].p, ps->mg[m].n, 0.0f));
  MSBG_TRY(multiply_impl(ps, MSBG_OPT_CALC_RESIDUAL, m, MSBG_PS_CH_X, MSBG_PS_CH_B, MSBG_PS_CH_R, nullptr));
  PsXferArgs a = xfer_args(*ps, m);
  MSBG_TRY(launch_ps_restrict(ps->stream, a, ps->mg[m].ch[MSBG_PS_CH_R].p, ps->mg[m + 1].ch[MSBG_PS_CH_B].p));
  MSBG_TRY(launch_ps_fill(ps->stream, ps->mg[m + 1].ch[MSBG_PS_CH_X].p, ps->mg[m + 1].n, 0.0f));
  MSBG_TRY(vcycle_level(ps, m + 1, nu1, nu2, coarseIter));
  MSBG_TRY(launch_ps_prolongate(ps->stream, a, ps->mg[m + 1].ch[MSBG_PS_CH_X].p, ps->mg[m].ch[MSBG_PS_CH_X].p));
  return relax_impl(ps, 1, m, MSBG_PS_CH_X, MSBG_PS_CH_B, MSBG_PS_CH_T, nu2);
}

}  // namespace

extern "C" {

int msbg_ps_create(const msbg_ps_desc *d, const int8_t *blockLevels, const uint8_t *cellType,
                   msbg_ps **out) {
  MSBG_CHECK(d && blockLevels && cellType && out, MSBG_EINVAL, "msbg_ps_create: null argument");
  *out = nullptr;
  const int bsx0 = d->bsx0;
  MSBG_CHECK(d->nbx >= 1 && d->nby >= 1 && d->nbz >= 1 && d->nbx <= 4096 && d->nby <= 4096 && d->nbz <= 4096,
             MSBG_EINVAL, "msbg_ps_create: block grid %dx%dx%d", d->nbx, d->nby, d->nbz);
  MSBG_CHECK(bsx0 == 4 || bsx0 == 8 || bsx0 == 16, MSBG_EINVAL,
             "msbg_ps_create: block edge %d (4, 8 or 16)", bsx0);
  MSBG_CHECK(d->blockIter >= 1 && d->denseLevels >= 0, MSBG_EINVAL, "msbg_ps_create: blockIter/denseLevels");
  MSBG_CHECK(d->dx0 > 0 && std::isfinite(d->dx0) && d->omega > 0 && d->omega < 2, MSBG_EINVAL,
             "msbg_ps_create: dx0/omega");
  for (int a = 0; a < 3; a++)
    for (int s = 0; s < 2; s++)
      MSBG_CHECK(d->domainBc[a][s] == MSBG_DBC_OPEN || d->domainBc[a][s] == MSBG_DBC_SOLID, MSBG_EUNSUPP,
                 "msbg_ps_create: domain boundary type %d (open or solid)", d->domainBc[a][s]);
  int nlev = 0;
  while ((2 << nlev) <= bsx0) nlev++;  // block edges bsx0 ... 2
  MSBG_CHECK(nlev + d->denseLevels <= kMaxMg, MSBG_EINVAL, "msbg_ps_create: too many levels");
  for (int k = 0; k < d->denseLevels; k++) {
    const int f = 1 << k;  // dense level k has nb >> k cells per axis
    MSBG_CHECK(d->nbx % f == 0 && d->nby % f == 0 && d->nbz % f == 0, MSBG_EINVAL,
               "msbg_ps_create: dense level %d needs block counts divisible by %d", k, f);
  }
  int ndev = 0;
  MSBG_CHECK(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, MSBG_EDEVICE, "msbg_ps_create: no HIP device");

  std::unique_ptr<msbg_ps> ps(new msbg_ps);
  ps->desc = *d;
  ps->nb = d->nbx * d->nby * d->nbz;
  ps->nlev = nlev;
  ps->nmg = nlev + d->denseLevels;
  ps->level.assign(blockLevels, blockLevels + ps->nb);
  ps->hnbr.resize(6 * ps->nb);
  for (int b = 0; b < ps->nb; b++) {
    MSBG_CHECK(ps->level[b] >= 0 && ps->level[b] < nlev, MSBG_EINVAL, "msbg_ps_create: block %d level %d", b,
               ps->level[b]);
    int c[3];
    block_xyz(*ps, b, c);
    for (int a = 0; a < 3; a++)
      for (int s = 0; s < 2; s++) {
        int q[3] = {c[0], c[1], c[2]};
        q[a] += s ? 1 : -1;
        const int lim[3] = {d->nbx, d->nby, d->nbz};
        int b2 = -1;
        if (q[a] >= 0 && q[a] < lim[a]) b2 = q[0] + d->nbx * (q[1] + d->nby * q[2]);
        ps->hnbr[6 * b + 2 * a + s] = b2;
      }
  }
  for (int b = 0; b < ps->nb; b++)
    for (int f = 0; f < 6; f++) {
      const int b2 = ps->hnbr[6 * b + f];
      MSBG_CHECK(b2 < 0 || std::abs(ps->level[b] - ps->level[b2]) <= 1, MSBG_EINVAL,
                 "msbg_ps_create: blocks %d and %d differ by more than one level", b, b2);
    }
  for (int l = 0; l < ps->nmg; l++) {
    ps->g[l].sx = (d->nbx * bsx0) >> l;
    ps->g[l].sy = (d->nby * bsx0) >> l;
    ps->g[l].sz = (d->nbz * bsx0) >> l;
  }
  const size_t n0 = (size_t)ps->g[0].sx * ps->g[0].sy * ps->g[0].sz;
  ps->g[0].type.assign(cellType, cellType + n0);
  for (size_t i = 0; i < n0; i++)
    MSBG_CHECK(cellType[i] <= MSBG_PS_VOID, MSBG_EINVAL, "msbg_ps_create: cell type %d", cellType[i]);
  MSBG_HIP(hipStreamCreate(&ps->stream));
  MSBG_HIP(hipEventCreate(&ps->ev[0]));
  MSBG_HIP(hipEventCreate(&ps->ev[1]));
  build_levels(*ps);
  for (int l = 0; l < nlev; l++) {
    const int bsx = bsx0 >> l;
    const HostLevel &G = ps->g[l];
    for (int a = 0; a < 3; a++) {
      std::vector<float> &p = ps->pfa[l][a];
      p.resize((size_t)ps->nb * bsx * bsx * bsx);
      for (int b = 0; b < ps->nb; b++) {
        int c[3];
        block_xyz(*ps, b, c);
        for (int z = 0; z < bsx; z++)
          for (int y = 0; y < bsx; y++)
            for (int x = 0; x < bsx; x++)
              p[(int64_t)b * bsx * bsx * bsx + vox(bsx, x, y, z)] =
                  G.fa[a][fa_index(G, a, c[0] * bsx + x, c[1] * bsx + y, c[2] * bsx + z)];
      }
      MSBG_TRY(ps->dpfa[l][a].alloc(p.size()));
      MSBG_HIP(hipMemcpy(ps->dpfa[l][a].p, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice));
    }
  }
  for (int m = 0; m < ps->nmg; m++) {
    MSBG_TRY(m < nlev ? build_block_level(*ps, m) : build_dense_level(*ps, m));
    MSBG_TRY(upload_level(*ps, m));
  }
  MSBG_HIP(hipStreamSynchronize(ps->stream));
  *out = ps.release();
  return 0;
}

int msbg_ps_destroy(msbg_ps *ps) {
  delete ps;
  return 0;
}

int msbg_ps_get_levels(const msbg_ps *ps, int *nLevelsMg, int *nLevelsBlock) {
  MSBG_CHECK(ps, MSBG_EINVAL, "msbg_ps_get_levels: null handle");
  if (nLevelsMg) *nLevelsMg = ps->nmg;
  if (nLevelsBlock) *nLevelsBlock = ps->nlev;
  return 0;
}

int msbg_ps_level_size(const msbg_ps *ps, int levelMg, int64_t *n) {
  MSBG_CHECK(ps && n && levelMg >= 0 && levelMg < ps->nmg, MSBG_EINVAL, "msbg_ps_level_size: bad args");
  *n = ps->mg[levelMg].n;
  return 0;
}

int msbg_ps_upload(msbg_ps *ps, int chan, int levelMg, const float *host) {
  MSBG_CHECK(ps && host && chan_ok(chan) && levelMg >= 0 && levelMg < ps->nmg, MSBG_EINVAL,
             "msbg_ps_upload: bad args");
  const MgLevel &g = ps->mg[levelMg];
  MSBG_HIP(hipMemcpyAsync(g.ch[chan].p, host, g.n * sizeof(float), hipMemcpyHostToDevice, ps->stream));
  MSBG_HIP(hipStreamSynchronize(ps->stream));
  return 0;
}

int msbg_ps_download(msbg_ps *ps, int chan, int levelMg, float *host) {
  MSBG_CHECK(ps && host && (chan_ok(chan) || chan == MSBG_PS_CH_D) && levelMg >= 0 && levelMg < ps->nmg,
             MSBG_EINVAL, "msbg_ps_download: bad args");
  const MgLevel &g = ps->mg[levelMg];
  const float *src = chan == MSBG_PS_CH_D ? g.D.p : g.ch[chan].p;
  MSBG_HIP(hipMemcpyAsync(host, src, g.n * sizeof(float), hipMemcpyDeviceToHost, ps->stream));
  MSBG_HIP(hipStreamSynchronize(ps->stream));
  return 0;
}

int msbg_ps_download_flags(msbg_ps *ps, int levelMg, uint16_t *cellFlags, uint16_t *blockFlags) {
  MSBG_CHECK(ps && levelMg >= 0 && levelMg < ps->nmg, MSBG_EINVAL, "msbg_ps_download_flags: bad args");
  const MgLevel &g = ps->mg[levelMg];
  if (cellFlags) {
    MSBG_HIP(hipMemcpyAsync(cellFlags, g.flags.p, g.n * sizeof(uint16_t), hipMemcpyDeviceToHost, ps->stream));
    MSBG_HIP(hipStreamSynchronize(ps->stream));
  }
  if (blockFlags) {
    MSBG_CHECK(!g.dense, MSBG_EINVAL, "msbg_ps_download_flags: level %d is dense", levelMg);
    std::memcpy(blockFlags, g.hbflags.data(), ps->nb * sizeof(uint16_t));
  }
  return 0;
}

int msbg_ps_fill(msbg_ps *ps, int chan, int levelMg, float v) {
  MSBG_CHECK(ps && chan_ok(chan) && levelMg >= 0 && levelMg < ps->nmg, MSBG_EINVAL, "msbg_ps_fill: bad args");
  return launch_ps_fill(ps->stream, ps->mg[levelMg].ch[chan].p, ps->mg[levelMg].n, v);
}

int msbg_ps_relax(msbg_ps *ps, int boundaryZoneOnly, int reverseOrder, int levelMg, int chanX, int chanB,
                  int chanTmp, int nIter) {
  MSBG_CHECK(ps && levelMg >= 0 && levelMg < ps->nmg && chan_ok(chanX) && chan_ok(chanB) && chanX != chanB &&
                 nIter >= 0,
             MSBG_EINVAL, "msbg_ps_relax: bad args");
  MSBG_CHECK(!boundaryZoneOnly, MSBG_EUNSUPP, "msbg_ps_relax: boundary-zone relaxation is not implemented");
  if (ps->mg[levelMg].dense)
    MSBG_CHECK(chan_ok(chanTmp) && chanTmp != chanX && chanTmp != chanB, MSBG_EINVAL,
               "msbg_ps_relax: a dense level needs a distinct chanTmp");
  if (!nIter) return 0;
  MSBG_HIP(hipEventRecord(ps->ev[0], ps->stream));
  MSBG_TRY(relax_impl(ps, reverseOrder, levelMg, chanX, chanB, chanTmp, nIter));
  MSBG_HIP(hipEventRecord(ps->ev[1], ps->stream));
  ps->timed = true;
  return 0;
}

int msbg_ps_multiply(msbg_ps *ps, unsigned options, int levelMg, int chanX, int chanB, float kap,
                     int chanDstY, double *dotprod) {
  (void)kap;  // accepted and unused, as in multiplyLaplacianMatrixOpt's body
  MSBG_CHECK(ps && levelMg >= 0 && levelMg < ps->nmg && chan_ok(chanX) && chan_ok(chanDstY) && chanX != chanDstY,
             MSBG_EINVAL, "msbg_ps_multiply: bad args");
  const bool resid = options & MSBG_OPT_CALC_RESIDUAL;
  MSBG_CHECK(!resid || (chan_ok(chanB) && chanB != chanDstY), MSBG_EINVAL, "msbg_ps_multiply: residual needs chanB");
  MSBG_CHECK(!(ps->mg[levelMg].dense && dotprod), MSBG_EINVAL,
             "msbg_ps_multiply: no dot product on a dense level (msbg3.cpp:1804)");
  MSBG_CHECK(!(options & MSBG_OPT_BOUNDZONE_ONLY), MSBG_EUNSUPP, "msbg_ps_multiply: boundary zone only");
  if (!resid) chanB = chanX;  // unused
  MSBG_HIP(hipEventRecord(ps->ev[0], ps->stream));
  MSBG_TRY(multiply_impl(ps, options, levelMg, chanX, chanB, chanDstY, dotprod));
  MSBG_HIP(hipEventRecord(ps->ev[1], ps->stream));
  ps->timed = true;
  return 0;
}

int msbg_ps_restrict(msbg_ps *ps, int levelMg, int chanSrc, int chanDst) {
  MSBG_CHECK(ps && levelMg >= 0 && levelMg + 1 < ps->nmg && chan_ok(chanSrc) && chan_ok(chanDst), MSBG_EINVAL,
             "msbg_ps_restrict: bad args");
  PsXferArgs a = xfer_args(*ps, levelMg);
  return launch_ps_restrict(ps->stream, a, ps->mg[levelMg].ch[chanSrc].p, ps->mg[levelMg + 1].ch[chanDst].p);
}

int msbg_ps_prolongate(msbg_ps *ps, int levelMg, int chanSrc, int chanDst) {
  MSBG_CHECK(ps && levelMg >= 0 && levelMg + 1 < ps->nmg && chan_ok(chanSrc) && chan_ok(chanDst), MSBG_EINVAL,
             "msbg_ps_prolongate: bad args");
  PsXferArgs a = xfer_args(*ps, levelMg);
  return launch_ps_prolongate(ps->stream, a, ps->mg[levelMg + 1].ch[chanSrc].p, ps->mg[levelMg].ch[chanDst].p);
}

int msbg_ps_vcycle(msbg_ps *ps, int nu1, int nu2, int coarseIter) {
  MSBG_CHECK(ps && nu1 >= 0 && nu2 >= 0 && coarseIter >= 0, MSBG_EINVAL, "msbg_ps_vcycle: bad args");
  MSBG_HIP(hipEventRecord(ps->ev[0], ps->stream));
  MSBG_TRY(vcycle_level(ps, 0, nu1, nu2, coarseIter));
  MSBG_HIP(hipEventRecord(ps->ev[1], ps->stream));
  ps->timed = true;
  return 0;
}

int msbg_ps_last_timing(msbg_ps *ps, double *ms) {
  MSBG_CHECK(ps && ms, MSBG_EINVAL, "msbg_ps_last_timing: bad args");
  MSBG_CHECK(ps->timed, MSBG_ESTATE, "msbg_ps_last_timing: no relax/multiply/vcycle call yet");
  MSBG_HIP(hipEventSynchronize(ps->ev[1]));
  float t = 0;
  MSBG_HIP(hipEventElapsedTime(&t, ps->ev[0], ps->ev[1]));
  *ms = t;
  return 0;
}

int msbg_ps_relaxed_cells(const msbg_ps *ps, int levelMg, int64_t *cells) {
  MSBG_CHECK(ps && cells && levelMg >= 0 && levelMg < ps->nmg, MSBG_EINVAL, "msbg_ps_relaxed_cells: bad args");
  const MgLevel &g = ps->mg[levelMg];
  int64_t n = 0;
  if (g.dense) {
    n = (int64_t)g.sx * g.sy * g.sz;
  } else {
    for (const PsItem &it : g.hitems) {
      const int e = ps->desc.bsx0 >> it.L;
      n += (int64_t)e * e * e;
    }
  }
  *cells = n;
  return 0;
}

int msbg_ps_synchronize(msbg_ps *ps) {
  MSBG_CHECK(ps, MSBG_EINVAL, "msbg_ps_synchronize: null handle");
  MSBG_HIP(hipStreamSynchronize(ps->stream));
  return 0;
}

}  // extern "C"
