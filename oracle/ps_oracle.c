/*
 * ps_oracle.c -- CPU restatement of MSBG's adaptive pressure relax.
 * TEST INFRASTRUCTURE ONLY (see ps_oracle.h). Built with -ffp-contract=off:
 * every fused multiply-add below is an explicit fmaf, every other operation
 * rounds on its own, so the result is fixed by the source.
 *
 * Arithmetic fixed by this restatement (the reference compiles with
 * -ffast-math, so its own rounding there is not fixed):
 *   relax   Y = fmaf(1-omega, F0, (omega*(B+S))*D)      msbg3.cpp:1292
 *   S       plain sum F1+..+F6, or fmaf(Fi, Wi, S) i=1..6 msbg3.cpp:1256-1266
 *   multiply Y = D != 0 ? (1/D)*F0 - S : 0 [, B - Y]    msbg3.cpp:1271-1278
 *   dense Jacobi Y = fmaf(1-omega, F0, ((B+S)*D)*(omega/h)), S = (sum F)*h
 *                                                       msbg3.cpp:355-389
 */
#include "ps_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 8
#define MAXMG 14

typedef struct {
  int dense;
  int sx, sy, sz;   /* dense levels: cells per axis */
  int64_t n;        /* floats per channel */
  int64_t *off;     /* block levels: page offset per block */
  int8_t *L;        /* block levels: effective level per block */
  float *ch[4];     /* X, B, R, T */
  float *D;
  uint16_t *flags;
  uint16_t *bflags; /* block levels */
  float *fa[3];     /* dense levels: face areas, padded layout */
} mgl_t;

struct or_ps {
  int nbx, nby, nbz, nb, bsx0, nlev, ndense, nmg;
  double dx0;
  float omega;
  int biter;
  int dbc[6];
  int8_t *level;
  /* per level l (all levels): global dense types and face areas */
  int gsx[MAXMG], gsy[MAXMG], gsz[MAXMG];
  uint8_t *type[MAXMG];
  float *gfa[MAXMG][3];      /* u: (sx+1)*sy*sz, v: sx*(sy+1)*sz, w: sx*sy*(sz+1) */
  uint16_t *gflags[MAXMG];
  /* block levels: page-form face areas (left faces) of every block */
  float *pfa[MAXL][3];
  mgl_t mg[MAXMG];
};

static int ilog2(int v) { int l = 0; while ((1 << l) < v) l++; return l; }
static int simdw(int bsx) { return bsx >= 8 ? 8 : bsx; } /* msbg3.cpp:1690-1738, float build */

/* ---------------------------------------------------------------- set-up */

static int gidx(const or_ps *ps, int l, int x, int y, int z) {
  return x + ps->gsx[l] * (y + ps->gsy[l] * z);
}

/* type at level l, outside the domain: open -> void, solid -> solid
 * (msbg.h:2880) */
static int type_at(const or_ps *ps, int l, int x, int y, int z) {
  int d = -1, s = 0;
  if (x < 0) { d = 0; s = 0; } else if (x >= ps->gsx[l]) { d = 0; s = 1; }
  else if (y < 0) { d = 1; s = 0; } else if (y >= ps->gsy[l]) { d = 1; s = 1; }
  else if (z < 0) { d = 2; s = 0; } else if (z >= ps->gsz[l]) { d = 2; s = 1; }
  if (d >= 0) return ps->dbc[2 * d + s] == PS_DBC_OPEN ? PS_VOID : PS_SOLID;
  return ps->type[l][gidx(ps, l, x, y, z)];
}

static int64_t fa_idx(const or_ps *ps, int l, int d, int x, int y, int z) {
  int sx = ps->gsx[l] + (d == 0), sy = ps->gsy[l] + (d == 1);
  return x + (int64_t)sx * (y + (int64_t)sy * z);
}

static void build_global(or_ps *ps) {
  int nl = ps->nmg;
  for (int l = 1; l < nl; l++) {
    int sx = ps->gsx[l], sy = ps->gsy[l], sz = ps->gsz[l];
    ps->type[l] = malloc((size_t)sx * sy * sz);
    for (int z = 0; z < sz; z++)
      for (int y = 0; y < sy; y++)
        for (int x = 0; x < sx; x++) {
          int anyf = 0, anyv = 0;
          for (int k = 0; k < 8; k++) {
            int t = ps->type[l - 1][gidx(ps, l - 1, 2 * x + (k & 1), 2 * y + ((k >> 1) & 1),
                                          2 * z + (k >> 2))];
            anyf |= t == PS_FLUID;
            anyv |= t == PS_VOID;
          }
          ps->type[l][gidx(ps, l, x, y, z)] = anyf ? PS_FLUID : anyv ? PS_VOID : PS_SOLID;
        }
  }
  for (int l = 0; l < nl; l++) {
    int sx = ps->gsx[l], sy = ps->gsy[l], sz = ps->gsz[l];
    for (int d = 0; d < 3; d++) {
      int ex = sx + (d == 0), ey = sy + (d == 1), ez = sz + (d == 2);
      float *fa = ps->gfa[l][d] = malloc(sizeof(float) * (size_t)ex * ey * ez);
      for (int z = 0; z < ez; z++)
        for (int y = 0; y < ey; y++)
          for (int x = 0; x < ex; x++) {
            float v;
            if (l == 0) {
              /* face between the cell and its lower neighbour along d */
              int ta = type_at(ps, 0, x - (d == 0), y - (d == 1), z - (d == 2));
              int tc = type_at(ps, 0, x, y, z);
              v = (ta == PS_SOLID || tc == PS_SOLID) ? 0.0f : 1.0f;
            } else {
              /* mean of the four child faces, in (a,b) order (0,0),(1,0),(0,1),(1,1)
               * over the two tangential axes (lower axis first) */
              const float *f = ps->gfa[l - 1][d];
              float s = 0.0f;
              for (int k = 0; k < 4; k++) {
                int a = k & 1, b = k >> 1;
                int cx = 2 * x, cy = 2 * y, cz = 2 * z;
                if (d == 0) { cy += a; cz += b; }
                else if (d == 1) { cx += a; cz += b; }
                else { cx += a; cy += b; }
                s = s + f[fa_idx(ps, l - 1, d, cx, cy, cz)];
              }
              v = s * 0.25f;
            }
            fa[x + (int64_t)ex * (y + (int64_t)ey * z)] = v;
          }
    }
    ps->gflags[l] = malloc(sizeof(uint16_t) * (size_t)sx * sy * sz);
    for (int z = 0; z < sz; z++)
      for (int y = 0; y < sy; y++)
        for (int x = 0; x < sx; x++) {
          int t = ps->type[l][gidx(ps, l, x, y, z)];
          uint16_t f = 0;
          if (t == PS_SOLID) f = PS_CELL_SOLID;
          else if (t == PS_VOID) f = PS_CELL_VOID;
          else {
            float m = 1.0f;
            for (int d = 0; d < 3; d++)
              for (int s = 0; s < 2; s++) {
                float v = ps->gfa[l][d][fa_idx(ps, l, d, x + s * (d == 0), y + s * (d == 1),
                                               z + s * (d == 2))];
                if (v < m) m = v;
              }
            if (m < 1.0f) f = PS_CELL_PARTIAL_SOLID;
          }
          ps->gflags[l][gidx(ps, l, x, y, z)] = f;
        }
  }
}

static void block_coords(const or_ps *ps, int b, int *bx, int *by, int *bz) {
  *bx = b % ps->nbx;
  *by = (b / ps->nbx) % ps->nby;
  *bz = b / (ps->nbx * ps->nby);
}

/* neighbour block across face (d, s), -1 outside the domain */
static int neigh_block(const or_ps *ps, int b, int d, int s) {
  int c[3];
  block_coords(ps, b, &c[0], &c[1], &c[2]);
  c[d] += s ? 1 : -1;
  if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= ps->nbx || c[1] >= ps->nby || c[2] >= ps->nbz)
    return -1;
  return c[0] + ps->nbx * (c[1] + ps->nby * c[2]);
}

static int vox(int bsx, int x, int y, int z) { return x + bsx * (y + bsx * z); }

/* per-vector gating of one block page: bit0 any SOLID|PARTIAL_SOLID,
 * bit1 no fluid lane (msbg3.cpp:580-603) */
static void vec_gates(const uint16_t *fl, int bsx, uint8_t *g) {
  int sw = simdw(bsx), nv = bsx * bsx * bsx / sw;
  for (int v = 0; v < nv; v++) {
    int any = 0, fluid = 0;
    for (int k = 0; k < sw; k++) {
      uint16_t f = fl[v * sw + k];
      any |= (f & (PS_CELL_SOLID | PS_CELL_PARTIAL_SOLID)) != 0;
      fluid |= (f & (PS_CELL_SOLID | PS_CELL_VOID)) == 0;
    }
    g[v] = (uint8_t)(any | ((!fluid) << 1));
  }
}

/* coefficient a of one coarse-fine contribution (DPMR_GRAD_COEFF_CF = 1/3,
 * msbg.h:85, procNeighMixedRes2 passes h = 1), times the face area when the
 * block cuts a solid (msbg3.cpp:48-56, 144-145) */
static float a_cf(void) { return (float)(1.0 * (1. / (4 * (3. / 4.)))); }
static float a_fc(void) { return (float)(1.0 * 2 * (1. / (4 * (3. / 4.)))); }

/* mixed-resolution face value of block b (level L, mg level m) at face cell
 * (vx,vy,vz) across face (d, s) with neighbour b2 at level L2; returns the
 * sum of a*x as procNeighMixedRes2 and the coefficient sum in *csum
 * (msbg3.cpp:59-160) */
static double mixed_sum(const or_ps *ps, int m, int chan, int b, int b2, int d, int s, int vx,
                        int vy, int vz, double *csum) {
  const mgl_t *g = &ps->mg[m];
  int L = g->L[b], L2 = g->L[b2], bsx = ps->bsx0 >> L;
  int cuts = (g->bflags[b] & PS_BLK_CUTS_SOLID) != 0;
  const float *X2 = g->ch[chan] + g->off[b2];
  double sum = 0, cs = 0;
  if (L2 < L) { /* neighbour finer */
    int bh = 2 * bsx;
    const float *fa = ps->pfa[L2][d] + (int64_t)(s ? b2 : b) * bh * bh * bh;
    for (int k = 0; k < 4; k++) {
      int p = k & 1, q = k >> 1, x2, y2, z2, fx, fy, fz;
      if (d == 0) { x2 = s ? 0 : bh - 1; y2 = 2 * vy + p; z2 = 2 * vz + q; fx = 0; fy = y2; fz = z2; }
      else if (d == 1) { y2 = s ? 0 : bh - 1; x2 = 2 * vx + p; z2 = 2 * vz + q; fx = x2; fy = 0; fz = z2; }
      else { z2 = s ? 0 : bh - 1; x2 = 2 * vx + p; y2 = 2 * vy + q; fx = x2; fy = y2; fz = 0; }
      float a = a_cf();
      if (cuts) a = a * fa[vox(bh, fx, fy, fz)];
      sum += (double)(a * X2[vox(bh, x2, y2, z2)]);
      cs += (double)a;
    }
  } else { /* neighbour coarser */
    int bl = bsx / 2, x2, y2, z2, fx = vx, fy = vy, fz = vz;
    if (d == 0) { x2 = s ? 0 : bl - 1; y2 = vy / 2; z2 = vz / 2; if (s) fx = 0; }
    else if (d == 1) { y2 = s ? 0 : bl - 1; x2 = vx / 2; z2 = vz / 2; if (s) fy = 0; }
    else { z2 = s ? 0 : bl - 1; x2 = vx / 2; y2 = vy / 2; if (s) fz = 0; }
    const float *fa = ps->pfa[L][d] + (int64_t)(s ? b2 : b) * bsx * bsx * bsx;
    float a = a_fc();
    if (cuts) a = a * fa[vox(bsx, fx, fy, fz)];
    sum += (double)(a * X2[vox(bl, x2, y2, z2)]);
    cs += (double)a;
  }
  *csum = cs;
  return sum;
}

/* face weight W the relax applies for face (d, s) of cell (x,y,z) of a
 * block, and the kind of the face: 0 same level, 1 outside the domain,
 * 2 mixed resolution (msbg3.cpp:606-633, 743-1030) */
static float face_w(const or_ps *ps, int m, int b, int d, int s, int x, int y, int z, int *kind) {
  const mgl_t *g = &ps->mg[m];
  int L = g->L[b], bsx = ps->bsx0 >> L, c[3] = {x, y, z};
  const float *fa = ps->pfa[L][d] + (int64_t)b * bsx * bsx * bsx;
  *kind = 0;
  if (!s && c[d] > 0) return fa[vox(bsx, x, y, z)];
  if (s && c[d] < bsx - 1) {
    c[d]++;
    return fa[vox(bsx, c[0], c[1], c[2])];
  }
  int b2 = neigh_block(ps, b, d, s);
  if (b2 < 0) {
    *kind = 1;
    if (!s) return fa[vox(bsx, x, y, z)];
    return ps->dbc[2 * d + 1] == PS_DBC_OPEN ? 1.0f : 0.0f;
  }
  if (g->L[b2] != L) { *kind = 2; return 1.0f; }
  if (!s) return fa[vox(bsx, x, y, z)];
  c[d] = 0;
  return ps->pfa[L][d][(int64_t)b2 * bsx * bsx * bsx + vox(bsx, c[0], c[1], c[2])];
}

static int build_mg(or_ps *ps) {
  for (int m = 0; m < ps->nmg; m++) {
    mgl_t *g = &ps->mg[m];
    if (m < ps->nlev) {
      g->dense = 0;
      g->off = malloc(sizeof(int64_t) * ps->nb);
      g->L = malloc(ps->nb);
      int64_t n = 0;
      for (int b = 0; b < ps->nb; b++) {
        int L = ps->level[b] > m ? ps->level[b] : m, bsx = ps->bsx0 >> L;
        g->L[b] = (int8_t)L;
        g->off[b] = n;
        n += (int64_t)bsx * bsx * bsx;
      }
      g->n = n;
    } else {
      g->dense = 1;
      g->sx = ps->gsx[m]; g->sy = ps->gsy[m]; g->sz = ps->gsz[m];
      g->n = (int64_t)(g->sx + 2) * (g->sy + 2) * (g->sz + 2);
    }
    for (int c = 0; c < 4; c++) g->ch[c] = calloc((size_t)g->n, sizeof(float));
    g->D = calloc((size_t)g->n, sizeof(float));
    g->flags = calloc((size_t)g->n, sizeof(uint16_t));
  }
  /* block levels: flags, block flags, diagonal */
  for (int m = 0; m < ps->nlev; m++) {
    mgl_t *g = &ps->mg[m];
    g->bflags = calloc(ps->nb, sizeof(uint16_t));
    for (int b = 0; b < ps->nb; b++) {
      int L = g->L[b], bsx = ps->bsx0 >> L, bx, by, bz, fluid = 0, cuts = 0;
      block_coords(ps, b, &bx, &by, &bz);
      uint16_t *fl = g->flags + g->off[b];
      for (int z = 0; z < bsx; z++)
        for (int y = 0; y < bsx; y++)
          for (int x = 0; x < bsx; x++) {
            uint16_t f = ps->gflags[L][gidx(ps, L, bx * bsx + x, by * bsx + y, bz * bsx + z)];
            fl[vox(bsx, x, y, z)] = f;
            fluid |= (f & (PS_CELL_SOLID | PS_CELL_VOID)) == 0;
            cuts |= (f & (PS_CELL_SOLID | PS_CELL_PARTIAL_SOLID)) != 0;
          }
      g->bflags[b] = (uint16_t)((fluid ? 0 : PS_BLK_NO_FLUID) | (cuts ? PS_BLK_CUTS_SOLID : 0));
    }
    for (int b = 0; b < ps->nb; b++) {
      int L = g->L[b], bsx = ps->bsx0 >> L, sw = simdw(bsx);
      const uint16_t *fl = g->flags + g->off[b];
      uint8_t gates[4096];
      vec_gates(fl, bsx, gates);
      int cuts = (g->bflags[b] & PS_BLK_CUTS_SOLID) != 0;
      for (int z = 0; z < bsx; z++)
        for (int y = 0; y < bsx; y++)
          for (int x = 0; x < bsx; x++) {
            int v = vox(bsx, x, y, z);
            if (fl[v] & (PS_CELL_SOLID | PS_CELL_VOID)) continue;
            int gated = cuts && (gates[v / sw] & 1);
            double sum = 0;
            for (int d = 0; d < 3; d++)
              for (int s = 0; s < 2; s++) {
                int kind;
                float w = face_w(ps, m, b, d, s, x, y, z, &kind);
                int c[3] = {x, y, z};
                int border = s ? c[d] == bsx - 1 : c[d] == 0;
                if (border && kind == 2) {
                  double cs;
                  mixed_sum(ps, m, PS_CH_X, b, neigh_block(ps, b, d, s), d, s, x, y, z, &cs);
                  sum += cs;
                } else {
                  sum += gated ? (double)w : 1.0;
                }
              }
            g->D[g->off[b] + v] = sum > 0 ? (float)(1.0 / sum) : 0.0f;
          }
    }
  }
  /* dense levels */
  for (int m = ps->nlev; m < ps->nmg; m++) {
    mgl_t *g = &ps->mg[m];
    int sx = g->sx, sy = g->sy, sz = g->sz, px = sx + 2, py = sy + 2;
    for (int d = 0; d < 3; d++) {
      g->fa[d] = calloc((size_t)g->n, sizeof(float));
      int ex = sx + (d == 0), ey = sy + (d == 1), ez = sz + (d == 2);
      for (int z = 0; z < ez; z++)
        for (int y = 0; y < ey; y++)
          for (int x = 0; x < ex; x++)
            g->fa[d][(x + 1) + (int64_t)px * ((y + 1) + (int64_t)py * (z + 1))] =
                ps->gfa[m][d][x + (int64_t)ex * (y + (int64_t)ey * z)];
    }
    for (int64_t i = 0; i < g->n; i++) g->flags[i] = PS_CELL_SOLID; /* ghost shell */
    for (int z = 0; z < sz; z++)
      for (int y = 0; y < sy; y++)
        for (int x = 0; x < sx; x++)
          g->flags[(x + 1) + (int64_t)px * ((y + 1) + (int64_t)py * (z + 1))] =
              ps->gflags[m][gidx(ps, m, x, y, z)];
    for (int z = 0; z < sz; z++)
      for (int y = 0; y < sy; y++)
        for (int x = 0; x < sx; x++) {
          int64_t p = (x + 1) + (int64_t)px * ((y + 1) + (int64_t)py * (z + 1));
          if (g->flags[p] & (PS_CELL_SOLID | PS_CELL_VOID)) continue;
          int gated = 0;
          for (int k = x & ~3; k < (x & ~3) + 4 && k < sx; k++)
            gated |= (g->flags[p - x + k] & (PS_CELL_SOLID | PS_CELL_PARTIAL_SOLID)) != 0;
          int64_t st[3] = {1, px, (int64_t)px * py};
          double sum = 0;
          for (int d = 0; d < 3; d++) {
            sum += gated ? (double)g->fa[d][p] : 1.0;
            sum += gated ? (double)g->fa[d][p + st[d]] : 1.0;
          }
          g->D[p] = sum > 0 ? (float)(1.0 / sum) : 0.0f;
        }
  }
  return 0;
}

or_ps *or_ps_create(int nbx, int nby, int nbz, int bsx0, double dx0, const int *dbc,
                    float omega, int block_iter, int ndense, const int8_t *level,
                    const uint8_t *cell_type, int *err) {
  *err = 1;
  if (nbx < 1 || nby < 1 || nbz < 1 || bsx0 < 4 || bsx0 > 16 || (bsx0 & (bsx0 - 1)) ||
      block_iter < 1 || ndense < 0)
    return NULL;
  int nlev = ilog2(bsx0); /* bsx0 ... 2 */
  if (nlev + ndense > MAXMG) return NULL;
  for (int k = 0; k < ndense; k++) {
    int f = 1 << (k + 1); /* dense level k has nb*bsx_min/2 >> k cells per axis */
    if (nbx % (f / 2) || nby % (f / 2) || nbz % (f / 2)) return NULL;
    if (nbx < f / 2 || nby < f / 2 || nbz < f / 2) return NULL;
  }
  or_ps *ps = calloc(1, sizeof(or_ps));
  ps->nbx = nbx; ps->nby = nby; ps->nbz = nbz; ps->nb = nbx * nby * nbz;
  ps->bsx0 = bsx0; ps->nlev = nlev; ps->ndense = ndense; ps->nmg = nlev + ndense;
  ps->dx0 = dx0; ps->omega = omega; ps->biter = block_iter;
  memcpy(ps->dbc, dbc, sizeof(ps->dbc));
  ps->level = malloc(ps->nb);
  memcpy(ps->level, level, ps->nb);
  for (int b = 0; b < ps->nb; b++) {
    if (level[b] < 0 || level[b] >= nlev) { or_ps_destroy(ps); return NULL; }
    for (int d = 0; d < 3; d++)
      for (int s = 0; s < 2; s++) {
        int b2 = neigh_block(ps, b, d, s);
        if (b2 >= 0 && abs(level[b] - level[b2]) > 1) { or_ps_destroy(ps); return NULL; }
      }
  }
  for (int l = 0; l < ps->nmg; l++) {
    ps->gsx[l] = (nbx * bsx0) >> l;
    ps->gsy[l] = (nby * bsx0) >> l;
    ps->gsz[l] = (nbz * bsx0) >> l;
  }
  size_t n0 = (size_t)ps->gsx[0] * ps->gsy[0] * ps->gsz[0];
  ps->type[0] = malloc(n0);
  memcpy(ps->type[0], cell_type, n0);
  for (size_t i = 0; i < n0; i++)
    if (cell_type[i] > PS_VOID) { or_ps_destroy(ps); return NULL; }
  build_global(ps);
  /* page-form face areas of every block at every block level */
  for (int l = 0; l < nlev; l++) {
    int bsx = bsx0 >> l;
    for (int d = 0; d < 3; d++) {
      float *p = ps->pfa[l][d] = malloc(sizeof(float) * (size_t)ps->nb * bsx * bsx * bsx);
      for (int b = 0; b < ps->nb; b++) {
        int bx, by, bz;
        block_coords(ps, b, &bx, &by, &bz);
        for (int z = 0; z < bsx; z++)
          for (int y = 0; y < bsx; y++)
            for (int x = 0; x < bsx; x++)
              p[(int64_t)b * bsx * bsx * bsx + vox(bsx, x, y, z)] =
                  ps->gfa[l][d][fa_idx(ps, l, d, bx * bsx + x, by * bsx + y, bz * bsx + z)];
      }
    }
  }
  build_mg(ps);
  *err = 0;
  return ps;
}

void or_ps_destroy(or_ps *ps) {
  if (!ps) return;
  for (int l = 0; l < MAXMG; l++) {
    free(ps->type[l]);
    free(ps->gflags[l]);
    for (int d = 0; d < 3; d++) free(ps->gfa[l][d]);
  }
  for (int l = 0; l < MAXL; l++)
    for (int d = 0; d < 3; d++) free(ps->pfa[l][d]);
  for (int m = 0; m < MAXMG; m++) {
    mgl_t *g = &ps->mg[m];
    free(g->off); free(g->L); free(g->D); free(g->flags); free(g->bflags);
    for (int c = 0; c < 4; c++) free(g->ch[c]);
    for (int d = 0; d < 3; d++) free(g->fa[d]);
  }
  free(ps->level);
  free(ps);
}

int or_ps_num_levels(const or_ps *ps, int *nmg, int *nlev) {
  *nmg = ps->nmg;
  *nlev = ps->nlev;
  return 0;
}

int64_t or_ps_level_size(const or_ps *ps, int m) {
  return (m < 0 || m >= ps->nmg) ? -1 : ps->mg[m].n;
}

int or_ps_get(const or_ps *ps, int chan, int m, float *out) {
  if (m < 0 || m >= ps->nmg || chan < 0 || chan > PS_CH_D) return 1;
  const float *src = chan == PS_CH_D ? ps->mg[m].D : ps->mg[m].ch[chan];
  memcpy(out, src, sizeof(float) * (size_t)ps->mg[m].n);
  return 0;
}

int or_ps_set(or_ps *ps, int chan, int m, const float *in) {
  if (m < 0 || m >= ps->nmg || chan < 0 || chan > PS_CH_T) return 1;
  memcpy(ps->mg[m].ch[chan], in, sizeof(float) * (size_t)ps->mg[m].n);
  return 0;
}

int or_ps_get_flags(const or_ps *ps, int m, uint16_t *out) {
  if (m < 0 || m >= ps->nmg) return 1;
  memcpy(out, ps->mg[m].flags, sizeof(uint16_t) * (size_t)ps->mg[m].n);
  return 0;
}

int or_ps_get_block_flags(const or_ps *ps, int m, uint16_t *out) {
  if (m < 0 || m >= ps->nlev) return 1;
  memcpy(out, ps->mg[m].bflags, sizeof(uint16_t) * (size_t)ps->nb);
  return 0;
}

int or_ps_fill(or_ps *ps, int chan, int m, float v) {
  if (m < 0 || m >= ps->nmg || chan < 0 || chan > PS_CH_T) return 1;
  for (int64_t i = 0; i < ps->mg[m].n; i++) ps->mg[m].ch[chan][i] = v;
  return 0;
}

/* ------------------------------------------------------- block operator */

/* One block: halo assembly (msbg3.cpp:530-1030) then the relax
 * sub-iterations or the single multiply pass (msbg3.cpp:1100-1470). */
static void block_op(or_ps *ps, int m, int b, int relax, unsigned options, int cx, int cb,
                     int cdst, double *dot) {
  mgl_t *g = &ps->mg[m];
  int L = g->L[b], bsx = ps->bsx0 >> L, sw = simdw(bsx), nv = bsx * bsx * bsx / sw;
  int hs = bsx + 2;
  float h = (float)((1 << L) * ps->dx0);
  float HR = 1.f / h;
  int cuts = (g->bflags[b] & PS_BLK_CUTS_SOLID) != 0;
  const uint16_t *fl = g->flags + g->off[b];
  const float *X = g->ch[cx] + g->off[b], *B = g->ch[cb] + g->off[b], *D = g->D + g->off[b];
  float *Y = g->ch[cdst] + g->off[b];
  static float buf[18 * 18 * 18];
  uint8_t gates[4096];
  memset(buf, 0, sizeof(buf));
  vec_gates(fl, bsx, gates);
#define HB(x, y, z) buf[((x) + 1) + hs * (((y) + 1) + hs * ((z) + 1))]
  for (int z = 0; z < bsx; z++)
    for (int y = 0; y < bsx; y++)
      for (int x = 0; x < bsx; x++) {
        int v = vox(bsx, x, y, z);
        HB(x, y, z) = (gates[v / sw] & 2) ? 0.0f : X[v] * h;
      }
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      int b2 = neigh_block(ps, b, d, s);
      for (int q = 0; q < bsx; q++)
        for (int p = 0; p < bsx; p++) {
          int c[3];
          c[d] = s ? bsx - 1 : 0;  /* tangential axes in (lower, upper) order */
          if (d == 0) { c[1] = p; c[2] = q; }
          else if (d == 1) { c[0] = p; c[2] = q; }
          else { c[0] = p; c[1] = q; }
          int o[3] = {c[0], c[1], c[2]};
          o[d] += s ? 1 : -1;
          float val;
          if (b2 < 0) {
            val = 0.0f;
          } else if (g->L[b2] == L) {
            int n[3] = {c[0], c[1], c[2]};
            n[d] = s ? 0 : bsx - 1;
            val = g->ch[cx][g->off[b2] + vox(bsx, n[0], n[1], n[2])] * h;
          } else {
            double cs;
            double sum = mixed_sum(ps, m, cx, b, b2, d, s, c[0], c[1], c[2], &cs);
            val = (float)(sum * (double)h);
          }
          HB(o[0], o[1], o[2]) = val;
        }
    }
  int its = relax ? ps->biter : 1;
  float omega = ps->omega, om1 = 1.0f - omega;
  int rev_opt = (options & PS_OPT_REVERSE_ORDER) != 0;
  for (int it = 0; it < its; it++) {
    int rev = relax && (rev_opt == (it & 1));
    int last = it == its - 1;
    for (int k = 0; k < nv; k++) {
      int kk = rev ? nv - 1 - k : k;   /* vectors in memory order, reversed or not */
      if (gates[kk] & 2) continue;
      int x0 = (kk * sw) % bsx, y = (kk * sw / bsx) % bsx, z = kk * sw / (bsx * bsx);
      float Yv[8];
      for (int l = 0; l < sw; l++) {
        int x = x0 + l, v = vox(bsx, x, y, z);
        float F0 = HB(x, y, z), F1 = HB(x - 1, y, z), F2 = HB(x + 1, y, z),
              F3 = HB(x, y - 1, z), F4 = HB(x, y + 1, z), F5 = HB(x, y, z - 1),
              F6 = HB(x, y, z + 1);
        float S;
        if (cuts && (gates[kk] & 1)) {
          float W[6];
          int kind;  /* a mixed-resolution face has W = 1, its value is premultiplied */
          for (int f = 0; f < 6; f++) W[f] = face_w(ps, m, b, f / 2, f & 1, x, y, z, &kind);
          S = 0.f;
          S = fmaf(F1, W[0], S);
          S = fmaf(F2, W[1], S);
          S = fmaf(F3, W[2], S);
          S = fmaf(F4, W[3], S);
          S = fmaf(F5, W[4], S);
          S = fmaf(F6, W[5], S);
        } else {
          S = F1 + F2 + F3 + F4 + F5 + F6;
        }
        if (!relax) {
          float r = D[v] != 0.0f ? (1.f / D[v]) * F0 - S : 0.0f;
          if (options & PS_OPT_CALC_RESIDUAL) r = B[v] - r;
          else if (dot) *dot += (double)r * (double)(F0 * HR);
          Yv[l] = r;
        } else {
          Yv[l] = fmaf(om1, F0, (omega * (B[v] + S)) * D[v]);
        }
      }
      for (int l = 0; l < sw; l++) {
        int x = x0 + l, v = vox(bsx, x, y, z);
        if (!relax) Y[v] = Yv[l];
        else if (!last) HB(x, y, z) = Yv[l];
        else Y[v] = Yv[l] * HR;
      }
    }
  }
#undef HB
}

/* ------------------------------------------------------- dense operator */

static void dense_op(or_ps *ps, int m, int jacobi, unsigned options, const float *X,
                     const float *B, float *Y) {
  mgl_t *g = &ps->mg[m];
  int sx = g->sx, sy = g->sy, sz = g->sz, px = sx + 2, py = sy + 2;
  int64_t ys = px, zs = (int64_t)px * py;
  float h = (float)((1 << m) * ps->dx0);
  float omega = ps->omega, HROmega = omega * 1.0f / h, om1 = 1.0f - omega;
  for (int z = 0; z < sz; z++)
    for (int y = 0; y < sy; y++)
      for (int x0 = 0; x0 < sx; x0 += 4) {
        int64_t p0 = (x0 + 1) + ys * (y + 1) + zs * (z + 1);
        int n = sx - x0 < 4 ? sx - x0 : 4, anyfluid = 0, gated = 0;
        for (int l = 0; l < n; l++) {
          uint16_t f = g->flags[p0 + l];
          anyfluid |= (f & (PS_CELL_SOLID | PS_CELL_VOID)) == 0;
          gated |= (f & (PS_CELL_SOLID | PS_CELL_PARTIAL_SOLID)) != 0;
        }
        if (!anyfluid) continue;
        for (int l = 0; l < n; l++) {
          int64_t p = p0 + l;
          float F0 = X[p], F1 = X[p - 1], F2 = X[p + 1], F3 = X[p - ys], F4 = X[p + ys],
                F5 = X[p - zs], F6 = X[p + zs];
          if (gated) {
            F1 = F1 * g->fa[0][p];
            F2 = F2 * g->fa[0][p + 1];
            F3 = F3 * g->fa[1][p];
            F4 = F4 * g->fa[1][p + ys];
            F5 = F5 * g->fa[2][p];
            F6 = F6 * g->fa[2][p + zs];
          }
          float S = (F1 + F2 + F3 + F4 + F5 + F6) * h, D = g->D[p], r;
          int fluid = (g->flags[p] & (PS_CELL_SOLID | PS_CELL_VOID)) == 0;
          if (jacobi) {
            r = fmaf(om1, F0, ((B[p] + S) * D) * HROmega);
            if (!fluid) r = 0.0f;
          } else {
            float N = 1.0f / D;
            r = (h * N) * F0 - S;
            if (!fluid) r = 0.0f;
            if (options & PS_OPT_CALC_RESIDUAL) r = B[p] - r;
          }
          Y[p] = r;
        }
      }
}

/* ------------------------------------------------------------ drivers */

int or_ps_relax(or_ps *ps, int reverse, int m, int cx, int cb, int ctmp, int n_iter) {
  if (m < 0 || m >= ps->nmg || cx < 0 || cx > PS_CH_T || cb < 0 || cb > PS_CH_T) return 1;
  if (n_iter <= 0) return 0;
  mgl_t *g = &ps->mg[m];
  if (g->dense) { /* msbg3.cpp:1917-2165: Jacobi, ping-pong through chanTmp */
    if (ctmp < 0 || ctmp > PS_CH_T || ctmp == cx || ctmp == cb) return 1;
    int rd = ctmp, wr = cx;
    for (int it = 0; it < n_iter; it++) {
      int t = rd; rd = wr; wr = t;
      dense_op(ps, m, 1, 0, g->ch[rd], g->ch[cb], g->ch[wr]);
    }
    if (wr != cx) memcpy(g->ch[cx], g->ch[wr], sizeof(float) * (size_t)g->n);
    return 0;
  }
  unsigned opt = reverse ? PS_OPT_REVERSE_ORDER : 0;
  for (int it = 0; it < n_iter; it++)
    for (int pass = 0; pass < 2; pass++)
      for (int b = 0; b < ps->nb; b++) {
        if (g->bflags[b] & PS_BLK_NO_FLUID) continue;
        int bx, by, bz;
        block_coords(ps, b, &bx, &by, &bz);
        int skip = ((bx + by + bz) & 1) == pass;
        if (reverse) skip = !skip;
        if (skip) continue;
        block_op(ps, m, b, 1, opt, cx, cb, cx, NULL);
      }
  return 0;
}

int or_ps_multiply(or_ps *ps, unsigned options, int m, int cx, int cb, int cdst, double *dot) {
  if (m < 0 || m >= ps->nmg || cx < 0 || cx > PS_CH_T || cb < 0 || cb > PS_CH_T ||
      cdst < 0 || cdst > PS_CH_T)
    return 1;
  mgl_t *g = &ps->mg[m];
  if (dot) *dot = 0;
  if (g->dense) {
    dense_op(ps, m, 0, options, g->ch[cx], g->ch[cb], g->ch[cdst]);
    return 0;
  }
  for (int b = 0; b < ps->nb; b++) {
    if (g->bflags[b] & PS_BLK_NO_FLUID) continue;
    block_op(ps, m, b, 0, options, cx, cb, cdst, dot);
  }
  return 0;
}

/* The reference names constant restriction and prolongation
 * (MG_CONST_RESTRICTION, MG_CONST_PROLONGATION with alpha 1, msbgcell.h:99-101)
 * but ships neither. The library's definition: a coarse cell receives the
 * float sum of its eight children, children in x-fastest order; a fine fluid
 * cell adds its parent's value. Cells that do not coarsen at m+1 (blocks
 * already coarser than m) are copied or added one to one. Non-fluid target
 * cells of the restriction are set to zero; non-fluid fine cells receive no
 * correction. */
static int fluid_flag(uint16_t f) { return (f & (PS_CELL_SOLID | PS_CELL_VOID)) == 0; }

/* global cell (x,y,z) of block level l -> pointer offset in mg level m */
static int64_t block_cell(const or_ps *ps, int m, int x, int y, int z) {
  const mgl_t *g = &ps->mg[m];
  int bsxL = ps->bsx0 >> m;  /* all blocks are at level m when this is used */
  int b = (x / bsxL) + ps->nbx * ((y / bsxL) + ps->nby * (z / bsxL));
  return g->off[b] + vox(bsxL, x % bsxL, y % bsxL, z % bsxL);
}

int or_ps_restrict(or_ps *ps, int m, int cs, int cd) {
  if (m < 0 || m + 1 >= ps->nmg || cs < 0 || cs > PS_CH_T || cd < 0 || cd > PS_CH_T) return 1;
  mgl_t *f = &ps->mg[m], *c = &ps->mg[m + 1];
  if (!c->dense) {
    for (int b = 0; b < ps->nb; b++) {
      int Lf = f->L[b], Lc = c->L[b], bf = ps->bsx0 >> Lf, bc = ps->bsx0 >> Lc;
      const float *src = f->ch[cs] + f->off[b];
      float *dst = c->ch[cd] + c->off[b];
      const uint16_t *cf = c->flags + c->off[b];
      for (int z = 0; z < bc; z++)
        for (int y = 0; y < bc; y++)
          for (int x = 0; x < bc; x++) {
            float s;
            if (Lc == Lf) {
              s = src[vox(bf, x, y, z)];
            } else {
              s = 0.0f;
              for (int k = 0; k < 8; k++)
                s = s + src[vox(bf, 2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + (k >> 2))];
            }
            dst[vox(bc, x, y, z)] = fluid_flag(cf[vox(bc, x, y, z)]) ? s : 0.0f;
          }
    }
    return 0;
  }
  int sx = c->sx, sy = c->sy, sz = c->sz, px = sx + 2, py = sy + 2;
  for (int z = 0; z < sz; z++)
    for (int y = 0; y < sy; y++)
      for (int x = 0; x < sx; x++) {
        float s = 0.0f;
        for (int k = 0; k < 8; k++) {
          int X = 2 * x + (k & 1), Y = 2 * y + ((k >> 1) & 1), Z = 2 * z + (k >> 2);
          int64_t p = f->dense ? (X + 1) + (int64_t)(f->sx + 2) * ((Y + 1) + (int64_t)(f->sy + 2) * (Z + 1))
                               : block_cell(ps, m, X, Y, Z);
          s = s + f->ch[cs][p];
        }
        int64_t q = (x + 1) + (int64_t)px * ((y + 1) + (int64_t)py * (z + 1));
        c->ch[cd][q] = fluid_flag(c->flags[q]) ? s : 0.0f;
      }
  return 0;
}

int or_ps_prolongate(or_ps *ps, int m, int cs, int cd) {
  if (m < 0 || m + 1 >= ps->nmg || cs < 0 || cs > PS_CH_T || cd < 0 || cd > PS_CH_T) return 1;
  mgl_t *f = &ps->mg[m], *c = &ps->mg[m + 1];
  if (!f->dense) {
    if (!c->dense) {
      for (int b = 0; b < ps->nb; b++) {
        int Lf = f->L[b], Lc = c->L[b], bf = ps->bsx0 >> Lf, bc = ps->bsx0 >> Lc;
        const float *src = c->ch[cs] + c->off[b];
        float *dst = f->ch[cd] + f->off[b];
        const uint16_t *ff = f->flags + f->off[b];
        for (int z = 0; z < bf; z++)
          for (int y = 0; y < bf; y++)
            for (int x = 0; x < bf; x++) {
              int v = vox(bf, x, y, z);
              if (!fluid_flag(ff[v])) continue;
              float e = Lc == Lf ? src[v] : src[vox(bc, x / 2, y / 2, z / 2)];
              dst[v] = dst[v] + e;
            }
      }
      return 0;
    }
    int bsxL = ps->bsx0 >> m, sx = bsxL * ps->nbx, sy = bsxL * ps->nby, sz = bsxL * ps->nbz;
    for (int z = 0; z < sz; z++)
      for (int y = 0; y < sy; y++)
        for (int x = 0; x < sx; x++) {
          int64_t p = block_cell(ps, m, x, y, z);
          if (!fluid_flag(f->flags[p])) continue;
          int64_t q = (x / 2 + 1) + (int64_t)(c->sx + 2) * ((y / 2 + 1) + (int64_t)(c->sy + 2) * (z / 2 + 1));
          f->ch[cd][p] = f->ch[cd][p] + c->ch[cs][q];
        }
    return 0;
  }
  for (int z = 0; z < f->sz; z++)
    for (int y = 0; y < f->sy; y++)
      for (int x = 0; x < f->sx; x++) {
        int64_t p = (x + 1) + (int64_t)(f->sx + 2) * ((y + 1) + (int64_t)(f->sy + 2) * (z + 1));
        if (!fluid_flag(f->flags[p])) continue;
        int64_t q = (x / 2 + 1) + (int64_t)(c->sx + 2) * ((y / 2 + 1) + (int64_t)(c->sy + 2) * (z / 2 + 1));
        f->ch[cd][p] = f->ch[cd][p] + c->ch[cs][q];
      }
  return 0;
}

/* The library's V-cycle (the reference's multigrid driver is not shipped):
 * pre-relax nu1 (forward order), residual, restrict, recurse with zero
 * initial correction, prolongate, post-relax nu2 (reverse order); the
 * coarsest level is relaxed coarse_iter times. */
static void vcycle_level(or_ps *ps, int m, int nu1, int nu2, int coarse_iter) {
  if (m == ps->nmg - 1) {
    or_ps_relax(ps, 0, m, PS_CH_X, PS_CH_B, PS_CH_T, coarse_iter);
    return;
  }
  or_ps_relax(ps, 0, m, PS_CH_X, PS_CH_B, PS_CH_T, nu1);
  or_ps_fill(ps, PS_CH_R, m, 0.0f);
  or_ps_multiply(ps, PS_OPT_CALC_RESIDUAL, m, PS_CH_X, PS_CH_B, PS_CH_R, NULL);
  or_ps_restrict(ps, m, PS_CH_R, PS_CH_B);
  or_ps_fill(ps, PS_CH_X, m + 1, 0.0f);
  vcycle_level(ps, m + 1, nu1, nu2, coarse_iter);
  or_ps_prolongate(ps, m, PS_CH_X, PS_CH_X);
  or_ps_relax(ps, 1, m, PS_CH_X, PS_CH_B, PS_CH_T, nu2);
}

int or_ps_vcycle(or_ps *ps, int nu1, int nu2, int coarse_iter) {
  if (nu1 < 0 || nu2 < 0 || coarse_iter < 0) return 1;
  vcycle_level(ps, 0, nu1, nu2, coarse_iter);
  return 0;
}
