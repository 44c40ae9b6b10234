/*
 * cnn_oracle.c -- FP64 direct (non-GEMM) restatement of the warm-start
 * network's layers. Test infrastructure only; see cnn_oracle.h.
 */
#include "cnn_oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline int64_t vox(int d, int z, int y, int x) { return ((int64_t)z * d + y) * d + x; }

void or_conv3d_fwd(const double *x, int n, int din, int cin, const double *w, const double *b,
                   int cout, int pad, int relu, double *y) {
  const int dout = din + 2 * pad - 2;
  const int64_t vin = (int64_t)din * din * din, vo = (int64_t)dout * dout * dout;
#pragma omp parallel for collapse(2) schedule(static)
  for (int s = 0; s < n; s++)
    for (int z = 0; z < dout; z++)
      for (int yy = 0; yy < dout; yy++)
        for (int xx = 0; xx < dout; xx++) {
          double *o = y + ((int64_t)s * vo + vox(dout, z, yy, xx)) * cout;
          for (int co = 0; co < cout; co++) {
            double acc = b ? b[co] : 0.0;
            for (int tap = 0; tap < 27; tap++) {
              const int iz = z + tap / 9 - pad, iy = yy + (tap / 3) % 3 - pad,
                        ix = xx + tap % 3 - pad;
              if (iz < 0 || iy < 0 || ix < 0 || iz >= din || iy >= din || ix >= din) continue;
              const double *xi = x + ((int64_t)s * vin + vox(din, iz, iy, ix)) * cin;
              const double *wr = w + (int64_t)co * cin * 27 + tap;
              for (int ci = 0; ci < cin; ci++) acc += wr[(int64_t)ci * 27] * xi[ci];
            }
            o[co] = (relu && acc < 0.0) ? 0.0 : acc;
          }
        }
}

void or_conv3d_bwd_data(const double *dy, int n, int dout, int cout, const double *w, int cin,
                        int pad, double *dx) {
  /* dx[i] = sum over outputs o with i = o + d - pad of dy[o] * w[d] */
  const int din = dout - 2 * pad + 2;
  const int64_t vin = (int64_t)din * din * din, vo = (int64_t)dout * dout * dout;
#pragma omp parallel for collapse(2) schedule(static)
  for (int s = 0; s < n; s++)
    for (int z = 0; z < din; z++)
      for (int yy = 0; yy < din; yy++)
        for (int xx = 0; xx < din; xx++) {
          double *o = dx + ((int64_t)s * vin + vox(din, z, yy, xx)) * cin;
          for (int ci = 0; ci < cin; ci++) o[ci] = 0.0;
          for (int tap = 0; tap < 27; tap++) {
            const int oz = z - tap / 9 + pad, oy = yy - (tap / 3) % 3 + pad,
                      ox = xx - tap % 3 + pad;
            if (oz < 0 || oy < 0 || ox < 0 || oz >= dout || oy >= dout || ox >= dout) continue;
            const double *g = dy + ((int64_t)s * vo + vox(dout, oz, oy, ox)) * cout;
            for (int ci = 0; ci < cin; ci++) {
              double acc = 0.0;
              for (int co = 0; co < cout; co++) acc += g[co] * w[((int64_t)co * cin + ci) * 27 + tap];
              o[ci] += acc;
            }
          }
        }
}

void or_conv3d_bwd_weights(const double *x, const double *dy, int n, int din, int cin, int cout,
                           int pad, double *dw, double *db) {
  const int dout = din + 2 * pad - 2;
  const int64_t vin = (int64_t)din * din * din, vo = (int64_t)dout * dout * dout;
#pragma omp parallel for collapse(2) schedule(static)
  for (int co = 0; co < cout; co++)
    for (int ci = 0; ci < cin; ci++) {
      double acc[27];
      for (int t = 0; t < 27; t++) acc[t] = 0.0;
      for (int s = 0; s < n; s++)
        for (int z = 0; z < dout; z++)
          for (int yy = 0; yy < dout; yy++)
            for (int xx = 0; xx < dout; xx++) {
              const double g = dy[((int64_t)s * vo + vox(dout, z, yy, xx)) * cout + co];
              if (g == 0.0) continue;
              for (int tap = 0; tap < 27; tap++) {
                const int iz = z + tap / 9 - pad, iy = yy + (tap / 3) % 3 - pad,
                          ix = xx + tap % 3 - pad;
                if (iz < 0 || iy < 0 || ix < 0 || iz >= din || iy >= din || ix >= din) continue;
                acc[tap] += g * x[((int64_t)s * vin + vox(din, iz, iy, ix)) * cin + ci];
              }
            }
      for (int t = 0; t < 27; t++) dw[((int64_t)co * cin + ci) * 27 + t] = acc[t];
    }
  if (db)
    for (int co = 0; co < cout; co++) {
      double acc = 0.0;
      for (int64_t i = 0; i < (int64_t)n * vo; i++) acc += dy[i * cout + co];
      db[co] = acc;
    }
}

void or_pool2_fwd(const double *x, int n, int d, int c, double *y) {
  const int h = d / 2;
  for (int s = 0; s < n; s++)
    for (int z = 0; z < h; z++)
      for (int yy = 0; yy < h; yy++)
        for (int xx = 0; xx < h; xx++)
          for (int ch = 0; ch < c; ch++) {
            double acc = 0.0;
            for (int k = 0; k < 8; k++)
              acc += x[((int64_t)s * d * d * d +
                        vox(d, 2 * z + (k >> 2), 2 * yy + ((k >> 1) & 1), 2 * xx + (k & 1))) *
                           c +
                       ch];
            y[((int64_t)s * h * h * h + vox(h, z, yy, xx)) * c + ch] = acc * 0.125;
          }
}

void or_pool2_bwd(const double *dy, int n, int d, int c, int accumulate, double *dx) {
  const int h = d / 2;
  for (int s = 0; s < n; s++)
    for (int z = 0; z < d; z++)
      for (int yy = 0; yy < d; yy++)
        for (int xx = 0; xx < d; xx++)
          for (int ch = 0; ch < c; ch++) {
            const double g =
                dy[((int64_t)s * h * h * h + vox(h, z / 2, yy / 2, xx / 2)) * c + ch] * 0.125;
            double *o = dx + ((int64_t)s * d * d * d + vox(d, z, yy, xx)) * c + ch;
            *o = accumulate ? *o + g : g;
          }
}

void or_upcat_fwd(const double *a, int ca, const double *b, int cb, int n, int d, double *out) {
  const int h = d / 2, c = ca + cb;
  for (int s = 0; s < n; s++)
    for (int z = 0; z < d; z++)
      for (int yy = 0; yy < d; yy++)
        for (int xx = 0; xx < d; xx++) {
          double *o = out + ((int64_t)s * d * d * d + vox(d, z, yy, xx)) * c;
          const double *pa = a + ((int64_t)s * h * h * h + vox(h, z / 2, yy / 2, xx / 2)) * ca;
          const double *pb = b + ((int64_t)s * d * d * d + vox(d, z, yy, xx)) * cb;
          for (int k = 0; k < ca; k++) o[k] = pa[k];
          for (int k = 0; k < cb; k++) o[ca + k] = pb[k];
        }
}

void or_upcat_bwd(const double *dout, int ca, int cb, int n, int d, double *da, double *db) {
  const int h = d / 2, c = ca + cb;
  memset(da, 0, sizeof(double) * (size_t)n * h * h * h * ca);
  for (int s = 0; s < n; s++)
    for (int z = 0; z < d; z++)
      for (int yy = 0; yy < d; yy++)
        for (int xx = 0; xx < d; xx++) {
          const double *g = dout + ((int64_t)s * d * d * d + vox(d, z, yy, xx)) * c;
          double *pa = da + ((int64_t)s * h * h * h + vox(h, z / 2, yy / 2, xx / 2)) * ca;
          double *pb = db + ((int64_t)s * d * d * d + vox(d, z, yy, xx)) * cb;
          for (int k = 0; k < ca; k++) pa[k] += g[k];
          for (int k = 0; k < cb; k++) pb[k] = g[ca + k];
        }
}

void or_round_bf16(double *x, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    float f = (float)x[i];
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) <= 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    x[i] = (double)f;
  }
}
