/*
 * cnn_oracle.h -- FP64 CPU restatement of the warm-start network's layers.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU's BF16 MFMA kernels. The
 * product library never calls into it. The network has no counterpart in
 * the reference (a build-defined addition on top of the MCF sweep), so these
 * functions restate textbook definitions; tests pin them against torch's CPU
 * conv3d / avg_pool3d / interpolate and against finite differences.
 *
 * Tensors are channels-last: x[n][z][y][x][c], cube edge d.
 * Weights: w[cout][cin][27], tap = (dz*3 + dy)*3 + dx, cross-correlation:
 *   y[n][z][y][x][co] = b[co] + sum_{ci,tap} w[co][ci][tap] *
 *                       x[n][z+dz-pad][y+dy-pad][x+dx-pad][ci]   (0 outside)
 * with dout = din + 2*pad - 2.
 */
#ifndef MSBG_CNN_ORACLE_H
#define MSBG_CNN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* forward; relu != 0 applies max(0, .) after the bias */
void or_conv3d_fwd(const double *x, int n, int din, int cin, const double *w, const double *b,
                   int cout, int pad, int relu, double *y);
/* dx = d(sum dy*y)/dx of the pre-activation conv (no relu) */
void or_conv3d_bwd_data(const double *dy, int n, int dout, int cout, const double *w, int cin,
                        int pad, double *dx);
/* dw[cout][cin][27], db[cout] (db may be NULL) */
void or_conv3d_bwd_weights(const double *x, const double *dy, int n, int din, int cin, int cout,
                           int pad, double *dw, double *db);
/* 2x2x2 average pool, d even: y has edge d/2 */
void or_pool2_fwd(const double *x, int n, int d, int c, double *y);
/* dx (edge d) = dy (edge d/2) / 8 per child, added to dx when accumulate != 0 */
void or_pool2_bwd(const double *dy, int n, int d, int c, int accumulate, double *dx);
/* out[n][v][0:ca] = a[n][v/2][:] (nearest upsample), out[n][v][ca:ca+cb] = b[n][v][:] */
void or_upcat_fwd(const double *a, int ca, const double *b, int cb, int n, int d, double *out);
void or_upcat_bwd(const double *dout, int ca, int cb, int n, int d, double *da, double *db);

/* round each element to the nearest bfloat16 (ties to even), as double */
void or_round_bf16(double *x, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
