// Launchers of the warm-start network's kernels (k_cnn.hip), used by cnn.cpp.
//
// Activations are BF16 (raw uint16 bits), channels-last: [n][z][y][x][c].
// Convolutions are 3x3x3 cross-correlations with zero padding `pad`:
// dout = din + 2*pad - 2. Weights are kept in FP32 as w[cout][cin][27]
// (tap = (dz*3+dy)*3+dx) and swizzled into MFMA A-operand fragments.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace msbg {
namespace cnn {

// uint4 fragments of a conv's swizzled weights: nch * ksteps * (cout/16) * 64
int64_t swizzled_size(int cin, int cout);
// forward weights (cin_true <= cin pads input channels with zeros), or the
// weights of the backward-data conv (flipped taps, cin and cout swapped)
int launch_swizzle(hipStream_t s, const float *w, int cin_true, int cin, int cout, bool flip,
                   uint4 *out);
// y = relu?(conv(x) + bias); bias may be null. cin 1 (the first layer) has
// its own kernel; other cin are multiples of 8.
int launch_conv(hipStream_t s, const uint16_t *x, int n, int cin, int cout, int dout, int pad,
                const uint4 *wsw, const float *bias, bool relu, uint16_t *y);
// the decoder's conv over [upsample(a), b] without materialising the concat
int launch_conv_cat(hipStream_t s, const uint16_t *a, int ca, const uint16_t *b, int cb, int n,
                    int cout, int dout, const uint4 *wsw, const float *bias, bool relu,
                    uint16_t *y);
// dw partials for the conv x(din=dout+2-2pad) -> dz(dout): x_tp is x transposed and
// zero padded to [cin16][dout+2]^3[n], dz_t is [cout][dout^3][n]; n % 8 == 0
int wgrad_chunks(int cin, int cout, int dout, int n);
int launch_wgrad(hipStream_t s, const uint16_t *dz_t, const uint16_t *x_tp, int n, int cin16,
                 int cout, int dout, float *partial);
// dw[co][ci][27] = sum of partials (ci < cin_true)
int launch_wgrad_reduce(hipStream_t s, const float *partial, int chunks, int cin16, int cin_true,
                        int cout, float *dw);
// x [n][din^3][c] -> x_tp [cpad][(din+2pad)^3][n], zero outside / beyond c
int launch_xpose_pad(hipStream_t s, const uint16_t *x, int n, int din, int c, int pad, int cpad,
                     uint16_t *x_tp);
// dy := dy * (y > 0) when relu, in place; dz_t [c][v][n] = the result transposed
int launch_relu_bwd_xpose(hipStream_t s, uint16_t *dy, const uint16_t *y, int n, int64_t v,
                          int c, bool relu, uint16_t *dz_t);
// out[c] = sum of row c of a [rows][len] BF16 matrix (FP32, fixed order);
// scratch holds rowsum_scratch(rows) floats (rows <= 128)
int rowsum_scratch(int rows);
int launch_rowsum(hipStream_t s, const uint16_t *a, int rows, int64_t len, float *scratch,
                  float *out);
int launch_pool2(hipStream_t s, const uint16_t *x, int n, int d, int c, uint16_t *y);
int launch_pool2_bwd(hipStream_t s, const uint16_t *dy, int n, int d, int c, bool accumulate,
                     uint16_t *dx);
int launch_upcat(hipStream_t s, const uint16_t *a, int ca, const uint16_t *b, int cb, int n, int d,
                 uint16_t *out);
int launch_upcat_bwd(hipStream_t s, const uint16_t *dout, int ca, int cb, int n, int d,
                     uint16_t *da, uint16_t *db);
// y[n][32^3] = hb + sum_c hw[c] * d[n][v][c]; head = {hw[16], hb}
int launch_head(hipStream_t s, const uint16_t *d, int n, const float *head, float *y);
// the last 16 -> 16 convolution at 32^3 with the head applied in its store:
// yh as launch_head on its output, bit for bit; y (the activation) may be
// null when no backward pass needs it
int launch_conv_head(hipStream_t s, const uint16_t *x, int n, int cin, int cout, int dout,
                     int pad, const uint4 *wsw, const float *bias, bool relu, uint16_t *y,
                     const float *head, float *yh);
// dd = dy * hw (BF16); partials [blocks][17] of (sum dy*d_c, sum dy)
int head_bwd_blocks(int n);
int launch_head_bwd(hipStream_t s, const float *dy, const uint16_t *d, int n, const float *head,
                    uint16_t *dd, float *partial);
int launch_sum_partials(hipStream_t s, const float *partial, int rows, int width, float *out);
// dy = 2 (y - t) / count; loss partials per block (FP64)
int mse_blocks(int64_t count);
int launch_mse(hipStream_t s, const float *y, const float *t, int64_t count, float *dy,
               double *partial);
int launch_sum_f64(hipStream_t s, const double *partial, int n, double *out);
// acc[0] += *sum * scale, acc[1] += 1
int launch_loss_acc(hipStream_t s, const double *sum, double scale, double *acc);
// dst[i] = src[rows[i]] for n rows of rowBytes (a multiple of 16); rows on the device
int launch_gather_rows(hipStream_t s, void *dst, const void *src, const int32_t *rows, int n,
                       int64_t rowBytes);
int launch_adam(hipStream_t s, float *p, const float *g, float *m, float *v, int64_t n, float lr,
                float b1, float b2, float eps, float bc1, float bc2);

}  // namespace cnn
}  // namespace msbg

namespace msbg {
namespace cnn {

// Network input windows of level-0 blocks: out[i][36^3] (BF16) holds block
// bids[i] with a 2-voxel halo under the sweep's halo rule (value blocks and
// const EMPTY/FULL neighbours are read, NULL neighbours and the outside of
// the domain repeat the block's own boundary voxels).
int launch_gather36(hipStream_t s, int storage, const int32_t *page_map, const uint8_t *pages,
                    double empty_raw, double full_raw, int nbx, int nby, int nbz,
                    const int32_t *bids, int n, uint16_t *out);
// y[i][32^3] = decode(page[bids[i]]) (FP32)
int launch_extract32(hipStream_t s, int storage, const int32_t *page_map, const uint8_t *pages,
                     const int32_t *bids, int n, float *y);
// page[bids[i]] := encode(clamp(decode(page) + dy[i], 0, 1)) for dy [n][32^3] (FP32)
int launch_scatter32(hipStream_t s, int storage, const int32_t *page_map, uint8_t *pages,
                     const int32_t *bids, int n, const float *y);

}  // namespace cnn
}  // namespace msbg
