// The warm-start network: a 3-stage 3D U-Net on 32^3 blocks with a 2-voxel
// halo (36^3 input), 16/32/64 channels, 3^3 convolutions, BF16 weights and
// activations with FP32 accumulation (k_cnn.hip), trained with Adam on the
// GPU. It predicts the change of a block over K plain MCF sweeps
// (msbg4.cpp:60-957); the warm start adds the change to the field at its own
// precision (the BF16 window only feeds the network) and the plain solver
// continues from there.
//
// Layers (input edge -> output edge, channels):
//   c0 36->34 1->16 (valid)   c1 34->32 16->16 (valid)      = e1
//   pool 32->16               c2 16 16->32, c3 16 32->32    = e2
//   pool 16->8                c4 8 32->64,  c5 8 64->64     = b
//   up+cat(b, e2) 16 96       c6 16 96->32, c7 16 32->32    = d2
//   up+cat(d2, e1) 32 48      c8 32 48->16, c9 32 16->16    = d1
//   head: y = hb + sum_c hw[c] d1[c]   (1x1x1, 16 -> 1), the predicted change
// ReLU follows every convolution. Parameters, flat FP32: for each conv
// w[cout][cin][27] then b[cout]; then hw[16], hb.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "cnn_kernels.h"
#include "internal.h"

using namespace msbg;
using namespace msbg::cnn;

namespace {

struct LayerSpec {
  int cin, cout, pad, din, dout;
};
constexpr LayerSpec kL[10] = {{1, 16, 0, 36, 34},  {16, 16, 0, 34, 32}, {16, 32, 1, 16, 16},
                              {32, 32, 1, 16, 16}, {32, 64, 1, 8, 8},   {64, 64, 1, 8, 8},
                              {96, 32, 1, 16, 16}, {32, 32, 1, 16, 16}, {48, 16, 1, 32, 32},
                              {16, 16, 1, 32, 32}};
constexpr int kHead = 17;
constexpr int64_t kVin = 36 * 36 * 36, kVout = 32 * 32 * 32;

// input channels as the conv kernel sees them, and as the weight-gradient
// kernel sees them (16-multiples: layer 0's single channel is padded)
inline int cin_k(int l) { return kL[l].cin; }
inline int cin16(int l) { return l == 0 ? 16 : kL[l].cin; }
inline int64_t vol(int d) { return (int64_t)d * d * d; }

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

struct msbg_unet {
  hipStream_t stream = nullptr;
  int64_t nparam = 0;
  int64_t offW[10], offB[10], offH;
  DevBuf<float> p, g, m, v;
  int64_t adam_t = 0;
  DevBuf<uint4> wf[10], wb[10];
  bool swz_valid = false;
  int cap = 0;
  // forward activations (channels-last BF16)
  DevBuf<uint16_t> x0, a[10], p1, p2, c7, c9;
  DevBuf<float> y, dy, tgt;
  // backward
  DevBuf<uint16_t> ga[10], gc9, gc7, gp2, gp1, dzt, xtp;
  DevBuf<float> wpart, hpart, hsum, rsum;
  DevBuf<double> lpart, lsum;
  DevBuf<double> lacc;     // running {loss sum, steps} of steps queued without a loss readback
  DevBuf<int32_t> rows;    // sample rows of the current *_rows call
};

namespace {

const uint16_t *conv_input(msbg_unet *u, int l) {
  switch (l) {
    case 0: return u->x0.p;
    case 2: return u->p1.p;
    case 4: return u->p2.p;
    case 6: return u->c7.p;
    case 8: return u->c9.p;
    default: return u->a[l - 1].p;
  }
}

int ensure_cap(msbg_unet *u, int n, bool train) {
  MSBG_CHECK(n >= 1 && n <= 4096, MSBG_EINVAL, "unet: batch %d outside [1, 4096]", n);
  if (n <= u->cap && (!train || u->ga[0].n >= (size_t)n * vol(34) * 16)) return 0;
  const int c = n > u->cap ? n : u->cap;
  MSBG_TRY(u->x0.ensure((size_t)c * kVin));
  for (int l = 0; l < 10; l++) MSBG_TRY(u->a[l].ensure((size_t)c * vol(kL[l].dout) * kL[l].cout));
  MSBG_TRY(u->p1.ensure((size_t)c * vol(16) * 16));
  MSBG_TRY(u->p2.ensure((size_t)c * vol(8) * 32));
  MSBG_TRY(u->c7.ensure((size_t)c * vol(16) * 96));
  MSBG_TRY(u->c9.ensure((size_t)c * vol(32) * 48));
  MSBG_TRY(u->y.ensure((size_t)c * kVout));
  if (train) {
    MSBG_TRY(u->dy.ensure((size_t)c * kVout));
    MSBG_TRY(u->tgt.ensure((size_t)c * kVout));
    for (int l = 0; l < 10; l++)
      MSBG_TRY(u->ga[l].ensure((size_t)c * vol(kL[l].dout) * kL[l].cout));
    MSBG_TRY(u->gc9.ensure((size_t)c * vol(32) * 48));
    MSBG_TRY(u->gc7.ensure((size_t)c * vol(16) * 96));
    MSBG_TRY(u->gp2.ensure((size_t)c * vol(8) * 32));
    MSBG_TRY(u->gp1.ensure((size_t)c * vol(16) * 16));
    size_t dz = 0, xt = 0, wp = 0;
    for (int l = 0; l < 10; l++) {
      dz = std::max(dz, (size_t)c * vol(kL[l].dout) * kL[l].cout);
      xt = std::max(xt, (size_t)c * vol(kL[l].dout + 2) * cin16(l));
      wp = std::max(wp, (size_t)wgrad_chunks(cin16(l), kL[l].cout, kL[l].dout, c) * 27 *
                            kL[l].cout * cin16(l));
    }
    MSBG_TRY(u->dzt.ensure(dz));
    MSBG_TRY(u->xtp.ensure(xt));
    MSBG_TRY(u->wpart.ensure(wp));
    MSBG_TRY(u->hpart.ensure((size_t)head_bwd_blocks(c) * kHead));
    MSBG_TRY(u->hsum.ensure(kHead));
    MSBG_TRY(u->rsum.ensure((size_t)rowsum_scratch(96)));
    MSBG_TRY(u->lpart.ensure((size_t)mse_blocks((int64_t)c * kVout)));
    MSBG_TRY(u->lsum.ensure(1));
  }
  u->cap = c;
  return 0;
}

int swizzle_all(msbg_unet *u) {
  if (u->swz_valid) return 0;
  for (int l = 0; l < 10; l++) {
    const LayerSpec &L = kL[l];
    MSBG_TRY(u->wf[l].ensure((size_t)swizzled_size(cin_k(l), L.cout)));
    MSBG_TRY(launch_swizzle(u->stream, u->p.p + u->offW[l], L.cin, cin_k(l), L.cout, false,
                            u->wf[l].p));
    if (l > 0) {  // backward data: a conv from cout to cin channels
      MSBG_TRY(u->wb[l].ensure((size_t)swizzled_size(L.cout, L.cin)));
      MSBG_TRY(launch_swizzle(u->stream, u->p.p + u->offW[l], L.cout, L.cout, L.cin, true,
                              u->wb[l].p));
    }
  }
  u->swz_valid = true;
  return 0;
}

// x0 already in u->x0 (n blocks); y -> u->y. Without `train` (inference)
// the last convolution applies the head in its store and its activation,
// which only the backward pass reads, is not written.
int forward(msbg_unet *u, int n, bool train) {
  hipStream_t s = u->stream;
  MSBG_TRY(swizzle_all(u));
  for (int l = 0; l < 10; l++) {
    const LayerSpec &L = kL[l];
    if (l == 6)  // decoder 2 reads [upsample(a5), a3] in place
      MSBG_TRY(launch_conv_cat(s, u->a[5].p, 64, u->a[3].p, 32, n, L.cout, L.dout, u->wf[l].p,
                               u->p.p + u->offB[l], true, u->a[l].p));
    else if (l == 8)  // decoder 1 reads [upsample(a7), a1] in place
      MSBG_TRY(launch_conv_cat(s, u->a[7].p, 32, u->a[1].p, 16, n, L.cout, L.dout, u->wf[l].p,
                               u->p.p + u->offB[l], true, u->a[l].p));
    else if (l == 9 && !train)
      MSBG_TRY(launch_conv_head(s, conv_input(u, l), n, cin_k(l), L.cout, L.dout, L.pad,
                                u->wf[l].p, u->p.p + u->offB[l], true, nullptr,
                                u->p.p + u->offH, u->y.p));
    else
      MSBG_TRY(launch_conv(s, conv_input(u, l), n, cin_k(l), L.cout, L.dout, L.pad, u->wf[l].p,
                           u->p.p + u->offB[l], true, u->a[l].p));
    if (l == 1) MSBG_TRY(launch_pool2(s, u->a[1].p, n, 32, 16, u->p1.p));
    if (l == 3) MSBG_TRY(launch_pool2(s, u->a[3].p, n, 16, 32, u->p2.p));
  }
  return train ? launch_head(s, u->a[9].p, n, u->p.p + u->offH, u->y.p) : 0;
}

// gradients of sum(dy * y) into u->g; dy in u->dy; forward state present
int backward(msbg_unet *u, int n) {
  hipStream_t s = u->stream;
  MSBG_TRY(launch_head_bwd(s, u->dy.p, u->a[9].p, n, u->p.p + u->offH, u->ga[9].p, u->hpart.p));
  MSBG_TRY(launch_sum_partials(s, u->hpart.p, head_bwd_blocks(n), kHead, u->g.p + u->offH));
  for (int l = 9; l >= 0; l--) {
    const LayerSpec &L = kL[l];
    const int64_t vo = vol(L.dout);
    uint16_t *gout = u->ga[l].p;  // dL/d a[l], becomes dL/d z[l] in place
    MSBG_TRY(launch_relu_bwd_xpose(s, gout, u->a[l].p, n, vo, L.cout, true, u->dzt.p));
    MSBG_TRY(launch_rowsum(s, u->dzt.p, L.cout, vo * n, u->rsum.p, u->g.p + u->offB[l]));
    // the weight gradient of a decoder conv needs its concatenated input:
    // built here, where only training needs it
    if (l == 8) MSBG_TRY(launch_upcat(s, u->a[7].p, 32, u->a[1].p, 16, n, 32, u->c9.p));
    if (l == 6) MSBG_TRY(launch_upcat(s, u->a[5].p, 64, u->a[3].p, 32, n, 16, u->c7.p));
    MSBG_TRY(launch_xpose_pad(s, conv_input(u, l), n, L.din, cin_k(l), L.pad, cin16(l),
                              u->xtp.p));
    MSBG_TRY(launch_wgrad(s, u->dzt.p, u->xtp.p, n, cin16(l), L.cout, L.dout, u->wpart.p));
    MSBG_TRY(launch_wgrad_reduce(s, u->wpart.p, wgrad_chunks(cin16(l), L.cout, L.dout, n),
                                 cin16(l), L.cin, L.cout, u->g.p + u->offW[l]));
    if (l == 0) break;
    // dL/d input of layer l
    uint16_t *gin;
    switch (l) {
      case 8: gin = u->gc9.p; break;
      case 6: gin = u->gc7.p; break;
      case 4: gin = u->gp2.p; break;
      case 2: gin = u->gp1.p; break;
      default: gin = u->ga[l - 1].p;
    }
    MSBG_TRY(launch_conv(s, gout, n, L.cout, L.cin, L.din, 2 - L.pad, u->wb[l].p, nullptr, false,
                         gin));
    if (l == 8) MSBG_TRY(launch_upcat_bwd(s, u->gc9.p, 32, 16, n, 32, u->ga[7].p, u->ga[1].p));
    if (l == 6) MSBG_TRY(launch_upcat_bwd(s, u->gc7.p, 64, 32, n, 16, u->ga[5].p, u->ga[3].p));
    if (l == 4) MSBG_TRY(launch_pool2_bwd(s, u->gp2.p, n, 16, 32, true, u->ga[3].p));
    if (l == 2) MSBG_TRY(launch_pool2_bwd(s, u->gp1.p, n, 32, 16, true, u->ga[1].p));
  }
  return 0;
}

int load_x(msbg_unet *u, const uint16_t *x, int n) {
  MSBG_HIP(hipMemcpyAsync(u->x0.p, x, sizeof(uint16_t) * n * kVin, hipMemcpyDefault, u->stream));
  return 0;
}

// the windows (and targets) of sample rows `rows` (host) of the resident
// sets X [N][36^3] and T [N][32^3], gathered on the device
int load_rows(msbg_unet *u, const uint16_t *X, const float *T, const int32_t *rows, int n) {
  MSBG_TRY(u->rows.ensure((size_t)std::max(n, u->cap)));
  MSBG_HIP(hipMemcpyAsync(u->rows.p, rows, sizeof(int32_t) * n, hipMemcpyHostToDevice, u->stream));
  MSBG_TRY(launch_gather_rows(u->stream, u->x0.p, X, u->rows.p, n, sizeof(uint16_t) * (int64_t)kVin));
  if (T)
    MSBG_TRY(launch_gather_rows(u->stream, u->tgt.p, T, u->rows.p, n, sizeof(float) * (int64_t)kVout));
  return 0;
}

// forward, MSE, backward and Adam on the loaded batch; with loss == NULL the
// step stays queued and its mean loss goes to the running accumulator
int train_core(msbg_unet *u, int n, float lr, double *loss) {
  MSBG_TRY(forward(u, n, true));
  const int64_t cnt = (int64_t)n * kVout;
  MSBG_TRY(launch_mse(u->stream, u->y.p, u->tgt.p, cnt, u->dy.p, u->lpart.p));
  MSBG_TRY(launch_sum_f64(u->stream, u->lpart.p, mse_blocks(cnt), u->lsum.p));
  MSBG_TRY(backward(u, n));
  u->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  MSBG_TRY(launch_adam(u->stream, u->p.p, u->g.p, u->m.p, u->v.p, u->nparam, lr, b1, b2, 1e-8f,
                       1.0f - std::pow(b1, (float)u->adam_t), 1.0f - std::pow(b2, (float)u->adam_t)));
  u->swz_valid = false;
  if (loss) {
    double sum = 0;
    MSBG_HIP(hipMemcpyAsync(&sum, u->lsum.p, sizeof(double), hipMemcpyDeviceToHost, u->stream));
    MSBG_HIP(hipStreamSynchronize(u->stream));
    *loss = sum / (double)cnt;
    return 0;
  }
  if (!u->lacc.p) {
    MSBG_TRY(u->lacc.alloc(2));
    MSBG_HIP(hipMemsetAsync(u->lacc.p, 0, 2 * sizeof(double), u->stream));
  }
  return launch_loss_acc(u->stream, u->lsum.p, 1.0 / (double)cnt, u->lacc.p);
}

}  // namespace

extern "C" {

int msbg_unet_create(uint64_t seed, msbg_unet **out) {
  MSBG_CHECK(out, MSBG_EINVAL, "msbg_unet_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("msbg_unet_create: no HIP device available");
    return MSBG_EDEVICE;
  }
  msbg_unet *u = new msbg_unet();
  if (hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking) != hipSuccess) {
    delete u;
    set_error("msbg_unet_create: stream creation failed");
    return MSBG_EDEVICE;
  }
  int64_t off = 0;
  for (int l = 0; l < 10; l++) {
    u->offW[l] = off;
    off += (int64_t)kL[l].cout * kL[l].cin * 27;
    u->offB[l] = off;
    off += kL[l].cout;
  }
  u->offH = off;
  off += kHead;
  u->nparam = off;
  // He-normal conv weights, zero biases, zero head: the initial network
  // predicts no change
  std::vector<float> h((size_t)off, 0.f);
  uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
  for (int l = 0; l < 10; l++) {
    const double sd = std::sqrt(2.0 / (kL[l].cin * 27.0));
    for (int64_t i = 0; i < (int64_t)kL[l].cout * kL[l].cin * 27; i++) {
      double u1 = ((splitmix(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      double u2 = (splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
      h[(size_t)(u->offW[l] + i)] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) *
                                            std::cos(6.283185307179586 * u2));
    }
  }
  int rc = 0;
  if ((rc = u->p.alloc((size_t)off)) || (rc = u->g.alloc((size_t)off)) ||
      (rc = u->m.alloc((size_t)off)) || (rc = u->v.alloc((size_t)off))) {
    msbg_unet_destroy(u);
    return rc;
  }
  if (hipMemcpy(u->p.p, h.data(), sizeof(float) * off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(u->g.p, 0, sizeof(float) * off) != hipSuccess ||
      hipMemset(u->m.p, 0, sizeof(float) * off) != hipSuccess ||
      hipMemset(u->v.p, 0, sizeof(float) * off) != hipSuccess) {
    msbg_unet_destroy(u);
    set_error("msbg_unet_create: parameter upload failed");
    return MSBG_EDEVICE;
  }
  *out = u;
  return 0;
}

int msbg_unet_destroy(msbg_unet *u) {
  if (!u) return 0;
  if (u->stream) (void)hipStreamSynchronize(u->stream);
  hipStream_t s = u->stream;
  delete u;
  if (s) (void)hipStreamDestroy(s);
  return 0;
}

int msbg_unet_num_params(msbg_unet *u, int64_t *n) {
  MSBG_CHECK(u && n, MSBG_EINVAL, "msbg_unet_num_params: bad args");
  *n = u->nparam;
  return 0;
}

int msbg_unet_get_params(msbg_unet *u, float *host, int64_t n) {
  MSBG_CHECK(u && host && n == u->nparam, MSBG_EINVAL, "msbg_unet_get_params: bad args");
  MSBG_HIP(hipStreamSynchronize(u->stream));
  MSBG_HIP(hipMemcpy(host, u->p.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  return 0;
}

int msbg_unet_set_params(msbg_unet *u, const float *host, int64_t n) {
  MSBG_CHECK(u && host && n == u->nparam, MSBG_EINVAL, "msbg_unet_set_params: bad args");
  MSBG_HIP(hipStreamSynchronize(u->stream));
  MSBG_HIP(hipMemcpy(u->p.p, host, sizeof(float) * n, hipMemcpyHostToDevice));
  MSBG_HIP(hipMemset(u->m.p, 0, sizeof(float) * n));
  MSBG_HIP(hipMemset(u->v.p, 0, sizeof(float) * n));
  u->adam_t = 0;
  u->swz_valid = false;
  return 0;
}

int msbg_unet_get_grads(msbg_unet *u, float *host, int64_t n) {
  MSBG_CHECK(u && host && n == u->nparam, MSBG_EINVAL, "msbg_unet_get_grads: bad args");
  MSBG_HIP(hipStreamSynchronize(u->stream));
  MSBG_HIP(hipMemcpy(host, u->g.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  return 0;
}

int msbg_unet_forward(msbg_unet *u, const uint16_t *x, int n, float *y) {
  MSBG_CHECK(u && x && y, MSBG_EINVAL, "msbg_unet_forward: bad args");
  MSBG_TRY(ensure_cap(u, n, false));
  MSBG_TRY(load_x(u, x, n));
  MSBG_TRY(forward(u, n, false));
  MSBG_HIP(hipMemcpyAsync(y, u->y.p, sizeof(float) * n * kVout, hipMemcpyDefault, u->stream));
  MSBG_HIP(hipStreamSynchronize(u->stream));
  return 0;
}

int msbg_unet_backward(msbg_unet *u, const uint16_t *x, const float *dy, int n) {
  MSBG_CHECK(u && x && dy, MSBG_EINVAL, "msbg_unet_backward: bad args");
  MSBG_CHECK(n % 8 == 0 && (n & (n - 1)) == 0, MSBG_EINVAL,
             "msbg_unet_backward: batch %d must be 8 * 2^k", n);
  MSBG_TRY(ensure_cap(u, n, true));
  MSBG_TRY(load_x(u, x, n));
  MSBG_TRY(forward(u, n, true));
  MSBG_HIP(hipMemcpyAsync(u->dy.p, dy, sizeof(float) * n * kVout, hipMemcpyDefault, u->stream));
  MSBG_TRY(backward(u, n));
  MSBG_HIP(hipStreamSynchronize(u->stream));
  return 0;
}

int msbg_unet_train_step(msbg_unet *u, const uint16_t *x, const float *target, int n, float lr,
                         double *loss) {
  MSBG_CHECK(u && x && target && lr >= 0.f, MSBG_EINVAL, "msbg_unet_train_step: bad args");
  MSBG_CHECK(n % 8 == 0 && (n & (n - 1)) == 0, MSBG_EINVAL,
             "msbg_unet_train_step: batch %d must be 8 * 2^k", n);
  MSBG_TRY(ensure_cap(u, n, true));
  MSBG_TRY(load_x(u, x, n));
  MSBG_HIP(hipMemcpyAsync(u->tgt.p, target, sizeof(float) * (int64_t)n * kVout, hipMemcpyDefault,
                          u->stream));
  return train_core(u, n, lr, loss);
}

int msbg_unet_train_step_rows(msbg_unet *u, const uint16_t *X, const float *T, const int32_t *rows,
                              int n, float lr, double *loss) {
  MSBG_CHECK(u && X && T && rows && lr >= 0.f, MSBG_EINVAL, "msbg_unet_train_step_rows: bad args");
  MSBG_CHECK(n % 8 == 0 && (n & (n - 1)) == 0, MSBG_EINVAL,
             "msbg_unet_train_step_rows: batch %d must be 8 * 2^k", n);
  for (int i = 0; i < n; i++)
    MSBG_CHECK(rows[i] >= 0, MSBG_EINVAL, "msbg_unet_train_step_rows: row %d negative", i);
  MSBG_TRY(ensure_cap(u, n, true));
  MSBG_TRY(load_rows(u, X, T, rows, n));
  return train_core(u, n, lr, loss);
}

int msbg_unet_forward_rows(msbg_unet *u, const uint16_t *X, const int32_t *rows, int n, float *y) {
  MSBG_CHECK(u && X && rows && y && n > 0, MSBG_EINVAL, "msbg_unet_forward_rows: bad args");
  for (int i = 0; i < n; i++)
    MSBG_CHECK(rows[i] >= 0, MSBG_EINVAL, "msbg_unet_forward_rows: row %d negative", i);
  MSBG_TRY(ensure_cap(u, n, false));
  MSBG_TRY(load_rows(u, X, nullptr, rows, n));
  MSBG_TRY(forward(u, n, false));
  MSBG_HIP(hipMemcpyAsync(y, u->y.p, sizeof(float) * n * kVout, hipMemcpyDefault, u->stream));
  MSBG_HIP(hipStreamSynchronize(u->stream));
  return 0;
}

int msbg_unet_take_loss(msbg_unet *u, double *sum, int64_t *steps) {
  MSBG_CHECK(u && sum, MSBG_EINVAL, "msbg_unet_take_loss: bad args");
  double h[2] = {0, 0};
  if (u->lacc.p) {
    MSBG_HIP(hipMemcpyAsync(h, u->lacc.p, sizeof(h), hipMemcpyDeviceToHost, u->stream));
    MSBG_HIP(hipStreamSynchronize(u->stream));
    MSBG_HIP(hipMemsetAsync(u->lacc.p, 0, sizeof(h), u->stream));
  }
  *sum = h[0];
  if (steps) *steps = (int64_t)h[1];
  return 0;
}

}  // extern "C"

namespace {

// validate host block ids (level-0 value blocks of a 32^3-block channel) and
// upload them; returns the channel level
int block_list(msbg_grid *g, int chan, const int32_t *blocks, int n, const char *who,
               ChanLevel **pcl, int *storage, DevBuf<int32_t> &d) {
  MSBG_CHECK(g && blocks && n >= 0, MSBG_EINVAL, "%s: bad args", who);
  MSBG_CHECK(g->bsx0 == 32, MSBG_EUNSUPP, "%s: the network takes 32^3 blocks", who);
  auto it = g->channels.find(chan);
  MSBG_CHECK(it != g->channels.end() && it->second.lv[0].defined, MSBG_ESTATE,
             "%s: channel %d has no level 0", who, chan);
  ChanLevel &cl = it->second.lv[0];
  for (int i = 0; i < n; i++)
    MSBG_CHECK(blocks[i] >= 0 && blocks[i] < g->nblocks && cl.h_page[blocks[i]] >= 0,
               MSBG_EINVAL, "%s: block %d is not a value block", who, blocks[i]);
  *pcl = &cl;
  *storage = it->second.storage;
  if (!n) return 0;
  MSBG_TRY(d.alloc((size_t)n));
  MSBG_HIP(hipMemcpyAsync(d.p, blocks, sizeof(int32_t) * n, hipMemcpyHostToDevice, g->stream));
  return 0;
}

}  // namespace

// ---- msbg_cnn_conv3d_noalloc's scratch ------------------------------------
// The swizzled weights of one convolution, in a file-scope device array
// instead of a per-call allocation: 16384 uint4 (256 KiB) per device, above
// the 13824 of the largest layer shape (64 -> 64 channels). Every call on a
// device shares that device's copy. hipGetSymbolAddress gives each device's
// own address, once per device (with the device current), and only that
// address goes to the kernels: the array's name is a host-side stand-in.
static constexpr int64_t kConvScratchElems = 16384;
static __device__ __align__(256) uint4 scratch[kConvScratchElems];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint4 *g_scratch_ptr[kMaxDevices] = {};

static int conv_scratch(uint4 **out) {
  int dev = -1;
  MSBG_HIP(hipGetDevice(&dev));
  MSBG_CHECK(dev >= 0 && dev < kMaxDevices, MSBG_EDEVICE, "conv scratch: device %d beyond %d",
             dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    MSBG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MSBG_CHECK(p, MSBG_EDEVICE, "conv scratch: no device address for device %d", dev);
    g_scratch_ptr[dev] = (uint4 *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

extern "C" {

int msbg_unet_gather(msbg_grid *g, int chan, const int32_t *blocks, int n, uint16_t *x) {
  ChanLevel *cl;
  int st;
  DevBuf<int32_t> d;
  MSBG_CHECK(x, MSBG_EINVAL, "msbg_unet_gather: x is NULL");
  MSBG_TRY(block_list(g, chan, blocks, n, "msbg_unet_gather", &cl, &st, d));
  if (!n) return 0;
  MSBG_TRY(launch_gather36(g->stream, st, cl->d_page.p, cl->d_pages.p, cl->empty_raw,
                           cl->full_raw, g->nbx, g->nby, g->nbz, d.p, n, x));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int msbg_unet_extract(msbg_grid *g, int chan, const int32_t *blocks, int n, float *y) {
  ChanLevel *cl;
  int st;
  DevBuf<int32_t> d;
  MSBG_CHECK(y, MSBG_EINVAL, "msbg_unet_extract: y is NULL");
  MSBG_TRY(block_list(g, chan, blocks, n, "msbg_unet_extract", &cl, &st, d));
  if (!n) return 0;
  MSBG_TRY(launch_extract32(g->stream, st, cl->d_page.p, cl->d_pages.p, d.p, n, y));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int msbg_unet_scatter(msbg_grid *g, int chan, const int32_t *blocks, int n, const float *y) {
  ChanLevel *cl;
  int st;
  DevBuf<int32_t> d;
  MSBG_CHECK(y, MSBG_EINVAL, "msbg_unet_scatter: y is NULL");
  MSBG_TRY(block_list(g, chan, blocks, n, "msbg_unet_scatter", &cl, &st, d));
  if (!n) return 0;
  MSBG_TRY(launch_scatter32(g->stream, st, cl->d_page.p, cl->d_pages.p, d.p, n, y));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  pages_written(*cl);
  return 0;
}

int msbg_warm_start(msbg_grid *g, int chan, msbg_unet *u, int batch, double *ms) {
  // predictions of all blocks are made from the unchanged field, then added
  MSBG_CHECK(g && u && batch >= 1, MSBG_EINVAL, "msbg_warm_start: bad args");
  MSBG_CHECK(g->bsx0 == 32, MSBG_EUNSUPP, "msbg_warm_start: the network takes 32^3 blocks");
  auto it = g->channels.find(chan);
  MSBG_CHECK(it != g->channels.end() && it->second.lv[0].defined, MSBG_ESTATE,
             "msbg_warm_start: channel %d has no level 0", chan);
  ChanLevel &cl = it->second.lv[0];
  // every level-0 value block (the blocks the sweep updates, msbg4.cpp:467-470)
  std::vector<int32_t> bids;
  for (int64_t b = 0; b < g->nblocks; b++)
    if (cl.h_page[b] >= 0 && g->h_level[b] == 0) bids.push_back((int32_t)b);
  const int n = (int)bids.size();
  if (ms) *ms = 0;
  if (!n) return 0;
  MSBG_TRY(ensure_cap(u, std::min(batch, n), false));
  DevBuf<int32_t> d;
  DevBuf<float> pred;  // predictions of all blocks before any is written back
  MSBG_TRY(d.alloc((size_t)n));
  MSBG_TRY(pred.alloc((size_t)n * kVout));
  MSBG_HIP(hipStreamSynchronize(g->stream));
  hipEvent_t e0, e1;
  MSBG_HIP(hipEventCreate(&e0));
  MSBG_HIP(hipEventCreate(&e1));
  int rc = 0;
  do {
    if (hipEventRecord(e0, u->stream) != hipSuccess ||
        hipMemcpyAsync(d.p, bids.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, u->stream) !=
            hipSuccess) {
      set_error("msbg_warm_start: upload failed");
      rc = MSBG_EDEVICE;
      break;
    }
    for (int b0 = 0; b0 < n && !rc; b0 += batch) {
      const int nb = std::min(batch, n - b0);
      rc = launch_gather36(u->stream, it->second.storage, cl.d_page.p, cl.d_pages.p, cl.empty_raw,
                           cl.full_raw, g->nbx, g->nby, g->nbz, d.p + b0, nb, u->x0.p);
      if (!rc) rc = forward(u, nb, false);
      if (!rc && hipMemcpyAsync(pred.p + (int64_t)b0 * kVout, u->y.p, sizeof(float) * nb * kVout,
                                hipMemcpyDeviceToDevice, u->stream) != hipSuccess) {
        set_error("msbg_warm_start: copy failed");
        rc = MSBG_EDEVICE;
      }
    }
    if (!rc)
      rc = launch_scatter32(u->stream, it->second.storage, cl.d_page.p, cl.d_pages.p, d.p, n,
                            pred.p);
    if (!rc && (hipEventRecord(e1, u->stream) != hipSuccess ||
                hipStreamSynchronize(u->stream) != hipSuccess)) {
      set_error("msbg_warm_start: device error");
      rc = MSBG_EDEVICE;
    }
  } while (0);
  float t = 0;
  if (!rc && ms && hipEventElapsedTime(&t, e0, e1) == hipSuccess) *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  pages_written(cl);
  return rc;
}

// Layer-level entry points for tests: one convolution, its backward-data
// and backward-weights products, on device tensors.
int msbg_cnn_conv3d(const uint16_t *x, int n, int cin, int cout, int dout, int pad,
                    const float *w, const float *b, int relu, uint16_t *y) {
  MSBG_CHECK(x && w && y && (cin == 1 || cin % 8 == 0), MSBG_EINVAL,
             "msbg_cnn_conv3d: bad args");
  DevBuf<uint4> sw;
  MSBG_TRY(sw.alloc((size_t)swizzled_size(cin, cout)));
  MSBG_TRY(launch_swizzle(nullptr, w, cin, cin, cout, false, sw.p));
  MSBG_TRY(launch_conv(nullptr, x, n, cin, cout, dout, pad, sw.p, b, relu != 0, y));
  MSBG_HIP(hipStreamSynchronize(nullptr));
  return 0;
}

int msbg_cnn_conv3d_noalloc(const uint16_t *x, int n, int cin, int cout, int dout, int pad,
                            const float *w, const float *b, int relu, uint16_t *y) {
  MSBG_CHECK(x && w && y && (cin == 1 || cin % 8 == 0) && cout > 0 && cout % 16 == 0,
             MSBG_EINVAL, "msbg_cnn_conv3d_noalloc: bad args");
  const int64_t need = swizzled_size(cin, cout);
  MSBG_CHECK(need <= kConvScratchElems, MSBG_EINVAL,
             "msbg_cnn_conv3d_noalloc: the swizzled weights of %d -> %d channels take %lld bytes, "
             "above the %lld-byte scratch",
             cin, cout, (long long)(need * 16), (long long)(kConvScratchElems * 16));
  uint4 *sw = nullptr;
  MSBG_TRY(conv_scratch(&sw));
  MSBG_TRY(launch_swizzle(nullptr, w, cin, cin, cout, false, sw));
  MSBG_TRY(launch_conv(nullptr, x, n, cin, cout, dout, pad, sw, b, relu != 0, y));
  MSBG_HIP(hipStreamSynchronize(nullptr));
  return 0;
}

int msbg_cnn_conv3d_bwd_data(const uint16_t *dy, int n, int cin, int cout, int din, int pad,
                             const float *w, uint16_t *dx) {
  MSBG_CHECK(dy && w && dx && cout % 8 == 0 && cin % 16 == 0, MSBG_EINVAL,
             "msbg_cnn_conv3d_bwd_data: bad args");
  DevBuf<uint4> sw;
  MSBG_TRY(sw.alloc((size_t)swizzled_size(cout, cin)));
  MSBG_TRY(launch_swizzle(nullptr, w, cout, cout, cin, true, sw.p));
  MSBG_TRY(launch_conv(nullptr, dy, n, cout, cin, din, 2 - pad, sw.p, nullptr, false, dx));
  MSBG_HIP(hipStreamSynchronize(nullptr));
  return 0;
}

int msbg_cnn_conv3d_bwd_weights(const uint16_t *x, const uint16_t *dy, int n, int cin, int cout,
                                int dout, int pad, float *dw, float *db) {
  MSBG_CHECK(x && dy && dw && (cin == 1 || cin % 8 == 0), MSBG_EINVAL,
             "msbg_cnn_conv3d_bwd_weights: bad args");
  MSBG_CHECK(n % 8 == 0 && (n & (n - 1)) == 0, MSBG_EINVAL,
             "msbg_cnn_conv3d_bwd_weights: batch %d must be 8 * 2^k", n);
  const int c16 = (cin + 15) / 16 * 16, din = dout + 2 - 2 * pad;
  const int chunks = wgrad_chunks(c16, cout, dout, n);
  DevBuf<uint16_t> dzt, xtp, dyc;
  DevBuf<float> part, rs;
  MSBG_CHECK(cout <= 128, MSBG_EINVAL, "msbg_cnn_conv3d_bwd_weights: cout %d > 128", cout);
  MSBG_TRY(rs.alloc((size_t)rowsum_scratch(cout)));
  MSBG_TRY(dyc.alloc((size_t)n * vol(dout) * cout));
  MSBG_TRY(dzt.alloc((size_t)n * vol(dout) * cout));
  MSBG_TRY(xtp.alloc((size_t)n * vol(dout + 2) * c16));
  MSBG_TRY(part.alloc((size_t)chunks * 27 * cout * c16));
  MSBG_HIP(hipMemcpy(dyc.p, dy, sizeof(uint16_t) * n * vol(dout) * cout, hipMemcpyDefault));
  MSBG_TRY(launch_relu_bwd_xpose(nullptr, dyc.p, nullptr, n, vol(dout), cout, false, dzt.p));
  MSBG_TRY(launch_xpose_pad(nullptr, x, n, din, cin, pad, c16, xtp.p));
  MSBG_TRY(launch_wgrad(nullptr, dzt.p, xtp.p, n, c16, cout, dout, part.p));
  MSBG_TRY(launch_wgrad_reduce(nullptr, part.p, chunks, c16, cin, cout, dw));
  if (db) MSBG_TRY(launch_rowsum(nullptr, dzt.p, cout, vol(dout) * n, rs.p, db));
  MSBG_HIP(hipStreamSynchronize(nullptr));
  return 0;
}

}  // extern "C"
