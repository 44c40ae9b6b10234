"""FP64 CPU oracle of the warm-start network (test infrastructure only).

The layer primitives are C (cnn_oracle.c, direct loops in double); this
module composes them into the network the product library builds in
msbg_amd/csrc/cnn.cpp and restates its backward pass. With round_bf16=True
every tensor the GPU stores as BF16 is rounded to BF16 at the same point, so
the remaining difference to the GPU is FP32-vs-FP64 accumulation.

The network has no counterpart in the reference; its definition is the
build's: layers (cin, cout, pad, din) below, ReLU after every convolution,
2x2x2 average pools, nearest-neighbour upsampling, channel concatenation
(upsampled first), a 1x1 head whose output is the predicted change.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

LAYERS = [(1, 16, 0, 36), (16, 16, 0, 34), (16, 32, 1, 16), (32, 32, 1, 16), (32, 64, 1, 8),
          (64, 64, 1, 8), (96, 32, 1, 16), (32, 32, 1, 16), (48, 16, 1, 32), (16, 16, 1, 32)]

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libcnn_oracle.so")
        if not os.path.exists(path):
            raise OSError(f"{path} not built; run `make -C oracle`")
        L = C.CDLL(path)
        P, I, I64 = C.c_void_p, C.c_int, C.c_int64
        sig = {
            "or_conv3d_fwd": [P, I, I, I, P, P, I, I, I, P],
            "or_conv3d_bwd_data": [P, I, I, I, P, I, I, P],
            "or_conv3d_bwd_weights": [P, P, I, I, I, I, I, P, P],
            "or_pool2_fwd": [P, I, I, I, P],
            "or_pool2_bwd": [P, I, I, I, I, P],
            "or_upcat_fwd": [P, I, P, I, I, I, P],
            "or_upcat_bwd": [P, I, I, I, I, P, P],
            "or_round_bf16": [P, I64],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.restype = None
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def round_bf16(a):
    """Round to the nearest bfloat16 (ties to even), as float64 (or_round_bf16)."""
    a = np.array(a, dtype=np.float64, copy=True, order="C")
    lib().or_round_bf16(_p(a), a.size)
    return a


def conv3d(x, w, b, pad, relu):
    """x [n][d][d][d][cin], w [cout][cin][27], b [cout] or None -> [n][dout]^3[cout]."""
    x, w = _c(x), _c(w)
    n, din, cin = x.shape[0], x.shape[1], x.shape[4]
    cout = w.shape[0]
    dout = din + 2 * pad - 2
    y = np.empty((n, dout, dout, dout, cout))
    lib().or_conv3d_fwd(_p(x), n, din, cin, _p(w), None if b is None else _p(_c(b)), cout, pad,
                        int(relu), _p(y))
    return y


def conv3d_bwd_data(dy, w, pad):
    dy, w = _c(dy), _c(w)
    n, dout, cout = dy.shape[0], dy.shape[1], dy.shape[4]
    cin = w.shape[1]
    din = dout - 2 * pad + 2
    dx = np.empty((n, din, din, din, cin))
    lib().or_conv3d_bwd_data(_p(dy), n, dout, cout, _p(w), cin, pad, _p(dx))
    return dx


def conv3d_bwd_weights(x, dy, pad):
    x, dy = _c(x), _c(dy)
    n, din, cin, cout = x.shape[0], x.shape[1], x.shape[4], dy.shape[4]
    dw = np.empty((cout, cin, 27))
    db = np.empty(cout)
    lib().or_conv3d_bwd_weights(_p(x), _p(dy), n, din, cin, cout, pad, _p(dw), _p(db))
    return dw, db


def pool2(x):
    x = _c(x)
    n, d, c = x.shape[0], x.shape[1], x.shape[4]
    y = np.empty((n, d // 2, d // 2, d // 2, c))
    lib().or_pool2_fwd(_p(x), n, d, c, _p(y))
    return y


def pool2_bwd(dy, acc=None):
    dy = _c(dy)
    n, h, c = dy.shape[0], dy.shape[1], dy.shape[4]
    d = 2 * h
    dx = np.zeros((n, d, d, d, c)) if acc is None else _c(acc).copy()
    lib().or_pool2_bwd(_p(dy), n, d, c, 0 if acc is None else 1, _p(dx))
    return dx


def upcat(a, b):
    a, b = _c(a), _c(b)
    n, d, ca, cb = b.shape[0], b.shape[1], a.shape[4], b.shape[4]
    out = np.empty((n, d, d, d, ca + cb))
    lib().or_upcat_fwd(_p(a), ca, _p(b), cb, n, d, _p(out))
    return out


def upcat_bwd(dout, ca):
    dout = _c(dout)
    n, d, c = dout.shape[0], dout.shape[1], dout.shape[4]
    h = d // 2
    da = np.empty((n, h, h, h, ca))
    db = np.empty((n, d, d, d, c - ca))
    lib().or_upcat_bwd(_p(dout), ca, c - ca, n, d, _p(da), _p(db))
    return da, db


def unpack(params):
    """Flat parameter vector -> ([(w, b)] * 10, hw, hb), msbg_gpu.h layout."""
    p = np.asarray(params, dtype=np.float64)
    out, o = [], 0
    for cin, cout, _, _ in LAYERS:
        w = p[o:o + cout * cin * 27].reshape(cout, cin, 27)
        o += cout * cin * 27
        out.append((w, p[o:o + cout]))
        o += cout
    hw, hb = p[o:o + 16], p[o + 16]
    assert o + 17 == p.size, (o + 17, p.size)
    return out, hw, hb


def num_params():
    return sum(cout * cin * 27 + cout for cin, cout, _, _ in LAYERS) + 17


def forward(params, x0, round_bf16_=False):
    """x0 [n][36][36][36] -> y [n][32][32][32]; also returns the activations."""
    R = round_bf16 if round_bf16_ else (lambda a: a)
    convs, hw, hb = unpack(params)
    x0 = _c(x0)
    a = [None] * 10
    inp = {0: x0[..., None]}

    def conv(l, xin):
        w, b = convs[l]
        return R(conv3d(xin, w, b, LAYERS[l][2], True))

    a[0] = conv(0, inp[0])
    a[1] = conv(1, a[0])
    inp[2] = R(pool2(a[1]))
    a[2] = conv(2, inp[2])
    a[3] = conv(3, a[2])
    inp[4] = R(pool2(a[3]))
    a[4] = conv(4, inp[4])
    a[5] = conv(5, a[4])
    inp[6] = upcat(a[5], a[3])
    a[6] = conv(6, inp[6])
    a[7] = conv(7, a[6])
    inp[8] = upcat(a[7], a[1])
    a[8] = conv(8, inp[8])
    a[9] = conv(9, a[8])
    for l in (1, 3, 5, 7, 9):
        inp[l] = a[l - 1]
    y = hb + a[9] @ hw
    if round_bf16_:
        y = y.astype(np.float32).astype(np.float64)
    return y, dict(a=a, inp=inp)


def backward(params, x0, dy, round_bf16_=False):
    """Gradient of sum(dy * y) with respect to the flat parameters."""
    R = round_bf16 if round_bf16_ else (lambda a: a)
    convs, hw, hb = unpack(params)
    _, cache = forward(params, x0, round_bf16_)
    a, inp = cache["a"], cache["inp"]
    dy = _c(dy)
    grads = [None] * 10
    ga = [None] * 10
    ga[9] = R(dy[..., None] * hw)
    dhw = np.einsum("nzyx,nzyxc->c", dy, a[9])
    dhb = dy.sum()
    for l in range(9, -1, -1):
        cin, cout, pad, din = LAYERS[l]
        dz = R(ga[l] * (a[l] > 0))
        dw, db = conv3d_bwd_weights(inp[l], dz, pad)
        grads[l] = (dw, db)
        if l == 0:
            break
        gin = R(conv3d_bwd_data(dz, convs[l][0], pad))
        if l == 8:
            da, db_ = upcat_bwd(gin, 32)
            ga[7], ga[1] = R(da), R(db_)
        elif l == 6:
            da, db_ = upcat_bwd(gin, 64)
            ga[5], ga[3] = R(da), R(db_)
        elif l == 4:
            ga[3] = R(pool2_bwd(gin, ga[3]))
        elif l == 2:
            ga[1] = R(pool2_bwd(gin, ga[1]))
        else:
            ga[l - 1] = gin
    flat = []
    for dw, db in grads:
        flat += [dw.ravel(), db.ravel()]
    flat += [dhw, [dhb]]
    return np.concatenate([np.asarray(f, dtype=np.float64).ravel() for f in flat])
