"""msbg_cnn_conv3d_noalloc (no device allocation; fixed per-device scratch)
against msbg_cnn_conv3d through the C ABI: bit-identical outputs on calls in
a row, each later one with a smaller input than the one before, and a clear
error for weights beyond the scratch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M = pytest.importorskip("msbg_amd")
from msbg_amd._lib import EINVAL, lib  # noqa: E402


def _bf16(a):
    return (np.ascontiguousarray(a, np.float32).view(np.uint32) >> 16).astype(np.uint16)


def _case(rng, n, cin, cout, dout, pad):
    din = dout + 2 - 2 * pad
    x = M.DeviceArray.from_numpy(_bf16(rng.random((n, din, din, din, cin)) * 2 - 1))
    w = M.DeviceArray.from_numpy((rng.normal(size=(cout, cin, 27)) * 0.2).astype(np.float32))
    b = M.DeviceArray.from_numpy((rng.normal(size=cout) * 0.1).astype(np.float32))
    return x, w, b, (n, dout, dout, dout, cout)


def _run(fn, x, n, cin, cout, dout, pad, w, b, relu, shape):
    y = M.DeviceArray(shape, np.uint16)
    rc = getattr(lib(), fn)(C.c_void_p(x.data_ptr()), n, cin, cout, dout, pad,
                            C.c_void_p(w.data_ptr()), C.c_void_p(b.data_ptr()), relu,
                            C.c_void_p(y.data_ptr()))
    assert rc == 0, lib().msbg_last_error()
    out = y.numpy()
    y.free()
    return out


def test_noalloc_conv_matches_allocating_conv():
    rng = np.random.default_rng(11)
    # (n, cin, cout, dout, pad): the largest weight set, then smaller inputs
    for n, cin, cout, dout, pad, relu in [(4, 64, 64, 8, 1, 1), (2, 64, 64, 8, 1, 0),
                                          (1, 16, 16, 32, 0, 1), (1, 1, 16, 34, 0, 1)]:
        print("[case] conv noalloc", n, cin, cout, dout, pad, relu, flush=True)
        x, w, b, shape = _case(rng, n, cin, cout, dout, pad)
        want = _run("msbg_cnn_conv3d", x, n, cin, cout, dout, pad, w, b, relu, shape)
        got = _run("msbg_cnn_conv3d_noalloc", x, n, cin, cout, dout, pad, w, b, relu, shape)
        assert np.array_equal(got, want)
        assert np.count_nonzero(want) > want.size // 8  # a real output, not zeros
        for a in (x, w, b):
            a.free()


def test_noalloc_conv_rejects_weights_beyond_scratch():
    # 128 -> 128 channels: the swizzled weights exceed the scratch; rejected
    # before any launch
    x = M.DeviceArray((16,), np.uint16)
    w = M.DeviceArray((16,), np.float32)
    y = M.DeviceArray((16,), np.uint16)
    rc = lib().msbg_cnn_conv3d_noalloc(C.c_void_p(x.data_ptr()), 1, 128, 128, 8, 1,
                                       C.c_void_p(w.data_ptr()), None, 1, C.c_void_p(y.data_ptr()))
    assert rc == EINVAL
    assert b"scratch" in lib().msbg_last_error()
