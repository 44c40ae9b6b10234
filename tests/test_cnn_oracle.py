"""CPU pins of the warm-start network's FP64 oracle (oracle/cnn.py,
oracle/cnn_oracle.c): each layer against torch's CPU library routines in
float64, the composed backward pass against torch autograd of an
independently built network, and finite differences."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
F = torch.nn.functional

from oracle import cnn as O  # noqa: E402


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


def _ncdhw(x):  # channels-last -> torch layout
    return _t(x).permute(0, 4, 1, 2, 3).contiguous()


def _cl(t):
    return t.permute(0, 2, 3, 4, 1).contiguous().numpy()


def _w5(w):  # [cout][cin][27] -> [cout][cin][3][3][3] (tap = (dz*3+dy)*3+dx)
    return _t(w.reshape(w.shape[0], w.shape[1], 3, 3, 3))


@pytest.mark.parametrize("cin,cout,d,pad,relu", [(1, 16, 10, 0, True), (16, 8, 6, 1, False),
                                                 (8, 16, 7, 1, True), (16, 16, 5, 2, False)])
def test_conv3d_matches_torch(cin, cout, d, pad, relu):
    rng = np.random.default_rng(cin * 100 + d)
    x = rng.standard_normal((2, d, d, d, cin))
    w = rng.standard_normal((cout, cin, 27))
    b = rng.standard_normal(cout)
    got = O.conv3d(x, w, b, pad, relu)
    ref = F.conv3d(_ncdhw(x), _w5(w), _t(b), padding=pad)
    if relu:
        ref = F.relu(ref)
    np.testing.assert_allclose(got, _cl(ref), rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("cin,cout,d,pad", [(8, 16, 8, 1), (16, 8, 8, 0), (4, 4, 6, 1)])
def test_conv3d_backward_matches_torch(cin, cout, d, pad):
    rng = np.random.default_rng(cin + cout + d)
    x = rng.standard_normal((2, d, d, d, cin))
    w = rng.standard_normal((cout, cin, 27))
    dout = d + 2 * pad - 2
    dy = rng.standard_normal((2, dout, dout, dout, cout))
    xt = _ncdhw(x).requires_grad_(True)
    wt = _w5(w).requires_grad_(True)
    bt = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    y = F.conv3d(xt, wt, bt, padding=pad)
    (y * _ncdhw(dy)).sum().backward()
    dx = O.conv3d_bwd_data(dy, w, pad)
    dw, db = O.conv3d_bwd_weights(x, dy, pad)
    np.testing.assert_allclose(dx, _cl(xt.grad), rtol=1e-11, atol=1e-10)
    np.testing.assert_allclose(dw, wt.grad.reshape(cout, cin, 27).numpy(), rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-11, atol=1e-9)


def test_pool_upcat_match_torch():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 8, 8, 8, 8))
    np.testing.assert_allclose(O.pool2(x), _cl(F.avg_pool3d(_ncdhw(x), 2)), rtol=1e-13)
    a = rng.standard_normal((2, 4, 4, 4, 16))
    b = rng.standard_normal((2, 8, 8, 8, 8))
    up = F.interpolate(_ncdhw(a), scale_factor=2, mode="nearest")
    np.testing.assert_array_equal(O.upcat(a, b), _cl(torch.cat([up, _ncdhw(b)], 1)))
    # backward of both through autograd
    g = rng.standard_normal((2, 8, 8, 8, 24))
    at = _ncdhw(a).requires_grad_(True)
    bt = _ncdhw(b).requires_grad_(True)
    (torch.cat([F.interpolate(at, scale_factor=2, mode="nearest"), bt], 1) * _ncdhw(g)).sum().backward()
    da, db = O.upcat_bwd(g, 16)
    np.testing.assert_allclose(da, _cl(at.grad), rtol=1e-12)
    np.testing.assert_allclose(db, _cl(bt.grad), rtol=1e-12)
    xt = _ncdhw(x).requires_grad_(True)
    gy = rng.standard_normal((2, 4, 4, 4, 8))
    (F.avg_pool3d(xt, 2) * _ncdhw(gy)).sum().backward()
    acc = rng.standard_normal(x.shape)
    np.testing.assert_allclose(O.pool2_bwd(gy, acc) - acc, _cl(xt.grad), rtol=1e-12, atol=1e-14)


def test_round_bf16():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0, 3.0e38])
    r = O.round_bf16(v)
    # ties to even at 2^-8 spacing near 1; exact values kept
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 4 * 2 ** -8 and r[3] == 1.0
    assert r[4] == -2.5 and r[5] == 0.0
    t = torch.tensor(v, dtype=torch.float64).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(r, t)


def _torch_net(params, x0):
    """The same network built from torch library layers (float64, CPU)."""
    convs, hw, hb = O.unpack(params)
    ws = [(_w5(w).requires_grad_(True), _t(b.copy()).requires_grad_(True)) for w, b in convs]
    hwt = _t(hw.copy()).requires_grad_(True)
    hbt = torch.tensor(float(hb), dtype=torch.float64, requires_grad=True)
    x = _t(x0)[:, None]

    def c(l, t):
        return F.relu(F.conv3d(t, ws[l][0], ws[l][1], padding=O.LAYERS[l][2]))

    e1 = c(1, c(0, x))
    e2 = c(3, c(2, F.avg_pool3d(e1, 2)))
    bb = c(5, c(4, F.avg_pool3d(e2, 2)))
    d2 = c(7, c(6, torch.cat([F.interpolate(bb, scale_factor=2, mode="nearest"), e2], 1)))
    d1 = c(9, c(8, torch.cat([F.interpolate(d2, scale_factor=2, mode="nearest"), e1], 1)))
    y = hbt + torch.einsum("nczyx,c->nzyx", d1, hwt)
    leaves = [t for pair in ws for t in pair] + [hwt, hbt]
    return y, leaves


def _params(seed):
    rng = np.random.default_rng(seed)
    p = []
    for cin, cout, _, _ in O.LAYERS:
        p += [rng.standard_normal(cout * cin * 27) * np.sqrt(2.0 / (cin * 27)),
              rng.standard_normal(cout) * 0.05]
    p += [rng.standard_normal(16) * 0.3, [0.01]]
    return np.concatenate([np.asarray(v, dtype=np.float64).ravel() for v in p])


def test_network_forward_backward_match_torch_autograd():
    params = _params(11)
    assert params.size == O.num_params()
    rng = np.random.default_rng(12)
    x0 = rng.uniform(0, 1, (1, 36, 36, 36))
    dy = rng.standard_normal((1, 32, 32, 32))
    y, _ = O.forward(params, x0)
    yt, leaves = _torch_net(params, x0)
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-10, atol=1e-10)
    (yt * _t(dy)).sum().backward()
    ref = np.concatenate([t.grad.numpy().ravel() if t.dim() else [t.grad.item()] for t in leaves])
    got = O.backward(params, x0, dy)
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_network_gradient_finite_differences():
    """Central differences of sum(dy * y) on parameters of every layer (the
    step is small enough that few ReLU kinks are crossed, so the bound is
    loose next to the autograd comparison above)."""
    params = _params(21)
    rng = np.random.default_rng(22)
    x0 = rng.uniform(0, 1, (1, 36, 36, 36))
    dy = rng.standard_normal((1, 32, 32, 32))
    g = O.backward(params, x0, dy)
    offs = np.cumsum([0] + [cout * cin * 27 + cout for cin, cout, _, _ in O.LAYERS])
    picks = [int(offs[l] + rng.integers(0, O.LAYERS[l][0] * O.LAYERS[l][1] * 27)) for l in
             (0, 4, 6, 9)] + [int(offs[3] - 1), params.size - 3, params.size - 1]
    for i in picks:
        h = 1e-6 * max(1.0, abs(params[i]))
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fp = (O.forward(pp, x0)[0] * dy).sum()
        fm = (O.forward(pm, x0)[0] * dy).sum()
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-4 * max(1.0, abs(g[i])), (i, fd, g[i])
