"""CPU pins of the oracle's other applyChannelPdeFast stencils and driver
modes (msbg4.cpp:60-957): closed forms on polynomial fields, mode
invariances, the hard constraints, the Jacobi temporary's NULL blocks and
the reference's fast exponential."""
import math

import numpy as np
import pytest

from oracle import liboracle as O


def _halo(fn, hw=1, hsx=None):
    hsx = hsx or 2 * hw + 8
    z, y, x = np.meshgrid(*(np.arange(hsx, dtype=np.float64) - hw - 3,) * 3, indexing="ij")
    return fn(x, y, z).astype(np.float32), hsx


def _hid(hsx, hw=1, v=(3, 3, 3)):
    return (v[0] + hw) + (v[1] + hw) * hsx + (v[2] + hw) * hsx * hsx


@pytest.mark.parametrize("op", [1, 3, 14, 15, 17])
def test_linear_fields_are_fixed_points(op):
    h, hsx = _halo(lambda x, y, z: 0.25 + 0.01 * x - 0.02 * y + 0.005 * z)
    hid = _hid(hsx)
    got = O.pde_voxel(op, h, hid, 0.1, alpha=1e30)
    assert abs(got - h.reshape(-1)[hid]) < 2e-7


def test_quadratic_field_closed_forms():
    dt = 0.125  # a power of two keeps the closed forms exact in FP32
    h, hsx = _halo(lambda x, y, z: 0.01 * x * x)
    hid = _hid(hsx)
    f0 = float(h.reshape(-1)[hid])  # x = 0: f0 = 0
    assert f0 == 0.0
    # diffusion: mean of the 6 neighbours is f0 + 0.02/6
    assert O.pde_voxel(3, h, hid, dt) == pytest.approx(f0 + dt * (0.02 / 6), rel=1e-6)
    # Jacobi diffusion: (f0 + dt * (6 f0 + 0.02)) / (1 + 6 dt)
    assert O.pde_voxel(17, h, hid, dt) == pytest.approx((f0 + dt * 0.02) / (1 + 6 * dt), rel=1e-6)
    # Perona-Malik with alpha -> inf is plain diffusion of the 6 differences
    assert O.pde_voxel(14, h, hid, dt, alpha=1e30) == pytest.approx(f0 + dt * 0.02, rel=1e-6)
    assert O.pde_voxel(15, h, hid, dt, alpha=1e30) == pytest.approx(
        (f0 + dt * 0.02) / (1 + 6 * dt), rel=1e-6)
    # laplacian flow at x = 1: L = 0.02 (second difference), |grad| = 0.02
    hid1 = _hid(hsx, v=(4, 3, 3))
    h1 = h
    f = float(h1.reshape(-1)[hid1])
    assert f == pytest.approx(0.01)
    assert O.pde_voxel(1, h1, hid1, dt) == pytest.approx(f + dt * 0.02 * 0.02, rel=1e-5)
    # Taubin: odd sweeps add D, even ones subtract taubin * D
    assert O.pde_voxel(1, h1, hid1, dt, taubin=0.5, it=0) == pytest.approx(
        f - 0.5 * dt * 0.02 * 0.02, rel=1e-5)


def test_bilaplacian_closed_forms():
    dt = 0.01
    # the 25-point operator is the squared 7-point laplacian: zero on
    # quadratics, 24 on x^4 (the fourth difference)
    h2, hsx = _halo(lambda x, y, z: 0.001 * (x * x + y * y - 2 * z * z + x * y), hw=2)
    hid = _hid(hsx, hw=2)
    assert O.pde_voxel(2, h2, hid, dt) == pytest.approx(float(h2.reshape(-1)[hid]), abs=1e-7)
    h4, hsx = _halo(lambda x, y, z: 1e-4 * x ** 4, hw=2)
    hid = _hid(hsx, hw=2, v=(4, 3, 3))  # x = 1: grad = 1e-4 * 0.5 * (16 - 0) = 8e-4
    f0 = float(h4.reshape(-1)[hid])
    H, G = 24e-4, 0.5 * 1e-4 * (2 ** 4 - 0)
    assert O.pde_voxel(2, h4, hid, dt) == pytest.approx(f0 - dt * H * G, rel=1e-4)
    # the limiter
    big, _ = _halo(lambda x, y, z: 10.0 * x ** 4, hw=2)
    assert O.pde_voxel(2, big, hid, 1.0) == pytest.approx(float(big.reshape(-1)[hid]) - 0.1)


def test_fast_exp():
    assert O.fast_exp(0.0) == 1.0
    assert O.fast_exp(-100.0) == 0.0 and math.isinf(O.fast_exp(100.0))
    xs = np.linspace(-20, 20, 4001)
    rel = max(abs(O.fast_exp(x) - math.exp(x)) / math.exp(x) for x in xs)
    # the repeated coefficient costs accuracy but keeps a smooth exponential
    assert rel < 1e-3
    # the weight of the exponential edge function at |g| = |alpha| is e^-1
    h, hsx = _halo(lambda x, y, z: np.where(x > 0, 0.5, 0.0) + 0 * y)
    hid = _hid(hsx)
    a = -0.5
    got = O.pde_voxel(14, h, hid, 0.1, alpha=a)
    assert got == pytest.approx(0.0 + 0.1 * O.fast_exp(-1.0) * 0.5, rel=1e-6)


def _level(S, bsx, vals, pm, storage=O.F32):
    return O.Level(S, S, S, bsx, pm, vals.copy(), storage, 0.0, 1.0)


def _field(S, bsx, seed=1):
    nb = S // bsx
    rng = np.random.default_rng(seed)
    pm = np.where(rng.random(nb ** 3) < 0.7, 0, O.EMPTY).astype(np.int32)
    pm[pm == 0] = np.arange((pm == 0).sum())
    vals = rng.random(((pm >= 0).sum(), bsx, bsx, bsx)).astype(np.float32)
    return pm, vals


@pytest.mark.parametrize("op", [3, 14, 17])
def test_single_block_modes_agree(op):
    # one value block surrounded by EMPTY: every mode performs the same updates
    S, bsx = 32, 16
    pm = np.full(8, O.EMPTY, np.int32)
    pm[0] = 0
    vals = np.random.default_rng(2).random((1, bsx, bsx, bsx)).astype(np.float32)
    outs = []
    for mode in (O.MODE_COLOR8, O.MODE_REDBLACK, O.MODE_JACOBI):
        L = _level(S, bsx, vals, pm)
        O.pde_sweeps(L, [0], None, 4, op, mode, 0.1, alpha=0.3)
        outs.append(L.pages)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_hard_constraints_pin_the_field():
    S, bsx = 32, 16
    pm, vals = _field(S, bsx)
    L = _level(S, bsx, vals, pm)
    O.pde_sweeps(L, np.nonzero(pm >= 0)[0], None, 4, 3, O.MODE_COLOR8, 0.2, densThr=0.5)
    assert np.array_equal(L.pages, vals)  # tol1 = tol2 = 0: phi0 <= phi <= phi0
    L = _level(S, bsx, vals, pm)
    O.pde_sweeps(L, np.nonzero(pm >= 0)[0], None, 4, 3, O.MODE_COLOR8, 0.2, densThr=0.5,
                 densThrTol1=0.01, densThrTol2=0.02)
    d = L.pages - vals
    assert d.min() >= -0.01 - 1e-7 and d.max() <= 0.02 + 1e-7 and np.abs(d).max() > 0


def test_jacobi_odd_count_copies_last():
    S, bsx = 32, 16
    pm, vals = _field(S, bsx, seed=3)
    bl = np.nonzero(pm >= 0)[0]
    a, b = _level(S, bsx, vals, pm), _level(S, bsx, vals, pm)
    O.pde_sweeps(a, bl, None, 3, 17, O.MODE_JACOBI, 0.1)
    O.pde_sweeps(b, bl, None, 4, 17, O.MODE_JACOBI, 0.1)
    assert not np.array_equal(a.pages, b.pages)
    # 3 updates then a copy: equal to 3 updates of a run that ends in place
    c = _level(S, bsx, vals, pm)
    r3 = O.pde_sweeps(c, bl, None, 3, 17, O.MODE_JACOBI, 0.1, residual=True)
    assert r3[-1] == 0.0 and r3[:3].min() > 0


def test_jacobi_temporary_has_null_unswept_blocks():
    # two value blocks side by side, only the first is swept: the first sweep
    # reads the second block's values across the face, the second sweep (from
    # the temporary, where that block is NULL) reads the Neumann clamp
    S, bsx = 32, 16
    pm = np.full(8, O.EMPTY, np.int32)
    pm[0], pm[1] = 0, 1
    vals = np.zeros((2, bsx, bsx, bsx), np.float32)
    vals[1] = 1.0
    L = _level(S, bsx, vals, pm)
    O.pde_sweeps(L, [0], None, 2, 17, O.MODE_JACOBI, 0.1)
    # by hand: x = bsx-1 face voxel of block 0, interior in y, z
    v1 = (0.0 + 0.1 * (0 + 1.0 + 0 + 0 + 0 + 0)) / (1 + 0.6)          # sweep 0, every face voxel
    # sweep 1: the x-1 neighbour stayed 0, the 4 in-face neighbours hold v1,
    # the x+1 neighbour is NULL in the temporary -> the voxel itself (v1);
    # reading block 1 instead would give (v1 + 0.1 * (1 + 4 v1)) / 1.6
    v2 = (v1 + 0.1 * (0.0 + v1 + v1 + v1 + v1 + v1)) / (1 + 0.6)
    got = L.pages[0, 8, 8, bsx - 1]
    assert got == pytest.approx(v2, rel=1e-6)
    assert np.array_equal(L.pages[1], vals[1])


@pytest.mark.parametrize("storage", [O.F32, O.U16])
def test_op4_colour_equals_mcf_sweeps(storage):
    S, bsx = 32, 16
    pm, vals = _field(S, bsx, seed=9)
    if storage == O.U16:
        vals = np.rint(vals * 65535).astype(np.uint16)
    bl = np.nonzero(pm >= 0)[0]
    a = O.Level(S, S, S, bsx, pm, vals.copy(), storage, 0.0)
    b = O.Level(S, S, S, bsx, pm, vals.copy(), storage, 0.0)
    ra = O.pde_sweeps(a, bl, None, 3, 4, O.MODE_COLOR8, 0.1, residual=True)
    rb = O.mcf_sweeps(b, bl, None, 3, 0.1, residual=True)
    assert np.array_equal(a.pages, b.pages) and np.array_equal(ra, rb)
