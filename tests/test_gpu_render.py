"""GPU parity of the field consumers (interpolation of a channel level and
the ray-march renderer, render.cpp:19-100) against the CPU oracle, through
the C ABI. Both run the same fixed operation order, so interpolated values,
gradients and rendered pixels must be bit-identical."""
import sys

import numpy as np
import pytest

from oracle import liboracle as O
from tests.helpers import oracle_pipeline, small_scene
from tests.test_render_oracle import sphere_level

pytestmark = pytest.mark.gpu
M = pytest.importorskip("msbg_amd")

CH = 40
DEMO_CAM = dict(camPos=(.5, .8, .7), lookAt=(.5, .5, .5), zoom=1.0, sunPos=(.5, 5, 5),
                color=(0.8, 0.6, 0.4))


def note(*a):
    print("[render]", *a, flush=True)
    sys.stdout.flush()


def ocam(camPos, lookAt, zoom, sunPos, color):
    return O.camera(camPos, lookAt, zoom, sunPos, color)


def _random_grid(storage, S=64, bsx=16, seed=0, nLevels=2):
    """A channel with value, FULL, EMPTY and NULL blocks of random values."""
    rng = np.random.default_rng(seed)
    nb = S // bsx
    g = M.MultiresSparseGrid.create("ip", S, S, S, bsx, nLevels=nLevels)
    g.setRefinementMap(np.zeros(nb ** 3, np.int32))
    kind = rng.choice([3, 2, 1, 0], nb ** 3, p=[0.55, 0.15, 0.15, 0.15]).astype(np.uint8)
    if storage == M.U16:
        empty, full = 1234.0, 65535.0
    else:
        empty, full = 0.03, 1.0
    g.channel_define(CH, storage, 0, kind, empty, full, 0.0)
    pm, npg = g.pagemap(CH)
    vals = rng.random((npg, bsx, bsx, bsx)).astype(np.float32)
    raw = {M.U16: lambda v: np.rint(v * 65535).astype(np.uint16),
           M.F32: lambda v: v,
           M.F16: lambda v: v.astype(np.float16),
           M.BF16: lambda v: np.array([O.load().or_bf16_from_float(float(x)) for x in v.ravel()],
                                      np.uint16).reshape(v.shape),
           M.F64: lambda v: v.astype(np.float64)}[storage](vals)
    g.upload(CH, 0, raw)
    ost = {M.U16: O.U16, M.F32: O.F32, M.F16: O.F16, M.BF16: O.BF16, M.F64: O.F64}[storage]
    pages = raw.view(np.uint16) if storage == M.F16 else raw
    L = O.Level(S, S, S, bsx, pm, pages, ost, empty, full)
    return g, L


def _points(S, n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-2.0, S + 2.0, (n, 3)).astype(np.float32)
    q = rng.integers(0, S, (n // 4, 3)).astype(np.float32) + 0.5
    e = rng.integers(0, S // 16 + 1, (n // 4, 3)).astype(np.float32) * 16  # block edges
    return np.concatenate([p, q, e + rng.uniform(-0.6, 0.6, e.shape).astype(np.float32)])


@pytest.mark.parametrize("storage", [M.U16, M.F32, M.F16, M.BF16, M.F64])
def test_interpolate_bitexact(storage):
    note("interpolate", storage)
    g, L = _random_grid(storage, seed=int(storage))
    with g:
        pts = _points(64, 20000, 1 + int(storage))
        got = g.interpolate(CH, pts, M.IP_LINEAR)
        want = O.interp_linear(L, pts)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        gv, gg = g.interpolate(CH, pts, M.IP_BSPLINE3)
        wv, wg = O.interp_bspline3(L, pts)
        assert np.array_equal(gv.view(np.uint32), wv.view(np.uint32))
        assert np.array_equal(gg.view(np.uint32), wg.view(np.uint32))


def test_interpolate_coarse_level():
    note("interpolate level 1")
    S, bsx = 64, 16
    g = M.MultiresSparseGrid.create("ipl", S, S, S, bsx, nLevels=2)
    with g:
        nb = S // bsx
        g.setRefinementMap(np.ones(nb ** 3, np.int32))
        kind = np.full(nb ** 3, 3, np.uint8)
        g.channel_define(CH, M.F32, 1, kind, 0.0, 1.0, 0.0)
        pm, npg = g.pagemap(CH, 1)
        b = bsx // 2
        vals = np.random.default_rng(3).random((npg, b, b, b)).astype(np.float32)
        g.upload(CH, 1, vals)
        L = O.Level(S // 2, S // 2, S // 2, b, pm, vals, O.F32, 0.0, 1.0)
        pts = _points(S // 2, 4000, 9)
        assert np.array_equal(g.interpolate(CH, pts, M.IP_LINEAR, level=1),
                              O.interp_linear(L, pts))


OUTSIDE_CAM = dict(camPos=(1.5, 0.2, -0.4), lookAt=(.62, .41, .47), zoom=2.5, sunPos=(4, 6, -2),
                   color=(0.9, 0.9, 0.9))


@pytest.mark.parametrize("storage,bsx", [(M.F32, 16), (M.U16, 16), (M.U16, 8), (M.F32, 32)])
def test_render_sphere_bitexact(storage, bsx):
    """Includes a camera outside the domain (rays enter through the faces)
    and three block sizes (the kernel advances through EMPTY blocks
    without sampling)."""
    note("render sphere", storage, bsx)
    S = 64
    ost = O.F32 if storage == M.F32 else O.U16
    L = sphere_level(S, bsx, storage=ost)
    with M.MultiresSparseGrid.create("rs", S, S, S, bsx, nLevels=2) as g:
        nb = S // bsx
        g.setRefinementMap(np.zeros(nb ** 3, np.int32))
        kind = np.where(L.page >= 0, 3, 1).astype(np.uint8)
        full = 65535.0 if storage == M.U16 else 1.0
        g.channel_define(CH, storage, 0, kind, 0.0, full, 0.0)
        pm, _ = g.pagemap(CH)
        assert np.array_equal(pm, L.page)
        g.upload(CH, 0, L.pages)
        for W, H, cam in ((160, 90, DEMO_CAM), (120, 80, OUTSIDE_CAM),
                          (97, 61, dict(camPos=(.2, .3, -.4), lookAt=(.6, .4, .5), zoom=1.7,
                                        sunPos=(-3, 2, -1), color=(1.0, 0.9, 0.2)))):
            img = g.render(CH, W, H, **cam)
            want = O.render(L, ocam(**cam), W, H)
            assert np.array_equal(img, want), int((img != want).sum())
            assert ((img & 0xffffff) != 0).sum() > 200
            img2, cnt = g.render_counts(CH, W, H, **cam)
            assert np.array_equal(img2, img)
            # a hit facing away from the sun shades to black
            assert cnt["hits"] >= int(((img & 0xffffff) != 0).sum())
            assert cnt["interp"] <= cnt["samples"] <= cnt["advanced"] + W * H


def test_render_demo_scene_bitexact():
    """The demo's last step: splat + finalize + MCF sweeps, then the 4K
    render with the demo camera (msbg_demo.cpp:122-166, 701-747)."""
    S, bsx = 256, 16
    xyz = small_scene(S, nObj=12, nPts=60000, seed=3)
    note("render demo scene", S, bsx, xyz.shape[0])
    o = oracle_pipeline(S, bsx, xyz, do_sweep=False)
    L = O.Level(S, S, S, bsx, o["pagemap"], o["splat"].copy(), O.U16)
    O.mcf_sweeps(L, o["active"], o["blv"], 4, 0.1, 1)
    with M.MultiresSparseGrid.create("rd", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL,
                                     nLevels=3) as g:
        g.set_arith(M.ARITH_EXACT)
        g.setRefinementMap(o["lv"])
        active = g.alloc_demo(M.CH_UINT16_1, M.U16)
        g.upload(M.CH_UINT16_1, 0, o["splat"])
        g.applyChannelPdeFast(M.CH_UINT16_1, 0, 0, active, M.MCF_8COLOR, 1, 0, 0, 0, 4, 1.0, 0.1)
        assert np.array_equal(g.download(M.CH_UINT16_1), L.pages)
        img, ms = g.render(M.CH_UINT16_1, 3840, 2160, timing=True, **DEMO_CAM)
        note("4K render", ms, "ms")
        want = O.render(L, ocam(**DEMO_CAM), 3840, 2160)
        assert np.array_equal(img, want), int((img != want).sum())
        assert 0.5 < ((img & 0xffffff) != 0).mean() < 0.95


def test_bad_arguments():
    g, L = _random_grid(M.F32, seed=5)
    with g:
        with pytest.raises(M.MsbgError) as e:
            g.interpolate(CH, np.zeros((4, 3), np.float32), kind=2)
        assert e.value.code == M._lib.EINVAL
        with pytest.raises(M.MsbgError) as e:
            g.interpolate(CH + 1, np.zeros((4, 3), np.float32))
        assert e.value.code == M._lib.ESTATE
        with pytest.raises(M.MsbgError) as e:
            g.render(CH, 0, 10)
        assert e.value.code == M._lib.EINVAL
        assert g.interpolate(CH, np.zeros((0, 3), np.float32)).shape == (0,)
