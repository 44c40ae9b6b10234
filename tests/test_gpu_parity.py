"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Integer/index outputs (activation counts, levels, flags, colour lists, halo
index tables, UNORM16 splat/finalize) must be bit-exact. The sweep in
MSBG_ARITH_EXACT mode must be bit-exact for every storage; the fast mode is
checked against the FP32 oracle with the tolerance stated in each test.
"""
import sys

import numpy as np
import pytest

from oracle import liboracle as O
from tests.helpers import RSCAN, dense_halo, edge_points, gpu_pipeline, oracle_pipeline, small_scene

pytestmark = pytest.mark.gpu

M = pytest.importorskip("msbg_amd")


def note(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _scene(S, seed=7, nObj=5, nPts=3000):
    return np.concatenate([small_scene(S, nObj=nObj, nPts=nPts, seed=seed), edge_points(S, 300, seed)])


# --------------------------------------------------------------------------
# block tables
# --------------------------------------------------------------------------
@pytest.mark.parametrize("S,bsx,nLevels", [(128, 16, 3), (128, 32, 4), (96, 8, 3), (64, 16, 2)])
def test_tables_bitexact(S, bsx, nLevels):
    note("tables", S, bsx, nLevels)
    xyz = _scene(S)
    o = oracle_pipeline(S, bsx, xyz, nLevels=nLevels, do_sweep=False)
    # the splat needs 2*(rScan+0.5) <= bsx (msbg_demo.cpp lock-free scheme)
    gp = gpu_pipeline(S, bsx, xyz, nLevels=nLevels, do_sweep=False, do_splat=bsx >= 9)
    assert gp["nk"] == int(o["keep"].sum())
    assert np.array_equal(gp["grid"].points_download(), o["xk"])
    assert np.array_equal(gp["cnt"], o["cnt"])
    assert np.array_equal(gp["lv0"], o["lv0"])
    assert np.array_equal(gp["lv"], o["lv"])
    assert np.array_equal(gp["blv"], o["blv"])
    ofl = o["bfl"].copy()
    # the demo's allocation loop then sets BLK_EXISTS for level-0 blocks
    _, gfl = gp["grid"].getBlockInfo()
    ex = np.uint16(M._lib.BLK_EXISTS)
    ofl = np.where(o["blv"] == 0, ofl | ex, ofl & ~ex).astype(np.uint16)
    assert np.array_equal(gfl, ofl)
    assert np.array_equal(gp["active"], o["active"])


def test_points_upload_filter_edges():
    """The device-side domain filter (isInDomainRange, then trunc(pos)
    inRange, msbg_demo.cpp:356-360) against the oracle's, on NaN, infinities,
    values next to both ends of the range and the domain's own faces; kept
    points stay in upload order."""
    note("upload filter edges")
    S = 64
    rng = np.random.default_rng(21)
    f = np.float32
    special = np.array([0.0, -0.0, 1e-30, -1e-30, 0.5, np.nextafter(f(0.5), f(0)), S - 1.0,
                        S - 0.5, np.nextafter(f(S - 0.5), f(S)), np.nextafter(f(S), f(0)), S,
                        S + 0.5, -0.5, np.nan, np.inf, -np.inf, 1e30, -1e30], np.float32)
    pts = rng.uniform(-4, S + 4, (4000, 3)).astype(np.float32)
    for k in range(3):  # every special value on every axis, the others inside
        blk = np.full((special.size, 3), S / 2, np.float32)
        blk[:, k] = special
        pts = np.concatenate([pts, blk])
    pts = pts[rng.permutation(pts.shape[0])]
    with M.MultiresSparseGrid.create("f", S, S, S, 16, options=M.OPT_SINGLE_LEVEL, nLevels=2) as g:
        nk = g.points_upload(pts)
        got = g.points_download()
    keep = O.domain_filter(S, S, S, pts) == 1
    assert nk == int(keep.sum()) and 0 < nk < pts.shape[0]
    assert np.array_equal(got.view(np.uint32), pts[keep].view(np.uint32))


@pytest.mark.parametrize("storage", [M.U16, M.F16])
def test_empty_point_sets(storage):
    """No particles (none uploaded, or all outside the domain): activation
    counts nothing, every band leaves the coarsest level, the splat leaves
    the codes at 65535 (UNORM 1.0), and a later non-empty upload replaces
    the cached cell sort."""
    note("empty point sets", storage)
    S, bsx = 64, 16
    nb = S // bsx
    with M.MultiresSparseGrid.create("e", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL, nLevels=2) as g:
        kind = np.full(nb ** 3, 3, np.uint8)
        g.channel_define(M.CH_UINT16_1, storage, 0, kind, 0.0, None,
                         65535.0 if storage == M.U16 else 0.0)
        for pts in (np.zeros((0, 3), np.float32), np.full((10, 3), -5.0, np.float32)):
            assert g.points_upload(pts) == 0
            assert g.points_download().shape == (0, 3)
            assert not g.activate(4.0).any()
            g.splat(M.CH_UINT16_1, 2.0, 2.0)
            raw = g.download(M.CH_UINT16_1).view(np.uint16)
            assert (raw == 65535).all()
        # a non-empty set after the empty ones: the splat sees it
        pts = np.array([[20.5, 20.5, 20.5]], np.float32)
        assert g.points_upload(pts) == 1
        assert g.activate(4.0).sum() > 0
        g.splat(M.CH_UINT16_1, 2.0, 2.0)
        raw = g.download(M.CH_UINT16_1).view(np.uint16)
        pm, npg = g.pagemap(M.CH_UINT16_1, 0)
        L = O.Level(S, S, S, bsx, pm, np.full((npg, bsx, bsx, bsx), 65535, np.uint16), O.U16)
        O.splat_u16(L, pts, 4.0)
        assert np.array_equal(raw.reshape(L.pages.shape), L.pages)
        assert (raw < 65535).sum() > 100


def test_regularize_multilevel_bitexact():
    note("regularize nLevels=4")
    nb = 16
    rng = np.random.default_rng(3)
    m = np.where(rng.random(nb ** 3) < 0.03, 0, 3).astype(np.int32)
    m[rng.random(nb ** 3) < 0.02] = 2
    with M.MultiresSparseGrid.create("r", nb * 16, nb * 16, nb * 16, 16, nLevels=4) as g:
        got = g.regularizeRefinementMap(m)
        want, rc = O.regularize(nb, nb, nb, 4, m)
        assert np.array_equal(got, want)
        flags_in = (rng.random(nb ** 3) < 0.5).astype(np.uint32) * (1 << 12) | 1
        g.setRefinementMap(got, flags_in)
        lv, fl = g.getBlockInfo()
        wl, wf = O.block_flags(nb, nb, nb, 4, want, flags_in)
        assert np.array_equal(lv, wl) and np.array_equal(fl, wf)
        for level in range(4):
            fc = g.blocksFineCoarse(level)
            want_fc = np.nonzero((wl == level) & ((wf & (1 << 3)) != 0))[0]
            assert np.array_equal(fc, want_fc)


def test_halo_table_matches_pagemap():
    note("halo table")
    S, bsx = 128, 16
    gp = gpu_pipeline(S, bsx, _scene(S), do_sweep=False)
    g = gp["grid"]
    pm, npg = g.pagemap(M.CH_UINT16_1)
    tab = g.halo_table(M.CH_UINT16_1)
    nb = S // bsx
    pages_bid = np.nonzero(pm >= 0)[0]
    assert np.array_equal(pm[pages_bid], np.arange(npg))
    for p, bid in enumerate(pages_bid):
        bz, rem = divmod(bid, nb * nb)
        by, bx = divmod(rem, nb)
        for k in range(27):
            dx, dy, dz = k % 3 - 1, (k // 3) % 3 - 1, k // 9 - 1
            x, y, z = bx + dx, by + dy, bz + dz
            want = pm[(z * nb + y) * nb + x] if 0 <= min(x, y, z) and max(x, y, z) < nb else -4
            assert tab[p, k] == want


# --------------------------------------------------------------------------
# splat / finalize
# --------------------------------------------------------------------------
# 8^3 blocks are left out: the splat's footprint, 2 (rScan + 0.5) = 9 voxels,
# is wider than such a block, which the reference's splat does not allow
# (msbg_demo.cpp:541-636); the argument check rejects it
@pytest.mark.parametrize("S,bsx", [(128, 16), (128, 32)])
def test_splat_finalize_u16_bitexact(S, bsx):
    note("splat", S, bsx)
    assert 2 * (RSCAN + 0.5) <= bsx
    xyz = _scene(S, seed=S + bsx, nObj=8, nPts=4000)
    o = oracle_pipeline(S, bsx, xyz, do_sweep=False)
    gp = gpu_pipeline(S, bsx, xyz, do_sweep=False)
    assert np.array_equal(gp["splat"], o["splat"])


def test_splat_abandon_quirk_gpu():
    note("splat goto")
    S, bsx = 32, 16
    with M.MultiresSparseGrid.create("g", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL, nLevels=2) as g:
        kind = np.full(8, 1, np.uint8)
        kind[0] = 3
        g.channel_define(M.CH_UINT16_1, M.U16, 0, kind, 0.0, 65535.0, 65535.0)
        pts = np.array([[14.5, 10.5, 10.5], [5.25, 5.5, 12.75], [9.5, 14.25, 2.0]], np.float32)
        g.points_upload(pts)
        g.splat(M.CH_UINT16_1, 2.0, 2.0)
        got = g.download(M.CH_UINT16_1)
        pm = np.full(8, O.EMPTY, np.int32)
        pm[0] = 0
        L = O.Level(S, S, S, bsx, pm, np.full((1, bsx, bsx, bsx), 65535, np.uint16), O.U16)
        O.splat_u16(L, pts, 4.0)
        assert np.array_equal(got, L.pages)


# The binned LDS splat (tile halos 1 to 4: rScan below 4.5) and the
# per-particle atomic one (wider footprints) against the oracle's
# colour-ordered splat: particles up to the domain's faces and integer
# positions, value/EMPTY/FULL/NULL blocks mixed (the early abandon), the code
# in every storage (the low 32 bits of FP32/FP64).
@pytest.mark.parametrize("S,bsx,rscan,storage", [
    (64, 16, 4.0, M.U16), (48, 8, 1.5, M.U16), (64, 32, 4.25, M.F32), (64, 16, 6.0, M.U16),
    (96, 32, 3.0, M.F64), (64, 16, 2.0, M.BF16), (64, 32, 4.5, M.F16), (128, 64, 4.0, M.U16),
    (48, 8, 1.0, M.F32)])
def test_splat_paths_bitexact(S, bsx, rscan, storage):
    note("splat paths", S, bsx, rscan, storage)
    nb = S // bsx
    rng = np.random.default_rng(int(S + 10 * rscan + storage))
    kind = rng.choice([3, 3, 3, 3, 1, 2, 0], nb ** 3).astype(np.uint8)
    c = rng.uniform(0, S, (12, 3))
    pts = np.concatenate([
        c[rng.integers(0, 12, 30000)] + rng.normal(0, 3.0, (30000, 3)),
        rng.uniform(-7, S + 7, (5000, 3)),
        np.floor(rng.uniform(-2, S + 2, (2000, 3)))]).astype(np.float32)
    with M.MultiresSparseGrid.create("s", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL, nLevels=2) as g:
        g.channel_define(M.CH_UINT16_1, storage, 0, kind, 0.0, None, 65535.0 if storage == M.U16
                         else 0.0)
        pm, npg = g.pagemap(M.CH_UINT16_1, 0)
        assert npg > 0
        nk = g.points_upload(pts)
        g.splat(M.CH_UINT16_1, rscan / 2, rscan / 2)
        got = g.download(M.CH_UINT16_1)
    # the upload keeps the points of the domain range (msbg_demo.cpp:356-360)
    pts = np.ascontiguousarray(pts[O.domain_filter(S, S, S, pts) == 1])
    assert nk == pts.shape[0] and nk > 20000
    if storage in (M.F32, M.F64):
        got = got.view(np.uint32).reshape(-1, 2 if storage == M.F64 else 1)[:, 0]
    else:
        got = got.view(np.uint16)
    L = O.Level(S, S, S, bsx, pm, np.full((npg, bsx, bsx, bsx), 65535, np.uint16), O.U16)
    O.splat_u16(L, pts, rscan)
    want = L.pages.reshape(-1)
    assert np.array_equal(got.reshape(-1), want)
    assert (want < 65535).sum() > 1000


@pytest.mark.parametrize("storage", [M.U16, M.F32, M.F16, M.BF16, M.F64])
def test_finalize_exhaustive(storage):
    note("finalize exhaustive", storage)
    S, bsx = 64, 32  # 2 value pages = 65536 voxels
    with M.MultiresSparseGrid.create("f", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL, nLevels=2) as g:
        kind = np.full(8, 1, np.uint8)
        kind[:2] = 3
        g.channel_define(M.CH_UINT16_1, storage, 0, kind, 0.0, None, 0.0)
        codes = np.arange(65536, dtype=np.uint16)
        if storage in (M.U16,):
            g.upload(M.CH_UINT16_1, 0, codes.reshape(2, bsx, bsx, bsx))
        elif storage in (M.F16, M.BF16):
            g.upload(M.CH_UINT16_1, 0, codes.view(np.float16 if storage == M.F16 else np.uint16)
                     .reshape(2, bsx, bsx, bsx))
        elif storage == M.F32:
            g.upload(M.CH_UINT16_1, 0, codes.astype(np.uint32).view(np.float32).reshape(2, bsx, bsx, bsx))
        else:
            g.upload(M.CH_UINT16_1, 0, codes.astype(np.uint64).view(np.float64).reshape(2, bsx, bsx, bsx))
        g.finalize(M.CH_UINT16_1, 2.0, 2.0)
        got = g.download(M.CH_UINT16_1).reshape(-1)
        if storage == M.U16:
            assert np.array_equal(got, O.finalize_table())
        else:
            f = O.finalize_density_table()
            if storage == M.F32:
                assert np.array_equal(got.view(np.uint32), f.view(np.uint32))
            elif storage == M.F64:
                assert np.array_equal(got, f.astype(np.float64))
            elif storage == M.F16:
                assert np.array_equal(got.view(np.uint16), f.astype(np.float16).view(np.uint16))
            else:
                Lb = O.load()
                want = np.array([Lb.or_bf16_from_float(float(v)) for v in f], np.uint16)
                assert np.array_equal(got.view(np.uint16), want)


# --------------------------------------------------------------------------
# sweep
# --------------------------------------------------------------------------
def _init_field(o, S, bsx, storage):
    """The oracle's finalized UNORM16 field converted to `storage`."""
    vals = o["splat"].astype(np.float32) * np.float32(1.0 / 65535.0)
    if storage == O.U16:
        return o["splat"].copy(), O.Level(S, S, S, bsx, o["pagemap"], o["splat"].copy(), O.U16)
    from tests.helpers import to_storage_level
    L = to_storage_level(S, bsx, o["pagemap"], vals, storage)
    return L.pages.copy(), L


def _np_view(pages, storage):
    if storage == M.F16:
        return pages.view(np.float16)
    return pages


@pytest.mark.parametrize("storage", [M.U16, M.F32, M.F16, M.BF16, M.F64, M.U8])
@pytest.mark.parametrize("S,bsx", [(128, 16), (128, 32), (64, 8)])
def test_sweep_exact_bitexact(S, bsx, storage):
    note("sweep exact", S, bsx, storage)
    xyz = _scene(S, seed=11 + bsx, nObj=6, nPts=3000)
    o = oracle_pipeline(S, bsx, xyz, do_sweep=False)
    init, L = _init_field(o, S, bsx, storage)
    res_o = O.mcf_sweeps(L, o["active"], o["blv"], 3, 0.1, 1, residual=True)
    with M.MultiresSparseGrid.create("s", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL, nLevels=3) as g:
        g.set_arith(M.ARITH_EXACT)
        g.setRefinementMap(o["lv"])
        active = g.alloc_demo(M.CH_UINT16_1, storage)
        assert np.array_equal(active, o["active"])
        g.upload(M.CH_UINT16_1, 0, _np_view(init, storage))
        res_g = g.applyChannelPdeFast(M.CH_UINT16_1, 0, 0, active, M.MCF_8COLOR, 1, 0, 0, 0, 3, 1.0,
                                      0.1, residual=True)
        got = g.download(M.CH_UINT16_1)
        want = L.pages
        if storage == M.F16:
            got = got.view(np.uint16)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8).reshape(got.view(np.uint8).shape))
        assert res_g.shape == res_o.shape == (4,)
        assert np.allclose(res_g, res_o, rtol=1e-12, atol=0)


@pytest.mark.parametrize("S,bsx", [(128, 16), (128, 32)])
def test_sweep_fast_close_to_exact(S, bsx):
    """FMA + reciprocal divide: within 2e-5 of the strict FP32 oracle on
    all but a tiny fraction of voxels (threshold flips near |grad|^2=1e-7)."""
    note("sweep fast", S, bsx)
    xyz = _scene(S, seed=21, nObj=6, nPts=3000)
    o = oracle_pipeline(S, bsx, xyz, do_sweep=False)
    init, L = _init_field(o, S, bsx, O.F32)
    O.mcf_sweeps(L, o["active"], o["blv"], 4, 0.1, 1)
    with M.MultiresSparseGrid.create("s", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL, nLevels=3) as g:
        g.set_arith(M.ARITH_FAST)
        g.setRefinementMap(o["lv"])
        active = g.alloc_demo(M.CH_FLOAT_1, M.F32)
        g.upload(M.CH_FLOAT_1, 0, init)
        g.applyChannelPdeFast(M.CH_FLOAT_1, 0, 0, active, M.MCF_8COLOR, 1, 0, 0, 0, 4, 1.0, 0.1)
        got = g.download(M.CH_FLOAT_1)
    err = np.abs(got.astype(np.float64) - L.pages)
    note("fast max err", err.max(), "frac>2e-5", (err > 2e-5).mean())
    assert (err > 2e-5).mean() < 1e-4
    assert np.sqrt((err ** 2).mean()) < 1e-6


# max |fast - exact| after 2 sweeps, per storage: FP32 2e-5 (2e-5 of the
# field's [0,1] range), UNORM16 2 LSB, FP16 and BF16 two units in the last
# place at 1.0 (storage rounding of the fast result may land one step away,
# and one neighbour's step can move a voxel by another)
_FAST_MAX = {M.F32: 2e-5, M.U16: 2.0 / 65535.0, M.F16: 2.0 ** -9, M.BF16: 2.0 ** -6,
             M.U8: 2.0 / 255.0}


# the unclamped mode on the float storages only: UNORM values outside [0,1]
# wrap in the integer store
@pytest.mark.parametrize("storage,doC01", [(M.U16, 1), (M.F32, 1), (M.F16, 1), (M.BF16, 1),
                                           (M.U8, 1), (M.F32, 0), (M.F16, 0), (M.BF16, 0)])
@pytest.mark.parametrize("S,bsx", [(128, 16), (128, 32)])
def test_sweep_fast_every_column_bounded(S, bsx, storage, doC01):
    """Every voxel-column variant of the fast kernel (2x2, 4x2, 4x4), with and
    without the [0,1] clamp, against the strict oracle after 2 sweeps: a max
    bound per storage (above) and an RMS bound. UNORM16 runs with the clamp
    only (unclamped values outside [0,1] wrap in the 16-bit store)."""
    xyz = _scene(S, seed=31 + bsx, nObj=6, nPts=3000)
    o = oracle_pipeline(S, bsx, xyz, do_sweep=False)
    init, L = _init_field(o, S, bsx, storage)
    O.mcf_sweeps(L, o["active"], o["blv"], 2, 0.1, doC01)
    want = L.pages
    if storage == M.U16:
        want = want.astype(np.float64) / 65535.0
    elif storage == M.U8:
        want = want.astype(np.float64) / 255.0
    elif storage == M.F16:
        want = want.view(np.float16).astype(np.float64)
    elif storage == M.BF16:
        want = (want.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    else:
        want = want.astype(np.float64)
    for tile in (0, 1, 2):
        note("sweep fast", S, bsx, storage, "tile", tile, "c01", doC01)
        with M.MultiresSparseGrid.create("s", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL,
                                         nLevels=3) as g:
            g.set_arith(M.ARITH_FAST)
            g.set_tile(tile)
            g.setRefinementMap(o["lv"])
            active = g.alloc_demo(M.CH_UINT16_1, storage)
            g.upload(M.CH_UINT16_1, 0, _np_view(init, storage))
            g.applyChannelPdeFast(M.CH_UINT16_1, 0, 0, active, M.MCF_8COLOR, doC01, 0, 0, 0, 2,
                                  1.0, 0.1)
            got = g.download(M.CH_UINT16_1)
        if storage == M.U16:
            got = got.astype(np.float64) / 65535.0
        elif storage == M.U8:
            got = got.astype(np.float64) / 255.0
        elif storage == M.F16:
            got = got.view(np.float16).astype(np.float64)
        elif storage == M.BF16:
            got = (got.view(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        else:
            got = got.astype(np.float64)
        d = np.abs(got.reshape(want.shape) - want)
        note("max", d.max(), "rms", np.sqrt((d ** 2).mean()))
        assert d.max() <= _FAST_MAX[storage], (tile, d.max())
        assert np.sqrt((d ** 2).mean()) <= _FAST_MAX[storage] / 8


def test_sweep_u8_stochastic_store_two_calls():
    """applyChannelPdeFast<uint8_t> (msbg4.cpp:1008-1030) stores UNORM8 with
    stochastic rounding (sbg.h:324-340). Exact arithmetic over two calls in a
    row, bit for bit against the oracle: the grid's sweep step carries over
    from call to call like the oracle level's. The entry points that have no
    UNORM8 form say so."""
    S, bsx = 128, 16
    xyz = _scene(S, seed=41, nObj=6, nPts=3000)
    o = oracle_pipeline(S, bsx, xyz, do_sweep=False)
    init, L = _init_field(o, S, bsx, O.U8)
    with M.MultiresSparseGrid.create("u8", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL,
                                     nLevels=3) as g:
        g.set_arith(M.ARITH_EXACT)
        g.setRefinementMap(o["lv"])
        active = g.alloc_demo(M.CH_UINT8_1, M.U8)
        g.upload(M.CH_UINT8_1, 0, init)
        for call in range(2):
            note("u8 call", call)
            r_o = O.mcf_sweeps(L, o["active"], o["blv"], 2, 0.1, 1, residual=True)
            r_g = g.applyChannelPdeFast(M.CH_UINT8_1, 0, 0, active, M.MCF_8COLOR, 1, 0, 0, 0, 2,
                                        1.0, 0.1, residual=True)
            assert np.array_equal(g.download(M.CH_UINT8_1), L.pages.reshape(-1, bsx, bsx, bsx))
            np.testing.assert_allclose(r_g, r_o, rtol=1e-12)
        with pytest.raises(M.MsbgError):
            g.splat(M.CH_UINT8_1, 2.0, 2.0)
        with pytest.raises(M.MsbgError):
            g.finalize(M.CH_UINT8_1, 2.0, 2.0)


def test_sweep_skips_non_level0_and_rejects_bad_lists():
    note("sweep list semantics")
    S, bsx = 128, 16
    xyz = small_scene(S, nObj=2, nPts=2000, seed=5, rmin=8, rmax=12)
    gp = gpu_pipeline(S, bsx, xyz, do_sweep=False)
    g = gp["grid"]
    before = g.download(M.CH_UINT16_1)
    lv, _ = g.getBlockInfo()
    empty_bid = int(np.nonzero(lv != 0)[0][0])
    with pytest.raises(M.MsbgError):
        g.applyChannelPdeFast(M.CH_UINT16_1, 0, 0, np.array([empty_bid]), M.MCF_8COLOR, 1, 0, 0, 0,
                              2, 1.0, 0.1)
    with pytest.raises(M.MsbgError):  # laplTyp 10 (redistancing) is not built
        g.applyChannelPdeFast(M.CH_UINT16_1, 0, 0, None, -(10 | M.OPT_8_COLOR_SCHEME), 1, 0, 0, 0,
                              2, 1.0, 0.1)
    g.applyChannelPdeFast(M.CH_UINT16_1, 0, 0, np.zeros(0, np.int32), M.MCF_8COLOR, 1, 0, 0, 0, 2,
                          1.0, 0.1)
    assert np.array_equal(g.download(M.CH_UINT16_1), before)


def test_sweep_deterministic_and_constant_field():
    note("determinism")
    S, bsx = 256, 32
    xyz = _scene(S, seed=2, nObj=10, nPts=8000)
    a = gpu_pipeline(S, bsx, xyz, storage=M.F16, numIter=6, arith=M.ARITH_FAST)["pages"]
    b = gpu_pipeline(S, bsx, xyz, storage=M.F16, numIter=6, arith=M.ARITH_FAST)["pages"]
    assert np.array_equal(a.view(np.uint16), b.view(np.uint16))
    with M.MultiresSparseGrid.create("c", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL, nLevels=2) as g:
        kind = np.full(g.nBlocks(), 3, np.uint8)
        g.channel_define(M.CH_FLOAT_1, M.F32, 0, kind, 0.0, 1.0, 0.375)
        g.setRefinementMap(np.zeros(g.nBlocks(), np.int32))
        g.applyChannelPdeFast(M.CH_FLOAT_1, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 4, 1.0, 0.1)
        assert np.all(g.download(M.CH_FLOAT_1) == np.float32(0.375))


@pytest.mark.parametrize("storage", [M.F32, M.F16, M.U16])
def test_sweep_repeatable_every_column_shape(storage):
    """Fast sweeps of one field give the same bits every time, for every
    voxel column (2x2, 4x2, 4x4: four, two and one waves per 32^3 block), so
    that the LDS planes a workgroup's waves share are never overwritten
    before every wave has read them (a race shows as run-to-run differences
    after many sweeps)."""
    S, bsx = 256, 32
    xyz = _scene(S, seed=5, nObj=12, nPts=8000)
    r = gpu_pipeline(S, bsx, xyz, storage=storage, do_sweep=False)
    g, active, splat = r["grid"], r["active"], r["splat"]
    try:
        for tile in (0, 1, 2):
            note("repeatable", storage, tile)
            g.set_tile(tile)
            runs = []
            for _ in range(3):
                g.upload(30, 0, splat)
                g.applyChannelPdeFast(30, 0, 0, active, M.MCF_8COLOR, 1, 0, 0, 0, 40, 1.0, 0.1)
                runs.append(g.download(30).view(np.uint8).copy())
            assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2]), tile
    finally:
        g.destroy()


# --------------------------------------------------------------------------
# halo fill (test hook) and cross-level ghosts
# --------------------------------------------------------------------------
def _multilevel_grid(S=64, bsx=16, seed=1, storage=M.F32):
    nb = S // bsx
    rng = np.random.default_rng(seed)
    m = np.where(rng.random(nb ** 3) < 0.15, 0, 2).astype(np.int32)
    g = M.MultiresSparseGrid.create("m", S, S, S, bsx, nLevels=3)
    lv = g.regularizeRefinementMap(m)
    g.setRefinementMap(lv)
    g.set_exists(None)
    levels, flags = g.getBlockInfo()
    pages = []
    pms = []
    for level in range(3):
        kind = np.where(levels == level, 3, 0).astype(np.uint8)
        if level > 0:
            kind[(levels < level)] = 0
        g.channel_define(M.CH_FLOAT_1, storage, level, kind, 0.125, 0.875, 0.0)
        pm, npg = g.pagemap(M.CH_FLOAT_1, level)
        b = bsx >> level
        data = rng.random((npg, b, b, b)).astype(np.float32)
        if storage == M.F16:  # raw binary16 bits
            data = data.astype(np.float16).view(np.uint16)
        elif storage == M.BF16:  # raw bfloat16 bits (truncated values)
            data = (data.view(np.uint32) >> 16).astype(np.uint16)
        g.upload(M.CH_FLOAT_1, level, data.view(np.float16) if storage == M.F16 else data)
        pages.append(data)
        pms.append(pm)
    return g, levels, flags, pms, pages


def _olevel(S, bsx, level, pm, pages, storage=O.F32):
    return O.Level(S >> level, S >> level, S >> level, bsx >> level, pm, pages, storage, 0.125, 0.875)


@pytest.mark.parametrize("opt", [M.OPT_BC_NEUMANN, M.OPT_BC_COARSE_LEVEL,
                                 M.OPT_BC_COARSE_LEVEL | M.OPT_INTERPOLATE, 0])
def test_halo_fill_hook_bitexact(opt):
    note("halo fill", opt)
    S, bsx = 64, 16
    g, levels, flags, pms, pages = _multilevel_grid(S, bsx)
    with g:
        for level in range(2):
            L = _olevel(S, bsx, level, pms[level], pages[level])
            Lo = _olevel(S, bsx, level + 1, pms[level + 1], pages[level + 1])
            bids = np.nonzero(levels == level)[0][:12]
            for bid in bids:
                got = g.fillHaloBlock(M.CH_FLOAT_1, int(bid), 1, opt)
                want = O.halo_fill(L, int(bid), 1, opt, lo=Lo)
                assert np.array_equal(got, want), (level, bid)
                if opt == M.OPT_BC_NEUMANN:
                    dense = dense_halo(pms[level], pages[level], S >> level, bsx >> level, int(bid),
                                       0.125, 0.875)
                    assert np.array_equal(got, dense)


_GHOST_OST = {M.F32: O.F32, M.F16: O.F16, M.BF16: O.BF16}


# block sizes 8..32 take the LDS-staged kernel, level 1 of 8^3 blocks (4^3)
# too; 64^3 blocks the per-tap one
@pytest.mark.parametrize("bsx,storage,seed", [
    (16, M.F32, 4), (32, M.F16, 5), (8, M.BF16, 6), (32, M.F32, 7), (16, M.F16, 8),
    (64, M.F32, 9)])
def test_prepare_ghost_blocks_bitexact(bsx, storage, seed):
    note("ghost blocks", bsx, storage)
    S = 4 * bsx
    g, levels, flags, pms, pages = _multilevel_grid(S, bsx, seed=seed, storage=storage)
    with g:
        g.prepareGhostBlocks(M.CH_FLOAT_1)
        _, fl = g.getBlockInfo()
        post_pm = [g.pagemap(M.CH_FLOAT_1, l)[0] for l in range(3)]
        post = [g.download(M.CH_FLOAT_1, l) for l in range(3)]
        if storage == M.F16:
            post = [p.view(np.uint16) for p in post]
        olv = []
        for l in range(3):
            b = bsx >> l
            pg = np.zeros((int((post_pm[l] >= 0).sum()), b, b, b), pages[l].dtype)
            pg[:pages[l].shape[0]] = pages[l]
            olv.append(_olevel(S, bsx, l, post_pm[l], pg, _GHOST_OST[storage]))
        fcs = [g.blocksFineCoarse(l) for l in range(3)]
        O.prepare_ghost_blocks(olv, fcs, fl)
        for l in range(3):
            assert np.array_equal(post[l], olv[l].pages), l
        assert sum(len(f) for f in fcs) > 0


# --------------------------------------------------------------------------
# the C++ facade (include/msbg/msbg.hpp) drives the same path
# --------------------------------------------------------------------------
def test_cpp_facade_demo_matches_binding(tmp_path):
    import os
    import subprocess
    note("facade demo")
    exe = os.path.join(os.path.dirname(__file__), "cpp", "facade_demo")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/facade_demo not built (run the build step)")
    out, frame, cnf = tmp_path / "pages.bin", tmp_path / "frame.bin", tmp_path / "cnf.bin"
    u8, pres = tmp_path / "u8.bin", tmp_path / "pressure.bin"
    r = subprocess.run([exe, str(out), str(frame), str(cnf), str(u8), str(pres)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, dtype=np.uint16)
    S, bsx = 128, 16
    xyz = M.scene_spheres(99, 6, 30, 98, 10, 25, 4000, 0.5, 0.3)
    gp = gpu_pipeline(S, bsx, xyz, numIter=5, dt=0.1, arith=M.ARITH_FAST)
    assert np.array_equal(got, gp["pages"].reshape(-1))
    # RDR::RaymarchRenderer through the facade = msbg_render through ctypes
    img = np.fromfile(frame, dtype=np.uint32).reshape(180, 320)
    want = gp["grid"].render(30, 320, 180, camPos=(.5, .8, .7), lookAt=(.5, .5, .5), zoom=1.0,
                             sunPos=(.5, 5, 5), color=(0.8, 0.6, 0.4))
    assert np.array_equal(img, want)
    assert ((img & 0xffffff) != 0).sum() > 1000
    # applyChannelPdeCnf (Taubin laplacian flow on a float copy) through the
    # facade = msbg_apply_channel_pde with lambda/taubin through ctypes
    g = gp["grid"]
    g.copyDataTo(30, M.CH_FLOAT_1, M.F32)
    g.applyChannelPdeFast(M.CH_FLOAT_1, 6, 7, None, -(1 + M.OPT_8_COLOR_SCHEME), 1, 0, 0, 0, 4,
                          1.0, 0.05, taubinStrength=0.5)
    got = np.fromfile(cnf, dtype=np.float32)
    assert np.array_equal(got, g.download(M.CH_FLOAT_1).reshape(-1))
    # applyChannelPdeFast<uint8_t>: the same stochastic store at the same
    # sweep steps of the grid
    g.copyDataTo(30, M.CH_UINT8_1, M.U8)
    g.applyChannelPdeFast(M.CH_UINT8_1, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 4, 1.0, 0.1)
    assert np.array_equal(np.fromfile(u8, dtype=np.uint8), g.download(M.CH_UINT8_1).reshape(-1))
    g.copyDataTo(30, 8, M.F32)  # without Taubin the result differs
    g.applyChannelPdeFast(8, 6, 7, None, -(1 + M.OPT_8_COLOR_SCHEME), 1, 0, 0, 0, 4, 1.0, 0.05)
    assert not np.array_equal(got, g.download(8).reshape(-1))
    # MSBG::PressureGrid (relax, multiplyLaplacianMatrixOpt, V-cycles) through
    # the facade = the Python binding on the same inputs
    from msbg_amd import pressure as PR
    raw = np.fromfile(pres, dtype=np.uint8)
    nb, b0 = (4, 4, 2), 16
    lev = raw[:32].view(np.int8)
    typ = raw[32:32 + 64 * 64 * 32].reshape(32, 64, 64)
    with PR.PressureGrid(nb, b0, lev, typ, dbc=(0, 0, 1, 1, 1, 0), dense_levels=1) as ps:
        n = ps.size(0)
        xr = raw[32 + typ.size:].view(np.float32)
        assert xr.size == 2 * n
        ps.upload(PR.CH_B, 0, np.float32(1e-4) * ((np.arange(n) % 13) - 6).astype(np.float32))
        ps.vcycle(2, 2, 20)
        ps.vcycle(2, 2, 20)
        ps.relax(0, 0, 0, PR.CH_X, PR.CH_B, PR.CH_T, 1)
        ps.multiplyLaplacianMatrixOpt(PR.OPT_CALC_RESIDUAL, 0, PR.CH_X, PR.CH_B, 0.0, PR.CH_R)
        assert np.array_equal(xr[:n].view(np.uint32), ps.download(PR.CH_X, 0).view(np.uint32))
        assert np.array_equal(xr[n:].view(np.uint32), ps.download(PR.CH_R, 0).view(np.uint32))
        assert np.abs(xr[:n]).max() > 0


# --------------------------------------------------------------------------
# storage conversion, FP64 differences and the precision study
# --------------------------------------------------------------------------
def test_channel_copy_and_diff():
    note("channel copy/diff")
    S, bsx = 64, 16
    rng = np.random.default_rng(12)
    with M.MultiresSparseGrid.create("c", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL, nLevels=2) as g:
        kind = np.where(rng.random(g.nBlocks()) < 0.5, 3, 1).astype(np.uint8)
        g.channel_define(1, M.F32, 0, kind, 0.0, 1.0, 0.0)
        vals = rng.random(g.download(1).shape).astype(np.float32)
        g.upload(1, 0, vals)
        g.copyDataTo(1, 2, M.F16)
        g.copyDataTo(1, 3, M.BF16)
        g.copyDataTo(1, 4, M.F64)
        g.copyDataTo(1, 5, M.U16)
        assert np.array_equal(g.download(2).view(np.uint16), vals.astype(np.float16).view(np.uint16))
        Lb = O.load()
        bf = np.array([Lb.or_bf16_from_float(float(v)) for v in vals.ravel()], np.uint16)
        assert np.array_equal(g.download(3).ravel(), bf)
        assert np.array_equal(g.download(4), vals.astype(np.float64))
        assert np.array_equal(g.download(5), np.rint(vals * np.float32(65535)).astype(np.uint16))
        mx, rms = g.diff(2, 4)
        d = vals.astype(np.float16).astype(np.float64) - vals.astype(np.float64)
        assert abs(mx - np.abs(d).max()) < 1e-12 and abs(rms - np.sqrt((d ** 2).mean())) < 1e-12


def test_precision_study_small():
    from msbg_amd.studies import precision_study
    note("precision study 256")
    r = precision_study(S=256, nObj=24, rmin=10, rmax=40, nPts=6000, sweeps=8, every=4)
    last = r["rows"][-1]
    # FP32 storage tracks FP64 closely; 16-bit formats in order of their step
    assert last["f32"]["max"] < 1e-4
    assert last["f32"]["rms"] < last["u16"]["rms"] < last["f16"]["rms"] < last["bf16"]["rms"]
    assert last["bf16"]["max"] < 0.1
    assert all(np.isfinite(last[k]["residual"]) for k in ("f32", "f16", "bf16", "u16"))


# --------------------------------------------------------------------------
# batched sweeps of independent grids (one graph launch)
# --------------------------------------------------------------------------
BCH = 30  # channel gpu_pipeline fills


def _batch_grids(specs, storage=M.F16, seed=0):
    grids = []
    for k, (S, bsx) in enumerate(specs):
        xyz = small_scene(S, nObj=4, nPts=2500, seed=seed + k, rmin=S / 12, rmax=S / 6)
        gp = gpu_pipeline(S, bsx, xyz, storage=storage, do_sweep=False, arith=M.ARITH_FAST)
        grids.append(gp["grid"])
    return grids


def test_batch_matches_single_grid_sweeps():
    note("batch vs single")
    specs = [(64, 16), (128, 32), (96, 16), (128, 16), (64, 32)]
    a = _batch_grids(specs, seed=3)
    b = _batch_grids(specs, seed=3)
    for g in b:
        g.applyChannelPdeFast(BCH, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 4, 1.0, 0.1)
    with M.Batch() as bt:
        for g in a:
            bt.add(g, BCH, 0.1, 1)
        bt.run(4, useGraph=False)
    for ga, gb in zip(a, b):
        assert np.array_equal(ga.download(BCH).view(np.uint16),
                              gb.download(BCH).view(np.uint16))


def test_batch_graph_replays_match_eager():
    """Each graph replay, with eager work and synchronisations between the
    replays, equals the same sweeps run eagerly."""
    note("batch graph replays")
    specs = [(64, 16), (128, 32), (128, 16)]
    a = _batch_grids(specs, seed=9)
    b = _batch_grids(specs, seed=9)
    with M.Batch() as ba, M.Batch() as bb:
        for g in a:
            ba.add(g, BCH, 0.1, 1)
        for g in b:
            bb.add(g, BCH, 0.1, 1)
        for rep in range(3):
            ba.run(2, useGraph=True)
            bb.run(2, useGraph=False)
            for g in a:  # eager work between the replays
                g.synchronize()
                g.download(BCH)
            for ga, gb in zip(a, b):
                assert np.array_equal(ga.download(BCH).view(np.uint16),
                                      gb.download(BCH).view(np.uint16)), rep


def test_batch_cfg4_grids_against_oracle():
    """cfg4's generator at test sizes: 16 grids of 128^3-256^3, 16^3 or 32^3
    blocks, 1-3 levels, 1-20 % active, FP16 and FP32. Exact arithmetic
    through one batched launch per pass as a captured graph and eagerly, and
    through the persistent launch: every grid bitwise equal to the oracle's
    own sweeps (msbg4.cpp:384-900). Fast arithmetic through the batch: within
    the per-storage sweep bound."""
    from msbg_amd.studies import _scene_grid
    rng = np.random.default_rng(2024)
    CH, CHE, CHF, CHP = 50, 51, 52, 53
    grids, want, want_fast, st = [], [], [], []
    try:
        for k in range(16):
            S = int(rng.choice([128, 256]))
            bsx = int(rng.choice([16, 32]))
            nl = int(rng.choice([1, 2, 3]))
            frac = float(rng.uniform(0.01, 0.20))
            nObj = int(rng.integers(8, 129))
            storage = (M.F16, M.F32)[k % 2]
            note("cfg4 grid", k, S, bsx, nl, round(frac, 3), nObj, storage)
            g, _ = _scene_grid(rng, S, bsx, nl, frac, nObj, CH, storage)
            grids.append(g)
            st.append(storage)
            g.copyDataTo(CH, CHE, storage)
            g.copyDataTo(CH, CHF, storage)
            g.copyDataTo(CH, CHP, storage)
            pm, _ = g.pagemap(CH)
            pages = g.download(CH)
            if storage == M.F16:
                pages = pages.view(np.uint16)
            levels, _ = g.getBlockInfo()
            blocks = np.nonzero(pm >= 0)[0].astype(np.int32)
            ost = O.F16 if storage == M.F16 else O.F32
            Lx = O.Level(S, S, S, bsx, pm, pages.copy(), ost)
            O.mcf_sweeps(Lx, blocks, levels, 3, 0.1, 1)  # 4 sweeps
            want.append(Lx.pages)
            Lf = O.Level(S, S, S, bsx, pm, pages.copy(), ost)
            O.mcf_sweeps(Lf, blocks, levels, 2, 0.1, 1)
            want_fast.append(Lf.pages)
        for g in grids:
            g.set_arith(M.ARITH_EXACT)
        with M.Batch() as bg, M.Batch() as be, M.Batch() as bp:
            for g in grids:
                bg.add(g, CH, 0.1, 1)
                be.add(g, CHE, 0.1, 1)
                bp.add(g, CHP, 0.1, 1)
            bg.run(3, useGraph=True)
            be.run(3, useGraph=False)
            bp.run(3, useGraph=M.Batch.PERSISTENT)
        for k, g in enumerate(grids):
            for ch in (CH, CHE, CHP):
                got = g.download(ch)
                if st[k] == M.F16:
                    got = got.view(np.uint16)
                assert np.array_equal(got, want[k]), (k, ch)
        for g in grids:
            g.set_arith(M.ARITH_FAST)
        with M.Batch() as bf:
            for g in grids:
                bf.add(g, CHF, 0.1, 1)
            bf.run(2, useGraph=True)
        for k, g in enumerate(grids):
            got = g.download(CHF).astype(np.float64)
            w = want_fast[k]
            w = w.view(np.float16).astype(np.float64) if st[k] == M.F16 else w.astype(np.float64)
            d = np.abs(got.reshape(w.shape) - w)
            assert d.max() <= _FAST_MAX[st[k]], (k, d.max())
    finally:
        for g in grids:
            g.destroy()


def test_batch_persistent_matches_per_pass_launches():
    """The persistent launch (work queue, per-grid waits) gives the bits of
    one launch per colour pass, in fast arithmetic, for 32^3 and 16^3 blocks
    of FP16 grids, on three replays of its captured graph with eager work
    between them."""
    note("batch persistent")
    specs = [(128, 32), (256, 32), (128, 16), (64, 16), (96, 32)]
    a = _batch_grids(specs, seed=31)
    b = _batch_grids(specs, seed=31)
    with M.Batch() as bq, M.Batch() as bl:
        for g in a:
            bq.add(g, BCH, 0.1, 1)
        for g in b:
            bl.add(g, BCH, 0.1, 1)
        for rep in range(3):
            bq.run(4, useGraph=M.Batch.PERSISTENT)
            bl.run(4, useGraph=False)
            for g in a:
                g.synchronize()
                g.download(BCH)
            for ga, gb in zip(a, b):
                assert np.array_equal(ga.download(BCH).view(np.uint16),
                                      gb.download(BCH).view(np.uint16)), rep


@pytest.mark.parametrize("spec", [(64, 16), (96, 32)])
def test_batch_persistent_single_grid_runs_cross_passes(spec):
    """One small grid alone in a persistent batch: its colour passes hold
    fewer blocks than a workgroup's run of queue items, so runs cross from
    one pass into the next of the same grid (the finished blocks are published
    before the next pass's wait); the bits equal one launch per colour pass."""
    note("batch persistent single grid", spec)
    a = _batch_grids([spec], seed=51)
    b = _batch_grids([spec], seed=51)
    with M.Batch() as bq, M.Batch() as bl:
        bq.add(a[0], BCH, 0.1, 1)
        bl.add(b[0], BCH, 0.1, 1)
        bq.run(6, useGraph=M.Batch.PERSISTENT)
        bl.run(6, useGraph=False)
        a[0].synchronize()
        assert np.array_equal(a[0].download(BCH).view(np.uint16),
                              b[0].download(BCH).view(np.uint16))


def test_batch_persistent_reports_a_wait_past_its_bound():
    """A dependency wait that runs past its poll bound ends the wait and is
    reported as a failed run (forced here with a bound of zero time, which
    lets blocks run before their grid's previous pass: no hang, an error);
    with the default bound the same batch runs correctly again."""
    note("batch persistent wait bound")
    specs = [(128, 32), (128, 32), (128, 16), (128, 16)]
    a = _batch_grids(specs, seed=41)
    b = _batch_grids(specs, seed=41)
    with M.Batch() as bq:
        for g in a:
            bq.add(g, BCH, 0.1, 1)
        bq.set_wait_limit(0)
        with pytest.raises(M.MsbgError):
            bq.run(8, useGraph=M.Batch.PERSISTENT)
        bq.set_wait_limit(2000000)
        for ga, gb in zip(a, b):  # start again from the same field
            ga.upload(BCH, 0, gb.download(BCH))
        bq.run(2, useGraph=M.Batch.PERSISTENT)
    for g in b:
        g.applyChannelPdeFast(BCH, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 2, 1.0, 0.1)
    for ga, gb in zip(a, b):
        assert np.array_equal(ga.download(BCH).view(np.uint16), gb.download(BCH).view(np.uint16))


def test_batch_autotune_keeps_results():
    note("batch autotune")
    specs = [(128, 32), (128, 16)]
    a = _batch_grids(specs, seed=21)
    b = _batch_grids(specs, seed=21)
    before = [g.download(BCH).copy() for g in a]
    with M.Batch() as bt:
        for g in a:
            bt.add(g, BCH, 0.1, 1)
        bt.autotune(passes=2)
        tiles = bt.tiles()
        assert len(tiles) == 2 and all(t in (0, 1, 2) for _, _, t in tiles)
        assert all(m in (0, 1) for m in bt.modes())
        # tuning restores the pages it sweeps
        for g, x in zip(a, before):
            assert np.array_equal(g.download(BCH).view(np.uint16), x.view(np.uint16))
        bt.run(2, useGraph=True)
    for g in b:
        g.applyChannelPdeFast(BCH, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 2, 1.0, 0.1)
    # the chosen variant computes the same update (up to its FP32 rounding order)
    for ga, gb in zip(a, b):
        d = np.abs(ga.download(BCH).astype(np.float32) -
                   gb.download(BCH).astype(np.float32))
        assert np.percentile(d, 99.9) <= 2e-3 and d.max() <= 0.2


def test_batched_study_small():
    """The batched-solve study at small sizes: every batched launch mode gives
    the bitwise result of sweeping the grids one at a time."""
    note("batched study small")
    from msbg_amd.studies import batched_study
    r = batched_study(nGrids=8, sizes=(64, 128), levels=(1, 2, 3), sweeps=4, seed=5)
    assert r["voxels"] > 0 and r["grids"] == 8
    assert sorted(int(k) for k in r["by_size"]) == [64, 128]
    assert all(t in (0, 1, 2) for _, _, t in r["tiles"])


def test_points_generate_on_device():
    note("points generate")
    S = 256
    g = M.MultiresSparseGrid.create("pg", S, S, S, 16, options=M.OPT_SINGLE_LEVEL, nLevels=2)
    with g:
        n = g.points_generate(11, 1, 128, 128, 40, 40, 20000, 0.5, 0.0)
        assert n == 20000
        p = g.points_download().astype(np.float64)
        r = np.linalg.norm(p - 128.0, axis=1)
        assert abs(r.mean() - 40) < 0.05 and 0.4 < r.std() < 0.6
        assert np.array_equal(g.points_download(), p.astype(np.float32))
        # deterministic, and append keeps the earlier points in front
        n2 = g.points_generate(11, 1, 128, 128, 40, 40, 20000, 0.5, 0.0)
        assert n2 == 20000 and np.array_equal(g.points_download(), p.astype(np.float32))
        n3 = g.points_generate(12, 2, 60, 200, 10, 20, 5000, 0.5, 1.0, 0.2, 0.05, append=True)
        q = g.points_download()
        assert n3 == 30000 and np.array_equal(q[:20000], p.astype(np.float32))
        # a sphere across the domain corner: the filter of msbg_points_upload
        n4 = g.points_generate(13, 1, 2, 2, 30, 30, 20000, 0.5, 0.0)
        kept = g.points_download()
        assert 1000 < n4 < 6000 and kept.shape == (n4, 3)
        assert O.domain_filter(S, S, S, kept).all()
