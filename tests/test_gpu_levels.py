"""GPU parity of the coarse-level sweeps (msbg_apply_channel_pde_level, an
extension of the level-0 sweep of msbg4.cpp:60-957) against the oracle's
or_mcf_sweeps_level, through the C ABI: exact arithmetic bit-exact for every
halo rule, storage and level; fast arithmetic within the sweep's tolerance."""
import sys

import numpy as np
import pytest

from oracle import liboracle as O

pytestmark = pytest.mark.gpu
M = pytest.importorskip("msbg_amd")

CH = 44
OST = {M.U16: O.U16, M.F32: O.F32, M.F16: O.F16, M.BF16: O.BF16}


def note(*a):
    print("[levels]", *a, flush=True)
    sys.stdout.flush()


def _raw(storage, vals):
    """float32 values in [0, 1] as the channel's raw elements"""
    return {M.U16: lambda v: np.rint(v * 65535).astype(np.uint16),
            M.F32: lambda v: v, M.F16: lambda v: v.astype(np.float16),
            M.BF16: lambda v: (v.view(np.uint32) >> 16).astype(np.uint16)}[storage](vals)


def _grid(S, bsx, storage, seed, ghosts):
    """Three levels; every level holds its own blocks as value pages and
    NULL elsewhere, then (float storages) prepareGhostBlocks adds the ghost
    pages of finer blocks."""
    nb = S // bsx
    rng = np.random.default_rng(seed)
    m = np.where(rng.random(nb ** 3) < 0.2, 0, 2).astype(np.int32)
    g = M.MultiresSparseGrid.create("lv", S, S, S, bsx, nLevels=3)
    lv = g.regularizeRefinementMap(m)
    g.setRefinementMap(lv)
    g.set_exists(None)
    levels, _ = g.getBlockInfo()
    full = 65535.0 if storage == M.U16 else 1.0
    empty = 1000.0 if storage == M.U16 else 0.03
    for level in range(3):
        kind = np.where(levels == level, 3, 0).astype(np.uint8)
        g.channel_define(CH, storage, level, kind, empty, full, 0.0)
        pm, npg = g.pagemap(CH, level)
        b = bsx >> level
        z, y, x = np.meshgrid(*(np.arange(b),) * 3, indexing="ij")
        vals = np.empty((npg, b, b, b), np.float32)
        for p in range(npg):
            c = rng.uniform(0, b, 3)
            r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
            vals[p] = np.clip(0.5 + 0.45 * np.cos(r / rng.uniform(1.5, 4)), 0, 1)
        g.upload(CH, level, _raw(storage, vals))
    if ghosts:
        g.prepareGhostBlocks(CH)
    return g, levels, empty, full


def _olevels(g, S, bsx, storage, empty, full):
    out = []
    for level in range(3):
        pm, _ = g.pagemap(CH, level)
        pages = g.download(CH, level)
        if storage == M.F16:
            pages = pages.view(np.uint16)
        s, b = S >> level, bsx >> level
        out.append(O.Level(s, s, s, b, pm, pages.copy(), OST[storage], empty, full))
    return out


CASES = [  # (storage, ghosts)
    (M.F32, True), (M.F16, True), (M.BF16, False), (M.U16, False),
]


@pytest.mark.parametrize("storage,ghosts", CASES)
@pytest.mark.parametrize("opt", [M.OPT_BC_NEUMANN, M.OPT_BC_COARSE_LEVEL,
                                 M.OPT_BC_COARSE_LEVEL | M.OPT_INTERPOLATE])
def test_level_sweeps_exact_bitexact(storage, ghosts, opt):
    S, bsx = 128, 32
    note("levels exact", storage, ghosts, opt)
    g, levels, empty, full = _grid(S, bsx, storage, seed=int(storage) + 7 * int(ghosts), ghosts=ghosts)
    with g:
        g.set_arith(M.ARITH_EXACT)
        for level in (1, 0) if opt != M.OPT_BC_NEUMANN else (2, 1, 0):
            L = _olevels(g, S, bsx, storage, empty, full)
            lo = L[level + 1] if level + 1 < 3 else None
            blocks = np.nonzero(levels == level)[0]
            assert blocks.size > 0
            want = O.mcf_sweeps_level(L[level], lo, blocks, levels, level, opt, 3, 0.1, 1,
                                      residual=True)
            got = g.applyChannelPdeLevel(CH, level, opt, None, 3, 0.1, 1, residual=True)
            out = g.download(CH, level)
            if storage == M.F16:
                out = out.view(np.uint16)
            assert np.array_equal(out, L[level].pages), (level, int((out != L[level].pages).sum()))
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
            assert got[0] > 0


def test_level0_neumann_equals_hot_path():
    S, bsx = 128, 32
    note("level 0 = hot path")
    g, levels, empty, full = _grid(S, bsx, M.F16, seed=3, ghosts=False)
    with g:
        before = g.download(CH, 0).copy()
        r1 = g.applyChannelPdeLevel(CH, 0, M.OPT_BC_NEUMANN, None, 2, 0.1, 1, residual=True)
        a = g.download(CH, 0).copy()
        g.upload(CH, 0, before)
        r2 = g.applyChannelPdeFast(CH, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 2, 1.0, 0.1,
                                   residual=True)
        b = g.download(CH, 0)
        assert np.array_equal(a.view(np.uint16), b.view(np.uint16))
        np.testing.assert_allclose(r1, r2, rtol=1e-12)


def test_level_sweeps_fast_close_to_exact():
    S, bsx = 128, 32
    note("levels fast")
    g, levels, empty, full = _grid(S, bsx, M.F32, seed=9, ghosts=True)
    with g:
        L = _olevels(g, S, bsx, M.F32, empty, full)
        blocks = np.nonzero(levels == 1)[0]
        O.mcf_sweeps_level(L[1], L[2], blocks, levels, 1, M.OPT_BC_COARSE_LEVEL, 4, 0.1, 1)
        g.applyChannelPdeLevel(CH, 1, M.OPT_BC_COARSE_LEVEL, None, 4, 0.1, 1)
        d = np.abs(g.download(CH, 1) - L[1].pages)
        assert d.max() < 1e-3 and np.sqrt((d.astype(np.float64) ** 2).mean()) < 1e-6


def test_level_sweep_block_list_and_errors():
    S, bsx = 128, 32
    g, levels, empty, full = _grid(S, bsx, M.F32, seed=5, ghosts=True)
    with g:
        g.set_arith(M.ARITH_EXACT)
        L = _olevels(g, S, bsx, M.F32, empty, full)
        b1 = np.nonzero(levels == 1)[0]
        sub = np.concatenate([b1[::3], np.nonzero(levels == 0)[0][:5]])  # other levels skipped
        O.mcf_sweeps_level(L[1], L[2], sub, levels, 1, M.OPT_BC_COARSE_LEVEL, 2, 0.1, 1)
        g.applyChannelPdeLevel(CH, 1, M.OPT_BC_COARSE_LEVEL, sub, 2, 0.1, 1)
        assert np.array_equal(g.download(CH, 1), L[1].pages)
        with pytest.raises(M.MsbgError) as e:  # no level 3 to read
            g.applyChannelPdeLevel(CH, 2, M.OPT_BC_COARSE_LEVEL, None, 2, 0.1, 1)
        assert e.value.code == M._lib.ESTATE
        with pytest.raises(M.MsbgError) as e:
            g.applyChannelPdeLevel(CH, 1, M.OPT_INTERPOLATE, None, 2, 0.1, 1)
        assert e.value.code == M._lib.EINVAL
        with pytest.raises(M.MsbgError) as e:
            g.applyChannelPdeLevel(CH, 3, M.OPT_BC_NEUMANN, None, 2, 0.1, 1)
        assert e.value.code == M._lib.EINVAL


def test_level_sweep_cached_setup_follows_changes():
    """The set-up (ghost list, halo table) is cached across calls; ghost
    values and pages are taken afresh each call, and a topology change
    rebuilds it."""
    S, bsx = 128, 32
    note("levels cache")
    g, levels, empty, full = _grid(S, bsx, M.F32, seed=12, ghosts=False)
    with g:
        g.set_arith(M.ARITH_EXACT)
        opt = M.OPT_BC_COARSE_LEVEL | M.OPT_INTERPOLATE
        b1 = np.nonzero(levels == 1)[0]
        L = _olevels(g, S, bsx, M.F32, empty, full)
        O.mcf_sweeps_level(L[1], L[2], b1, levels, 1, opt, 2, 0.1, 1)
        g.applyChannelPdeLevel(CH, 1, opt, None, 2, 0.1, 1)
        assert np.array_equal(g.download(CH, 1), L[1].pages)
        # new coarse values: the second call must read them
        new2 = np.random.default_rng(1).random(L[2].pages.shape).astype(np.float32)
        g.upload(CH, 2, new2)
        L[2].pages[...] = new2
        O.mcf_sweeps_level(L[1], L[2], b1, levels, 1, opt, 2, 0.1, 1)
        g.applyChannelPdeLevel(CH, 1, opt, None, 2, 0.1, 1)
        assert np.array_equal(g.download(CH, 1), L[1].pages)
        # the level's own side-cars carry over between calls only while
        # nothing else writes the level: an upload, or a sweep of it under
        # another halo rule (a set-up of its own), makes the next call
        # rebuild them
        new1 = np.random.default_rng(2).random(L[1].pages.shape).astype(np.float32)
        g.upload(CH, 1, new1)
        L[1].pages[...] = new1
        for o in (opt, M.OPT_BC_NEUMANN, opt, opt):
            O.mcf_sweeps_level(L[1], L[2], b1, levels, 1, o, 2, 0.1, 1)
            g.applyChannelPdeLevel(CH, 1, o, None, 2, 0.1, 1)
            assert np.array_equal(g.download(CH, 1), L[1].pages), o
        # a topology change (ghost pages of finer blocks appear) rebuilds it
        g.prepareGhostBlocks(CH)
        L = _olevels(g, S, bsx, M.F32, empty, full)
        O.mcf_sweeps_level(L[1], L[2], b1, levels, 1, opt, 2, 0.1, 1)
        g.applyChannelPdeLevel(CH, 1, opt, None, 2, 0.1, 1)
        assert np.array_equal(g.download(CH, 1), L[1].pages)


@pytest.mark.parametrize("S,bsx", [(128, 32), (64, 16)])
@pytest.mark.parametrize("storage", [M.F32, M.F16, M.BF16, M.U16])
def test_level_sweep_carried_sidecars_interpolate(storage, S, bsx):
    """Three COARSE_LEVEL|INTERPOLATE calls in a row on one cached set-up,
    only the coarser level changed between them: from the second call on the
    level's own side-cars carry over and the ghost fill rewrites the ghost
    pages' side-car entries. Every call equals the oracle bit for bit, in
    every storage, for 16^3 and 8^3 pages of level 1."""
    note("levels carried side-cars", storage, S, bsx)
    ghosts = storage != M.U16  # prepareGhostBlocks downsamples float storages
    g, levels, empty, full = _grid(S, bsx, storage, seed=21, ghosts=ghosts)
    with g:
        g.set_arith(M.ARITH_EXACT)
        opt = M.OPT_BC_COARSE_LEVEL | M.OPT_INTERPOLATE
        b1 = np.nonzero(levels == 1)[0]
        assert len(b1)
        L = _olevels(g, S, bsx, storage, empty, full)
        rng = np.random.default_rng(5)
        for call in range(3):
            if call:
                shape = g.download(CH, 2).shape
                new2 = _raw(storage, rng.random(shape).astype(np.float32))
                g.upload(CH, 2, new2)
                L[2].pages[...] = new2.view(np.uint16) if storage == M.F16 else new2
            O.mcf_sweeps_level(L[1], L[2], b1, levels, 1, opt, 2, 0.1, 1)
            g.applyChannelPdeLevel(CH, 1, opt, None, 2, 0.1, 1)
            got = g.download(CH, 1)
            if storage == M.F16:
                got = got.view(np.uint16)
            assert np.array_equal(got, L[1].pages), (storage, S, bsx, call)
