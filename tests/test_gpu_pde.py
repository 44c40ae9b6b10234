"""GPU parity of the other applyChannelPdeFast operators, driver modes and
constraints (msbg4.cpp:60-957) against the CPU oracle, through the C ABI.
Every combination here runs the exact arithmetic, so results must be
bit-identical and residuals equal to FP64 rounding."""
import sys

import numpy as np
import pytest

from oracle import liboracle as O

pytestmark = pytest.mark.gpu
M = pytest.importorskip("msbg_amd")

CH, CH_T, CH_T2 = 50, 51, 52
FLAG8 = M.OPT_8_COLOR_SCHEME


def note(*a):
    print("[pde]", *a, flush=True)
    sys.stdout.flush()


def _grid(S=64, bsx=16, storage=M.F32, seed=1, nLevels=2):
    nb = S // bsx
    rng = np.random.default_rng(seed)
    g = M.MultiresSparseGrid.create("pde", S, S, S, bsx, nLevels=nLevels)
    lv = np.where(rng.random(nb ** 3) < 0.15, 1, 0).astype(np.int32)
    g.setRefinementMap(lv)
    kind = rng.choice([3, 1, 2, 0], nb ** 3, p=[0.6, 0.15, 0.1, 0.15]).astype(np.uint8)
    full = {M.U16: 65535.0, M.U8: 255.0}.get(storage, 1.0)
    empty = {M.U16: 2000.0, M.U8: 8.0}.get(storage, 0.03)
    g.channel_define(CH, storage, 0, kind, empty, full, 0.0)
    pm, npg = g.pagemap(CH)
    # a smooth field plus noise, in [0, 1]
    z, y, x = np.meshgrid(*(np.arange(bsx),) * 3, indexing="ij")
    vals = np.empty((npg, bsx, bsx, bsx), np.float32)
    for p in range(npg):
        c = rng.uniform(0, bsx, 3)
        r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
        vals[p] = np.clip(0.5 + 0.4 * np.cos(r / rng.uniform(2, 5)) + rng.normal(0, 0.03, r.shape),
                          0, 1)
    if storage == M.U16:
        raw = np.rint(vals * 65535).astype(np.uint16)
    elif storage == M.U8:
        raw = np.rint(vals * 255).astype(np.uint8)
    elif storage == M.F16:
        raw = vals.astype(np.float16)
    else:
        raw = vals
    g.upload(CH, 0, raw)
    ost = {M.U16: O.U16, M.F32: O.F32, M.F16: O.F16, M.U8: O.U8}[storage]
    levels, _ = g.getBlockInfo()
    return g, pm, ost, empty, full, levels


def _run_both(g, pm, ost, empty, full, levels, S, bsx, laplTyp, numIter, blocks=None, dt=0.1,
              c01=1, densThr=0.0, tol1=0.0, tol2=0.0, taubin=0.0, alpha=0.0):
    pages = g.download(CH).copy()
    if ost == O.F16:
        pages = pages.view(np.uint16)
    L = O.Level(S, S, S, bsx, pm, pages, ost, empty, full)
    op, mode = laplTyp, O.MODE_JACOBI
    if op < 0:
        op = -op
        if op & FLAG8:
            op &= ~FLAG8
            mode = O.MODE_COLOR8
        else:
            mode = O.MODE_REDBLACK
    bl = np.nonzero(pm >= 0)[0] if blocks is None else np.asarray(blocks)
    want_r = O.pde_sweeps(L, bl, levels, numIter, op, mode, dt, c01, densThr, tol1, tol2, taubin,
                          alpha, residual=True)
    got_r = g.applyChannelPdeFast(CH, CH_T, CH_T2, blocks, laplTyp, c01, densThr, tol1, tol2,
                                  numIter, 1.0, dt, residual=True, taubinStrength=taubin,
                                  anisDiffAlpha=alpha)
    got = g.download(CH)
    if ost == O.F16:
        got = got.view(np.uint16)
    return got, L.pages, got_r, want_r


CASES = [
    # (op, alpha, taubin)
    (1, 0.0, 0.0), (1, 0.0, 0.63), (2, 0.0, 0.0), (3, 0.0, 0.0), (14, 0.2, 0.0), (14, -0.15, 0.0),
    (15, 0.3, 0.0), (17, 0.0, 0.0),
]


# every (operator, mode) pair but the bi-laplacian (op 2) in red-black mode:
# blocks of one red-black pass share edges, which the bi-laplacian's stencil
# reads, so its result depends on block order, in the reference as here
MODES = ["color8", "redblack", "jacobi"]
OP_MODES = [(c, m) for c in CASES for m in MODES if not (c[0] == 2 and m == "redblack")]


@pytest.mark.parametrize("case,mode", OP_MODES)
def test_operators_bitexact(case, mode):
    op, alpha, taubin = case
    note("op", op, alpha, taubin, mode)
    S, bsx = 64, 16
    lt = {"color8": -(op | FLAG8), "redblack": -op, "jacobi": op}[mode]
    # UNORM8 (applyChannelPdeFast<uint8_t>): the stochastic store, same steps
    # on both sides; the bi-laplacian has no UNORM8 form here
    for storage in (M.F32, M.U16) if op == 2 else (M.F32, M.U16, M.U8):
        g, pm, ost, empty, full, levels = _grid(S, bsx, storage, seed=op + 7)
        with g:
            got, want, gr, wr = _run_both(g, pm, ost, empty, full, levels, S, bsx, lt, 3,
                                          alpha=alpha, taubin=taubin)
            assert np.array_equal(got, want), (storage, int((got != want).sum()))
            np.testing.assert_allclose(gr, wr, rtol=1e-12, atol=1e-15)
            # the temporaries are gone afterwards, as the reference resets them
            with pytest.raises(M.MsbgError):
                g.pagemap(CH_T)


@pytest.mark.parametrize("storage", [M.F32, M.U16, M.F16, M.U8])
def test_mcf_jacobi_and_hard_constraints_bitexact(storage):
    note("mcf jacobi / hard", storage)
    S, bsx = 64, 16
    for lt, kw in ((4, {}), (-(4 | FLAG8), dict(densThr=0.3, tol1=0.02, tol2=0.05)),
                   (4, dict(densThr=0.3, tol1=0.01, tol2=0.0)),
                   (3, dict(densThr=0.3, tol1=0.03, tol2=0.03))):
        g, pm, ost, empty, full, levels = _grid(S, bsx, storage, seed=5)
        with g:
            got, want, gr, wr = _run_both(g, pm, ost, empty, full, levels, S, bsx, lt, 4, **kw)
            assert np.array_equal(got, want), (lt, kw)
            np.testing.assert_allclose(gr, wr, rtol=1e-12, atol=1e-15)


def test_jacobi_block_list_and_odd_count():
    note("jacobi list")
    S, bsx = 64, 16
    g, pm, ost, empty, full, levels = _grid(S, bsx, M.F32, seed=11)
    with g:
        vb = np.nonzero(pm >= 0)[0]
        sub = vb[::2]  # unswept value blocks are NULL in the temporary
        got, want, gr, wr = _run_both(g, pm, ost, empty, full, levels, S, bsx, 17, 5, blocks=sub)
        assert np.array_equal(got, want)
        assert gr[-1] == 0.0 and np.allclose(gr, wr, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("bsx", [8, 32])
def test_other_block_sizes(bsx):
    note("pm / bilap bsx", bsx)
    S = 4 * bsx
    for lt, kw in ((-(14 | FLAG8), dict(alpha=0.25)), (-(2 | FLAG8), {}),
                   (2, dict(densThr=0.5, tol1=0.01, tol2=0.02))):
        g, pm, ost, empty, full, levels = _grid(S, bsx, M.F32, seed=13)
        with g:
            got, want, gr, wr = _run_both(g, pm, ost, empty, full, levels, S, bsx, lt, 2, **kw)
            assert np.array_equal(got, want), lt
            np.testing.assert_allclose(gr, wr, rtol=1e-12, atol=1e-15)


def test_mcf_red_black_runs():
    # 19-point stencil with red-black passes: blocks of one pass share edges,
    # so the result depends on the order blocks run in (as in the
    # reference's threaded loop); check it runs and stays in [0, 1]
    g, pm, ost, empty, full, levels = _grid(64, 16, M.F32, seed=3)
    with g:
        r = g.applyChannelPdeFast(CH, 0, 0, None, -4, 1, 0, 0, 0, 2, 1.0, 0.1, residual=True)
        v = g.download(CH)
        assert np.isfinite(r).all() and v.min() >= 0 and v.max() <= 1


def test_unsupported_and_bad_arguments():
    g, pm, ost, empty, full, levels = _grid(64, 32, M.F32, seed=2)
    with g:
        for lt in (-(1 | FLAG8), -(3 | FLAG8), 17):  # fixed 16^3-halo offsets
            with pytest.raises(M.MsbgError) as e:
                g.applyChannelPdeFast(CH, CH_T, 0, None, lt, 1, 0, 0, 0, 2, 1.0, 0.1)
            assert e.value.code == M._lib.EUNSUPP
        with pytest.raises(M.MsbgError) as e:  # redistancing (10) is not built
            g.applyChannelPdeFast(CH, CH_T, 0, None, -(10 | FLAG8), 1, 0, 0, 0, 2, 1.0, 0.1)
        assert e.value.code == M._lib.EUNSUPP
        with pytest.raises(M.MsbgError) as e:  # Jacobi without a temporary
            g.applyChannelPdeFast(CH, 0, 0, None, 4, 1, 0, 0, 0, 2, 1.0, 0.1)
        assert e.value.code == M._lib.EINVAL
        with pytest.raises(M.MsbgError) as e:  # Perona-Malik needs alpha
            g.applyChannelPdeFast(CH, 0, 0, None, -(14 | FLAG8), 1, 0, 0, 0, 2, 1.0, 0.1)
        assert e.value.code == M._lib.EINVAL
