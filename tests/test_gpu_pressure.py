"""GPU parity of the adaptive pressure relax (msbg_ps_*, k_ps.hip) against
the CPU oracle (oracle/ps_oracle.c): set-up (cell flags, block flags,
diagonal), relax on every multigrid level in both orders, Laplacian multiply
and residual, restriction, prolongation and V-cycles, all bit-exact; the dot
product of the multiply to 1e-12 relative (its summation order differs)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
M = pytest.importorskip("msbg_amd")

from msbg_amd import pressure as G  # noqa: E402
from oracle import ps as P  # noqa: E402
from tests import ps_cases as PC  # noqa: E402


def note(*a):
    print("[pressure]", *a, flush=True)
    sys.stdout.flush()


def pair(inp, **kw):
    o = P.PsOracle(inp["nb"], inp["bsx0"], inp["level"], inp["cell_type"], dbc=inp["dbc"],
                   ndense=inp["ndense"], **kw)
    g = G.PressureGrid(inp["nb"], inp["bsx0"], inp["level"], inp["cell_type"], dbc=inp["dbc"],
                       dense_levels=inp["ndense"],
                       block_iter=kw.get("block_iter", G.BLOCK_ITER))
    assert (g.nmg, g.nlev) == (o.nmg, o.nlev)
    return o, g


def same(a, b, what):
    assert a.shape == b.shape, what
    bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
    assert bad.size == 0, f"{what}: {bad.size} differ, first {bad[:5]} gpu {a[bad[:5]]} cpu {b[bad[:5]]}"


def random_fields(o, m, seed):
    fl = o.flags(m)
    x = PC.zero_nonfluid(PC.pages_random(o.size(m), seed), fl)
    b = PC.zero_nonfluid(PC.pages_random(o.size(m), seed + 1, 0.01), fl)
    return x, b


@pytest.mark.parametrize("case", PC.CASES, ids=lambda c: c[0])
def test_setup_matches_oracle(case):
    inp = PC.build_inputs(case)
    note("setup", case[0])
    o, g = pair(inp)
    with g:
        for m in range(o.nmg):
            assert g.size(m) == o.size(m)
            assert np.array_equal(g.flags(m), o.flags(m)), f"flags level {m}"
            same(g.download(G.CH_D, m), o.get(P.CH_D, m), f"diagonal level {m}")
            if m < o.nlev:
                assert np.array_equal(g.block_flags(m), o.block_flags(m)), f"block flags {m}"


@pytest.mark.parametrize("case", PC.CASES, ids=lambda c: c[0])
def test_relax_every_level_bitexact(case):
    inp = PC.build_inputs(case)
    o, g = pair(inp)
    with g:
        for m in range(o.nmg):
            for rev in (0, 1):
                note("relax", case[0], "level", m, "reverse", rev)
                x, b = random_fields(o, m, 100 + 7 * m + rev)
                for c, v in ((P.CH_X, x), (P.CH_B, b)):
                    o.set(c, m, v)
                    g.upload(c, m, v)
                o.relax(m, 2, reverse=rev)
                g.relax(0, rev, m, G.CH_X, G.CH_B, G.CH_T, 2)
                same(g.download(G.CH_X, m), o.get(P.CH_X, m), f"relax level {m} rev {rev}")


@pytest.mark.parametrize("block_iter", [1, 2, 4])
def test_relax_sub_iteration_counts(block_iter):
    inp = PC.build_inputs(PC.CASES[1])
    o, g = pair(inp, block_iter=block_iter)
    with g:
        for m in range(o.nlev):
            note("relax block_iter", block_iter, "level", m)
            x, b = random_fields(o, m, 300 + m)
            for c, v in ((P.CH_X, x), (P.CH_B, b)):
                o.set(c, m, v)
                g.upload(c, m, v)
            o.relax(m, 1)
            g.relax(0, 0, m, G.CH_X, G.CH_B, G.CH_T, 1)
            same(g.download(G.CH_X, m), o.get(P.CH_X, m), f"relax level {m}")


@pytest.mark.parametrize("case", PC.CASES, ids=lambda c: c[0])
def test_multiply_and_residual_bitexact(case):
    inp = PC.build_inputs(case)
    o, g = pair(inp)
    with g:
        for m in range(o.nmg):
            note("multiply", case[0], "level", m)
            x, b = random_fields(o, m, 200 + m)
            for c, v in ((P.CH_X, x), (P.CH_B, b)):
                o.set(c, m, v)
                g.upload(c, m, v)
            o.fill(P.CH_R, m, 0.0)
            g.fill(G.CH_R, m, 0.0)
            dense = m >= o.nlev
            d_o = o.multiply(m, 0)
            d_g = g.multiplyLaplacianMatrixOpt(0, m, G.CH_X, G.CH_B, 0.0, G.CH_R, want_dot=not dense)
            same(g.download(G.CH_R, m), o.get(P.CH_R, m), f"multiply level {m}")
            if not dense:
                assert abs(d_g - d_o) <= 1e-12 * max(abs(d_o), 1e-30) + 1e-300, (d_g, d_o)
            o.multiply(m, P.OPT_CALC_RESIDUAL)
            g.multiplyLaplacianMatrixOpt(G.OPT_CALC_RESIDUAL, m, G.CH_X, G.CH_B, 0.0, G.CH_R)
            same(g.download(G.CH_R, m), o.get(P.CH_R, m), f"residual level {m}")


@pytest.mark.parametrize("case", PC.CASES, ids=lambda c: c[0])
def test_transfers_bitexact(case):
    inp = PC.build_inputs(case)
    o, g = pair(inp)
    with g:
        for m in range(o.nmg - 1):
            note("transfers", case[0], "level", m)
            r, _ = random_fields(o, m, 400 + m)
            for c in (P.CH_R,):
                o.set(c, m, r)
                g.upload(c, m, r)
            o.restrict(m, P.CH_R, P.CH_B)
            g.restrict(m, G.CH_R, G.CH_B)
            same(g.download(G.CH_B, m + 1), o.get(P.CH_B, m + 1), f"restrict {m}")
            e, _ = random_fields(o, m + 1, 500 + m)
            x, _ = random_fields(o, m, 600 + m)
            for c, v, lv in ((P.CH_X, e, m + 1), (P.CH_X, x, m)):
                o.set(c, lv, v)
                g.upload(c, lv, v)
            o.prolongate(m, P.CH_X, P.CH_X)
            g.prolongate(m, G.CH_X, G.CH_X)
            same(g.download(G.CH_X, m), o.get(P.CH_X, m), f"prolongate {m}")


@pytest.mark.parametrize("case", PC.CASES, ids=lambda c: c[0])
def test_vcycle_bitexact(case):
    inp = PC.build_inputs(case)
    o, g = pair(inp)
    with g:
        x, b = random_fields(o, 0, 700)
        o.set(P.CH_B, 0, b)
        g.upload(G.CH_B, 0, b)
        for cyc in range(2):
            note("vcycle", case[0], cyc)
            o.vcycle(2, 2, 30)
            g.vcycle(2, 2, 30)
            for m in range(o.nmg):
                same(g.download(G.CH_X, m), o.get(P.CH_X, m), f"cycle {cyc} X level {m}")
                same(g.download(G.CH_B, m), o.get(P.CH_B, m), f"cycle {cyc} B level {m}")


def test_large_adaptive_grid_relax_and_vcycle():
    """16x16x8 blocks of 16^3 (256x256x128 cells) on four levels, a solid
    sphere, a free surface: one relax pass and one V-cycle, bit-exact."""
    nb, bsx0 = (16, 16, 8), 16
    lev = PC.levels_random(nb, 4, 21)
    t = PC.cell_types(nb, bsx0, 21)
    inp = dict(nb=nb, bsx0=bsx0, level=lev, cell_type=t, dbc=(0, 0, 1, 1, 1, 0), ndense=2)
    o, g = pair(inp)
    with g:
        x, b = random_fields(o, 0, 800)
        for c, v in ((P.CH_X, x), (P.CH_B, b)):
            o.set(c, 0, v)
            g.upload(c, 0, v)
        note("large relax")
        o.relax(0, 1)
        g.relax(0, 0, 0, G.CH_X, G.CH_B, G.CH_T, 1)
        same(g.download(G.CH_X, 0), o.get(P.CH_X, 0), "large relax")
        note("large vcycle")
        o.vcycle(1, 1, 20)
        g.vcycle(1, 1, 20)
        same(g.download(G.CH_X, 0), o.get(P.CH_X, 0), "large vcycle")


def test_gpu_vcycle_converges():
    """Manufactured solution: b = A x*, V-cycles from zero reach x*."""
    inp = PC.build_inputs(PC.CASES[1])
    o, g = pair(inp)
    with g:
        xs = PC.zero_nonfluid(PC.pages_random(g.size(0), 11), g.flags(0))
        g.upload(G.CH_X, 0, xs)
        g.fill(G.CH_B, 0, 0.0)
        g.multiplyLaplacianMatrixOpt(0, 0, G.CH_X, G.CH_B, 0.0, G.CH_B)
        b = g.download(G.CH_B, 0)
        g.fill(G.CH_X, 0, 0.0)
        for _ in range(8):
            g.vcycle(2, 2, 60)
        g.fill(G.CH_R, 0, 0.0)
        g.multiplyLaplacianMatrixOpt(G.OPT_CALC_RESIDUAL, 0, G.CH_X, G.CH_B, 0.0, G.CH_R)
        r = g.download(G.CH_R, 0)
        assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-3
        assert np.abs(g.download(G.CH_X, 0) - xs).max() < 0.02 * np.abs(xs).max()


def test_argument_errors():
    inp = PC.build_inputs(PC.CASES[0])
    o, g = pair(inp)
    with g:
        with pytest.raises(M.MsbgError) as e:
            g.relax(1, 0, 0, G.CH_X, G.CH_B, G.CH_T, 1)  # boundary zone only
        assert "EUNSUPP" in str(e.value)
        with pytest.raises(M.MsbgError):
            g.relax(0, 0, 0, G.CH_X, G.CH_X, G.CH_T, 1)
        with pytest.raises(M.MsbgError):
            g.restrict(g.nmg - 1, G.CH_R, G.CH_B)
    with pytest.raises(M.MsbgError):  # two levels apart
        G.PressureGrid((2, 1, 1), 16, np.array([0, 2], np.int8), np.zeros((16, 16, 32), np.uint8))
    with pytest.raises(M.MsbgError):  # all-solid grid is fine, a bad type is not
        G.PressureGrid((1, 1, 1), 8, np.zeros(1, np.int8), np.full((8, 8, 8), 7, np.uint8))
    with G.PressureGrid((1, 1, 1), 8, np.zeros(1, np.int8), np.full((8, 8, 8), 1, np.uint8)) as s:
        s.relax(0, 0, 0, G.CH_X, G.CH_B, G.CH_T, 3)  # no fluid: nothing to relax
        s.vcycle(1, 1, 2)
        assert np.all(s.download(G.CH_X, 0) == 0)
