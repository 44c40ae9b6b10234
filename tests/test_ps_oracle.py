"""Pins of the pressure-relax oracle (oracle/ps_oracle.c) against what the
mathematics fixes, independent of how the oracle computes it:
  - the discrete Laplacian of quadratics (closed form, exact in binary
    arithmetic) and of linear fields across resolution changes;
  - the assembled operator is exactly symmetric (coarse-fine coefficients
    1/3 and 2/3 pair up) and its rows sum to zero away from Dirichlet cells;
  - one relax pass equals a float64 statement of block Gauss-Seidel with
    Jacobi inside each SIMD vector and a Jacobi last sub-iteration, written
    on the assembled matrix;
  - the dense-level relax is textbook damped Jacobi on the assembled matrix;
  - restriction conserves sums, prolongation of a constant adds it to every
    fluid cell; V-cycles converge to a manufactured solution.
"""
import numpy as np
import pytest

from oracle import ps as P
from tests import ps_cases as PC

FLUIDMASK = P.CELL_SOLID | P.CELL_VOID


def make(case_or_inputs, **kw):
    inp = PC.build_inputs(case_or_inputs) if isinstance(case_or_inputs, tuple) else case_or_inputs
    return P.PsOracle(inp["nb"], inp["bsx0"], inp["level"], inp["cell_type"], dbc=inp["dbc"],
                      ndense=inp["ndense"], **kw), inp


def assemble(o, m):
    """Matrix of the multiply at mg level m, by columns (fluid cells only)."""
    fl = o.flags(m)
    idx = np.nonzero((fl & FLUIDMASK) == 0)[0]
    n = o.size(m)
    A = np.zeros((idx.size, idx.size))
    for k, j in enumerate(idx):
        e = np.zeros(n, np.float32)
        e[j] = 1
        o.set(P.CH_X, m, e)
        o.fill(P.CH_R, m, 0.0)
        o.multiply(m, 0)
        A[:, k] = o.get(P.CH_R, m)[idx]
    return A, idx


def test_quadratic_closed_form():
    # h = 1/32 (a power of two): F = h*x and every sum are exact
    nb, bsx0 = (4, 4, 4), 8
    S = 32
    o = P.PsOracle(nb, bsx0, np.zeros(64, np.int8), np.zeros((S, S, S), np.uint8))
    z, y, x = np.meshgrid(np.arange(S), np.arange(S), np.arange(S), indexing="ij")
    for a, b, c in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 2, -1)]:
        f = (a * x ** 2 + b * y ** 2 + c * z ** 2).astype(np.float32)
        o.set(P.CH_X, 0, P.pages_from_field(lambda l: f, nb, bsx0, o._level, 0))
        o.fill(P.CH_R, 0, 0.0)
        o.multiply(0, 0)
        r = P.field_from_pages(o.get(P.CH_R, 0), nb, bsx0, o._level, 0, 0)
        inner = r[1:-1, 1:-1, 1:-1]
        assert np.all(inner == -2.0 * (a + b + c) / 32), np.unique(inner)
    # a constant has zero Laplacian inside, and the open border's Dirichlet
    # face gives h*c on the face cells (one missing neighbour)
    o.set(P.CH_X, 0, np.full(o.size(0), 3.0, np.float32))
    o.fill(P.CH_R, 0, 0.0)
    o.multiply(0, 0)
    r = P.field_from_pages(o.get(P.CH_R, 0), nb, bsx0, o._level, 0, 0)
    assert np.all(r[1:-1, 1:-1, 1:-1] == 0)
    assert np.all(r[0, 1:-1, 1:-1] == 3.0 / 32) and np.all(r[1:-1, -1, 1:-1] == 3.0 / 32)


@pytest.mark.parametrize("case", [PC.CASES[1], PC.CASES[3], PC.CASES[4]], ids=lambda c: c[0])
def test_operator_symmetric_rows_sum_to_zero(case):
    name, nb, bsx0 = case[0], case[1], case[2]
    small = (min(nb[0], 2), min(nb[1], 2), min(nb[2], 2)) if bsx0 == 16 else nb
    inp = PC.build_inputs((name, small, min(bsx0, 8)) + case[3:])
    o, _ = make(inp)
    for m in range(o.nmg):
        if o.size(m) > 1100:
            continue
        A, idx = assemble(o, m)
        assert np.array_equal(A, A.T), f"level {m}: not symmetric"
        # diagonal h/D
        D = o.get(P.CH_D, m)[idx]
        h = np.float32((1 << min(m, 10)) * o.dx0)
        if m < o.nlev:
            off, L = P.page_offsets(inp["level"], inp["bsx0"], m)
            cellL = np.repeat(L, (inp["bsx0"] >> L) ** 3)[idx]
            h = ((1 << cellL) * o.dx0).astype(np.float32)
        # a fluid cell walled in by solid faces on all sides has D = 0 and an
        # empty row (the multiply's select(D != 0, ...))
        iso = D == 0
        assert np.all(A[iso] == 0)
        np.testing.assert_allclose(np.diag(A)[~iso], (h / np.where(iso, 1, D))[~iso], rtol=1e-6)
        assert np.all(A - np.diag(np.diag(A)) <= 0)
        # rows without a Dirichlet (void / open border) neighbour sum to zero
        rs = A.sum(axis=1)[~iso]
        diri = np.abs(rs) > 1e-5 * np.abs(np.diag(A)[~iso])
        assert diri.mean() < 0.9
        assert np.all(rs[diri] > 0)  # a Dirichlet neighbour only adds to the diagonal


def test_linear_field_across_resolution_changes():
    nb, bsx0 = (4, 2, 2), 16
    lev = PC.levels_by_x(nb, 4)
    S = (32, 32, 64)
    t = np.zeros(S, np.uint8)
    # Dirichlet in x, Neumann (solid) in y and z
    o = P.PsOracle(nb, bsx0, lev, t, dbc=(0, 0, 1, 1, 1, 1))

    def lin(g):
        def f(l):
            s = 1 << l
            z, y, x = np.meshgrid((np.arange(32 >> l) + 0.5) * s, (np.arange(32 >> l) + 0.5) * s,
                                  (np.arange(64 >> l) + 0.5) * s, indexing="ij")
            return g[0] * x + g[1] * y + g[2] * z
        return f

    # linear in x: zero Laplacian everywhere but the Dirichlet x faces
    o.set(P.CH_X, 0, P.pages_from_field(lin((0.25, 0, 0)), nb, bsx0, lev, 0))
    o.fill(P.CH_R, 0, 0.0)
    o.multiply(0, 0)
    r = o.get(P.CH_R, 0)
    for l in range(4):
        f = P.field_from_pages(r, nb, bsx0, lev, 0, l)
        inner = f[:, :, 1:-1]
        if np.isfinite(inner).any():
            assert np.nanmax(np.abs(inner)) < 2e-5, (l, np.nanmax(np.abs(inner)))
    # linear in y: the coarse side of every transition sums its four fine
    # neighbours symmetrically, so it stays at zero too (away from the
    # Neumann y faces)
    o.set(P.CH_X, 0, P.pages_from_field(lin((0, 0.25, 0)), nb, bsx0, lev, 0))
    o.fill(P.CH_R, 0, 0.0)
    o.multiply(0, 0)
    r = o.get(P.CH_R, 0)
    for l in range(1, 4):
        f = P.field_from_pages(r, nb, bsx0, lev, 0, l)
        cols = np.nonzero(np.isfinite(f[0, 0]))[0]
        # drop the region's last column: the fine side of the next transition
        # (or the Dirichlet border)
        inner = f[:, 1:-1, cols[:-1]]
        assert np.abs(inner).max() < 2e-5, (l, np.abs(inner).max())
    # ... while a fine cell next to a coarse one sees the coarse centre half a
    # fine cell off in y: A x = h_f * 2/3 * (x_f - x_c) = -/+ h_f * 2/3 * 0.25 * 0.5
    f0 = P.field_from_pages(r, nb, bsx0, lev, 0, 0)
    edge = f0[:, 1:-1, 15]  # level-0 cells left of the level-1 blocks
    hf = np.float32(1 / 64)
    expect = hf * (2 / 3) * 0.25 * 0.5
    np.testing.assert_allclose(np.abs(edge), expect, rtol=1e-4)


def relax_reference(A, idx, x, b, order_vecs, blocks, omega, K, rev0):
    """One relax pass in float64 on the assembled matrix: blocks in the
    given colour order, K sub-iterations per block, the first K-1 sweeping
    SIMD vectors in sequence (Jacobi inside a vector), alternately reversed
    starting with reversed when rev0, the last one Jacobi."""
    pos = {j: k for k, j in enumerate(idx)}
    x = x.astype(np.float64).copy()
    diag = np.diag(A)
    for blk in blocks:
        vecs = order_vecs[blk]
        for it in range(K):
            rev = (it % 2 == 0) == bool(rev0)
            if it < K - 1:
                seq = vecs[::-1] if rev else vecs
                for v in seq:
                    ks = [pos[j] for j in v if j in pos]
                    if not ks:
                        continue
                    new = {}
                    for k in ks:
                        s = b[k] - A[k] @ x + diag[k] * x[k]
                        new[k] = (1 - omega) * x[k] + omega * s / diag[k]
                    for k, val in new.items():
                        x[k] = val
            else:
                ks = [pos[j] for v in vecs for j in v if j in pos]
                new = {k: (1 - omega) * x[k] + omega * (b[k] - A[k] @ x + diag[k] * x[k]) / diag[k]
                       for k in ks}
                for k, val in new.items():
                    x[k] = val
    return x


@pytest.mark.parametrize("reverse", [0, 1])
def test_relax_pass_is_block_gauss_seidel(reverse):
    nb, bsx0 = (2, 2, 1), 8
    lev = np.array([0, 1, 1, 0], np.int8)
    t = PC.cell_types(nb, bsx0, 9, solid=True, void=True)
    o = P.PsOracle(nb, bsx0, lev, t, dbc=(0, 1, 0, 1, 1, 0), block_iter=3)
    m = 0
    A, idx = assemble(o, m)
    fl = o.flags(m)
    rng = np.random.default_rng(3)
    xs = PC.zero_nonfluid(rng.standard_normal(o.size(m)).astype(np.float32), fl)
    bs = PC.zero_nonfluid(rng.standard_normal(o.size(m)).astype(np.float32), fl)
    off, L = P.page_offsets(lev, bsx0, m)
    order_vecs, colours = {}, {0: [], 1: []}
    for blk in range(4):
        bsx = bsx0 >> L[blk]
        sw = 8 if bsx >= 8 else bsx
        order_vecs[blk] = [list(range(off[blk] + v * sw, off[blk] + (v + 1) * sw))
                           for v in range(bsx ** 3 // sw)]
        bx, by = blk % 2, blk // 2
        colours[(bx + by) & 1].append(blk)
    first = 0 if reverse else 1
    blocks = colours[first] + colours[1 - first]
    o.set(P.CH_X, m, xs)
    o.set(P.CH_B, m, bs)
    o.relax(m, 1, reverse=reverse)
    got = o.get(P.CH_X, m)[idx]
    want = relax_reference(A, idx, xs[idx], bs[idx].astype(np.float64), order_vecs, blocks,
                           float(P.OMEGA), 3, rev0=not reverse)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-5 * np.abs(want).max())
    # a plain Jacobi pass (one sub-iteration) differs from the sweep: the
    # order matters
    o.set(P.CH_X, m, xs)
    o2 = P.PsOracle(nb, bsx0, lev, t, dbc=(0, 1, 0, 1, 1, 0), block_iter=1)
    o2.set(P.CH_X, m, xs)
    o2.set(P.CH_B, m, bs)
    o2.relax(m, 1, reverse=reverse)
    assert np.abs(o2.get(P.CH_X, m)[idx] - got).max() > 1e-3


def test_dense_relax_is_damped_jacobi():
    inp = PC.build_inputs(PC.CASES[2])
    o, _ = make(inp)
    m = o.nlev  # first dense level
    A, idx = assemble(o, m)
    fl = o.flags(m)
    rng = np.random.default_rng(5)
    xs = PC.zero_nonfluid(rng.standard_normal(o.size(m)).astype(np.float32), fl)
    bs = PC.zero_nonfluid(rng.standard_normal(o.size(m)).astype(np.float32), fl)
    o.set(P.CH_X, m, xs)
    o.set(P.CH_B, m, bs)
    o.fill(P.CH_T, m, 0.0)
    o.relax(m, 1)
    got = o.get(P.CH_X, m)[idx]
    x = xs[idx].astype(np.float64)
    want = x + float(P.OMEGA) * (bs[idx] - A @ x) / np.diag(A)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())


def test_restriction_conserves_and_prolongation_adds_constant():
    inp = PC.build_inputs(PC.CASES[2])
    o, _ = make(inp)
    rng = np.random.default_rng(7)
    for m in range(o.nmg - 1):
        fl = o.flags(m)
        r = PC.zero_nonfluid(rng.standard_normal(o.size(m)).astype(np.float32), fl)
        o.set(P.CH_R, m, r)
        o.fill(P.CH_B, m + 1, 0.0)
        o.restrict(m, P.CH_R, P.CH_B)
        bc = o.get(P.CH_B, m + 1)
        assert abs(bc.astype(np.float64).sum() - r.astype(np.float64).sum()) < 1e-4 * np.abs(r).sum()
        assert np.all(bc[(o.flags(m + 1) & FLUIDMASK) != 0] == 0)
        # constant correction
        o.fill(P.CH_X, m + 1, 0.0)
        o.set(P.CH_X, m + 1, np.where((o.flags(m + 1) & FLUIDMASK) == 0, 0.75, 0).astype(np.float32))
        o.fill(P.CH_X, m, 0.5)
        o.prolongate(m, P.CH_X, P.CH_X)
        x = o.get(P.CH_X, m)
        fluid = (fl & FLUIDMASK) == 0
        assert np.all(x[fluid] == 1.25) and np.all(x[~fluid] == 0.5)


@pytest.mark.parametrize("case", PC.CASES, ids=lambda c: c[0])
def test_vcycle_converges_to_manufactured_solution(case):
    o, inp = make(case)
    fl = o.flags(0)
    xs = PC.zero_nonfluid(PC.pages_random(o.size(0), 11), fl)
    o.set(P.CH_X, 0, xs)
    o.fill(P.CH_B, 0, 0.0)
    o.multiply(0, 0, chan_dst=P.CH_B)
    b = o.get(P.CH_B, 0)
    o.fill(P.CH_X, 0, 0.0)
    res = []
    for _ in range(6):
        o.vcycle(2, 2, 60)
        o.fill(P.CH_R, 0, 0.0)
        o.multiply(0, P.OPT_CALC_RESIDUAL)
        res.append(np.linalg.norm(o.get(P.CH_R, 0)) / np.linalg.norm(b))
    assert res[-1] < 2e-3 and res[-1] < 0.05 * res[0], res
    closed = all(v == P.DBC_SOLID for v in inp["dbc"]) and not (inp["cell_type"] == P.VOID).any()
    if not closed:  # a Dirichlet part fixes the solution
        err = np.abs(o.get(P.CH_X, 0) - xs).max() / np.abs(xs).max()
        assert err < 0.05, err


def test_invalid_hierarchies_rejected():
    t = np.zeros((16, 16, 32), np.uint8)
    with pytest.raises(ValueError):  # neighbours two levels apart
        P.PsOracle((2, 1, 1), 16, np.array([0, 2], np.int8), t)
    with pytest.raises(ValueError):  # block edge 32
        P.PsOracle((1, 1, 1), 32, np.array([0], np.int8), np.zeros((32, 32, 32), np.uint8))
    with pytest.raises(ValueError):  # dense levels need divisible block counts
        P.PsOracle((3, 1, 1), 8, np.zeros(3, np.int8), np.zeros((8, 8, 24), np.uint8), ndense=2)
