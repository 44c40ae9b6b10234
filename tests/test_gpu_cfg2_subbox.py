"""Parity on the benchmark scene itself: the centre 512^3 sub-box of the
4096^3 sphere/ellipsoid scene (2048 objects, 65.5 M points, seed 1234), run
as its own domain with 32^3 blocks and three levels by distance to the
particles (<= 8 voxels -> 0, <= 32 -> 1, else 2). GPU and oracle must agree
bit for bit on the level map, flags, active list, UNORM16 splat and
finalize, and on exact-mode FP16 sweeps over the workload's 100 sweeps (with
every sweep's FP64 residual); fast mode within the sweep tolerance."""
import sys

import numpy as np
import pytest

from oracle import liboracle as O

pytestmark = pytest.mark.gpu
M = pytest.importorskip("msbg_amd")

S, BSX, LO = 512, 32, 1792  # sub-box origin: (4096 - 512) / 2, block aligned
CH, CH16 = 30, 31


def note(*a):
    print("[cfg2 sub-box]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def subbox_points():
    # the benchmark's scene (bench.py build_scene for the 4096^3 workload)
    xyz = M.scene_spheres(1234, 2048, 208.0, 4096 - 208.0, 30.0, 200.0, 32000, 0.5, 0.5)
    p = xyz - np.float32(LO)
    m = np.all((p >= -8) & (p < S + 8), axis=1)  # a margin, so the domain filter is exercised
    note("points in the sub-box margin", int(m.sum()), "of", xyz.shape[0])
    return np.ascontiguousarray(p[m], np.float32)


def _oracle(p):
    nb = S // BSX
    keep = O.domain_filter(S, S, S, p)
    xk = np.ascontiguousarray(p[keep == 1])
    c8, _ = O.activate(BSX, nb, nb, nb, xk, 8.0)
    c32, _ = O.activate(BSX, nb, nb, nb, xk, 32.0)
    lv0 = np.where(c8 > 0, 0, np.where(c32 > 0, 1, 2)).astype(np.int32)
    lv, rc = O.regularize(nb, nb, nb, 3, lv0)
    assert rc == 0
    blv, bfl = O.block_flags(nb, nb, nb, 3, lv)
    active = np.nonzero(blv == 0)[0].astype(np.int32)
    pm = np.full(nb ** 3, O.EMPTY, np.int32)
    pm[active] = np.arange(active.size, dtype=np.int32)
    L = O.Level(S, S, S, BSX, pm, np.full((active.size, BSX, BSX, BSX), 65535, np.uint16), O.U16)
    O.splat_u16(L, xk, 4.0)
    O.finalize_u16(L.pages, 2.0, 2.0)
    return dict(nk=int(keep.sum()), lv=lv, blv=blv, bfl=bfl, active=active, pm=pm, splat=L.pages)


def test_cfg2_subbox_bitexact(subbox_points):
    p = subbox_points
    o = _oracle(p)
    note("level-0 blocks", o["active"].size)
    assert o["active"].size > 200
    with M.MultiresSparseGrid.create("sub", S, S, S, BSX, options=0, nLevels=3) as g:
        g.set_arith(M.ARITH_EXACT)
        assert g.points_upload(p) == o["nk"]
        lv = g.regularizeRefinementMap(g.levels_from_points([8.0, 32.0]))
        assert np.array_equal(lv, o["lv"])
        g.setRefinementMap(lv)
        blv, bfl = g.getBlockInfo()
        assert np.array_equal(blv, o["blv"]) and np.array_equal(bfl, o["bfl"])
        kind = np.where(blv == 0, 3, 1).astype(np.uint8)  # level 0 values, EMPTY elsewhere
        g.channel_define(CH, M.U16, 0, kind, 0.0, None, 65535.0)
        pm, _ = g.pagemap(CH)
        assert np.array_equal(pm, o["pm"])
        g.splat(CH, 2.0, 2.0)
        g.finalize(CH, 2.0, 2.0)
        assert np.array_equal(g.download(CH), o["splat"])
        note("splat + finalize bit-exact")
        # FP16 storage, FP32 arithmetic: the workload's 100 sweeps, exact
        g.copyDataTo(CH, CH16, M.F16)
        vals = o["splat"].astype(np.float32) * np.float32(1.0 / 65535.0)
        from tests.helpers import to_storage_level
        L16 = to_storage_level(S, BSX, o["pm"], vals, O.F16)
        assert np.array_equal(g.download(CH16).view(np.uint16), L16.pages)
        want = O.mcf_sweeps(L16, o["active"], o["blv"], 100, 0.1, 1, residual=True)
        got = g.applyChannelPdeFast(CH16, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 100, 1.0, 0.1,
                                    residual=True)
        assert np.array_equal(g.download(CH16).view(np.uint16), L16.pages)
        assert len(got) == 100
        np.testing.assert_allclose(got, want, rtol=1e-12)
        note("100 exact FP16 sweeps bit-exact; residuals", got[0], "...", got[-1])
        # fast arithmetic from the same start, 4 sweeps: within the sweep's
        # tolerance of the oracle's own 4 sweeps
        L4 = to_storage_level(S, BSX, o["pm"], vals, O.F16)
        O.mcf_sweeps(L4, o["active"], o["blv"], 4, 0.1, 1)
        g.copyDataTo(CH, CH16, M.F16)
        g.set_arith(M.ARITH_FAST)
        g.applyChannelPdeFast(CH16, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 4, 1.0, 0.1)
        fast = g.download(CH16).astype(np.float32)
        exact = L4.pages.view(np.float16).astype(np.float32)
        d = np.abs(fast - exact)
        # FP16 storage: one rounding step is up to 2^-11 of the value
        assert np.mean(d > 2e-3) < 1e-4 and np.sqrt(np.mean(d.astype(np.float64) ** 2)) < 1e-4
