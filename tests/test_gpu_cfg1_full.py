"""cfg1 at its specified size: a 512^3 domain in 16^3 blocks, 64 seeded
spheres with radii in [20, 60] and centres in [64, 448]^3, 20 000 jittered
points each (1.28 M), rScan 4, one level of active blocks, then 20 MCF sweeps
with dt 0.1 in UNORM16 and in FP32 storage (msbg_demo.cpp:368-680,
msbg4.cpp:60-957). In exact arithmetic the GPU must equal the oracle bit for
bit: domain filter count, activation counts, level map, flags, active list,
splat and finalize codes, the swept field after all 20 sweeps and every
sweep's FP64 residual."""
import sys

import numpy as np
import pytest

from oracle import liboracle as O
from tests.helpers import gpu_pipeline, oracle_pipeline

pytestmark = pytest.mark.gpu
M = pytest.importorskip("msbg_amd")

S, BSX, RSCAN = 512, 16, 4.0


def note(*a):
    print("[cfg1 full]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg1_points():
    # the scene of bench.py's cfg1 workload (64 spheres, 20k points each)
    xyz = M.scene_spheres(1235, 64, 64.0, 448.0, 20.0, 60.0, 20000, 0.5, 0.0)
    assert xyz.shape == (1280000, 3)
    return xyz


def _oracle_f32(xyz):
    """The FP32 run of the oracle: the splat's distance codes finalized by the
    float rule (or_finalize_density_f), then 20 FP32 sweeps."""
    o = oracle_pipeline(S, BSX, xyz, rScan=RSCAN, do_sweep=False)
    nb = S // BSX
    codes = O.Level(S, S, S, BSX, o["pagemap"],
                    np.full((o["active"].size, BSX, BSX, BSX), 65535, np.uint16), O.U16)
    O.splat_u16(codes, o["xk"], RSCAN)
    dens = O.finalize_density_table(2.0, 2.0).astype(np.float32)[codes.pages]
    L = O.Level(S, S, S, BSX, o["pagemap"], np.ascontiguousarray(dens), O.F32)
    o["splat"] = dens.copy()
    o["res"] = O.mcf_sweeps(L, o["active"], o["blv"], 20, 0.1, 1, residual=True)
    o["level"] = L
    assert nb ** 3 == o["blv"].size
    return o


@pytest.mark.parametrize("storage", [M.U16, M.F32])
def test_cfg1_full_size_bitexact(cfg1_points, storage):
    xyz = cfg1_points
    note("oracle", storage)
    if storage == M.U16:
        o = oracle_pipeline(S, BSX, xyz, rScan=RSCAN, storage=O.U16, numIter=20, dt=0.1,
                            residual=True)
    else:
        o = _oracle_f32(xyz)
    assert o["active"].size > 5000
    note("gpu", storage, "active blocks", o["active"].size)
    g = gpu_pipeline(S, BSX, xyz, rScan=RSCAN, storage=storage, numIter=20, dt=0.1,
                     residual=True, arith=M.ARITH_EXACT)
    try:
        assert g["nk"] == int(o["keep"].sum())
        assert np.array_equal(g["cnt"], o["cnt"])
        assert np.array_equal(g["lv"], o["lv"])
        assert np.array_equal(g["blv"], o["blv"]) and np.array_equal(g["bfl"], o["bfl"])
        assert np.array_equal(g["active"], o["active"])
        assert np.array_equal(g["splat"].view(np.uint8), o["splat"].view(np.uint8))
        assert np.array_equal(g["pages"].view(np.uint8), o["level"].pages.view(np.uint8))
        assert len(g["res"]) == 20
        np.testing.assert_allclose(g["res"], o["res"], rtol=1e-12)
        note("bit-exact over 20 sweeps; residuals", g["res"][0], "...", g["res"][-1])
    finally:
        g["grid"].destroy()
