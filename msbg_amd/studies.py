"""Studies of the MCF sweep on the GPU.

precision_study: the sweep over storage formats (FP32 / FP16 / BF16 storage
with FP32 arithmetic, UNORM16 as the reference format) against an FP64 run of
the same sweep on the GPU.

Every run starts from the same FP32 density field (splat + finalize into an
FP32 channel, then converted on the device). Every `every` sweeps the study
records, per storage: max and RMS difference to the FP64 field (accumulated
in FP64 on the device) and the FP64 RMS of the last sweep's update.

batched_study: many independent sparse grids of uneven size, level count and
active fraction swept together (one graph launch per call), against the same
grids swept one after another through applyChannelPdeFast.

Run: python -m msbg_amd.studies [precision|batched] [options]; prints JSON.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import _lib as L
from .msbg import MCF_8COLOR, Batch, MultiresSparseGrid, scene_spheres

CH_INIT, CH_F64 = 40, 41
STUDY = {"f32": (L.F32, 42), "f16": (L.F16, 43), "bf16": (L.BF16, 44), "u16": (L.U16, 45)}


def precision_study(S=1024, bsx=32, nObj=512, rmin=10.0, rmax=80.0, nPts=16000, sweeps=200,
                    every=20, dt=0.1, seed=4321, storages=("f32", "f16", "bf16", "u16"),
                    arith=L.ARITH_FAST, log=None):
    t0 = time.perf_counter()
    xyz = scene_spheres(seed, nObj, rmax + 8, S - rmax - 8, rmin, rmax, nPts, 0.5, 0.0)
    g = MultiresSparseGrid.create("precision", S, S, S, bsx, options=L.OPT_SINGLE_LEVEL, nLevels=3)
    try:
        g.points_upload(xyz)
        lv = g.regularizeRefinementMap(g.levels_from_points([4.0]))
        g.setRefinementMap(lv)
        g.alloc_demo(CH_INIT, L.F32)
        g.splat(CH_INIT, 2.0, 2.0)
        g.finalize(CH_INIT, 2.0, 2.0)
        g.copyDataTo(CH_INIT, CH_F64, L.F64)
        for name in storages:
            st, ch = STUDY[name]
            g.copyDataTo(CH_INIT, ch, st)
        levels, _ = g.getBlockInfo()
        nblk = int((levels == 0).sum())
        out = {"scene": dict(S=S, bsx=bsx, nObj=nObj, rmin=rmin, rmax=rmax, nPts=nPts, seed=seed,
                             level0_blocks=nblk, voxels=nblk * bsx ** 3),
               "dt": dt, "sweeps": sweeps, "every": every,
               "arith": "fast" if arith == L.ARITH_FAST else "exact", "rows": []}
        # initial quantisation error of each storage
        row = {"sweep": 0}
        for name in storages:
            mx, rms = g.diff(STUDY[name][1], CH_F64)
            row[name] = {"max": mx, "rms": rms}
        out["rows"].append(row)
        done = 0
        while done < sweeps:
            n = min(every, sweeps - done)
            n += n % 2  # the driver runs sweeps in pairs (msbg4.cpp:150-152)
            g.set_arith(L.ARITH_EXACT)
            r64 = g.applyChannelPdeFast(CH_F64, 0, 0, None, MCF_8COLOR, 1, 0, 0, 0, n, 1.0, dt,
                                        residual=True)
            g.set_arith(arith)
            row = {"sweep": done + n, "f64": {"residual": float(r64[-1])}}
            for name in storages:
                ch = STUDY[name][1]
                res = g.applyChannelPdeFast(ch, 0, 0, None, MCF_8COLOR, 1, 0, 0, 0, n, 1.0, dt,
                                            residual=True)
                mx, rms = g.diff(ch, CH_F64)
                row[name] = {"max": mx, "rms": rms, "residual": float(res[-1])}
            done += n
            out["rows"].append(row)
            if log:
                log(json.dumps(row))
        out["seconds"] = time.perf_counter() - t0
        return out
    finally:
        g.destroy()


RSCAN_ACT = 3.0  # activation distance of a block to the nearest particle, voxels


def _scene_grid(rng, S, bsx, nLevels, frac, nObj, chan, storage):
    """One sparse grid: nObj seeded spheres and ellipsoids sized so that the
    blocks within RSCAN_ACT of their surfaces cover about `frac` of the
    domain; level-0 blocks hold pages, the rest of the channel is EMPTY.
    Returns (grid, objects used)."""
    shell = 1.5 * bsx + 2 * RSCAN_ACT  # block-band thickness around a surface
    r = float(np.sqrt(frac * S ** 3 / (nObj * 4 * np.pi * shell)))
    seed = int(rng.integers(1 << 31))
    g = MultiresSparseGrid.create("batched", S, S, S, bsx, options=L.OPT_SINGLE_LEVEL,
                                  nLevels=nLevels)
    try:
        # the shell estimate misses block granularity; correct the radius (and,
        # below the smallest radius, the object count) from the measured
        # fraction a few times
        for _ in range(5):
            r = min(max(r, 3.0), S / 4.0)
            rmin, rmax = 0.7 * r, min(1.3 * r, S / 2.0 - 2.0)
            nPts = int(min(max(np.pi * rmax ** 2, 200), 200000))  # ~2-voxel point spacing
            xyz = scene_spheres(seed, nObj, rmax + 2, S - rmax - 2, rmin, rmax, nPts, 0.5, 0.5)
            g.points_upload(xyz)
            act = g.activate(RSCAN_ACT) > 0
            got = max(float(act.mean()), 1e-9)
            if abs(got / frac - 1.0) < 0.15:
                break
            if r <= 3.0 and got > frac and nObj > 1:
                nObj = max(1, int(nObj * frac / got))
            else:
                r *= float(np.sqrt(frac / got))
        if nLevels == 1:
            g.setRefinementMap(np.zeros(g.nBlocks(), np.int32))
            kind = np.where(act, L.KIND_VALUE, L.KIND_EMPTY).astype(np.uint8)
            g.channel_define(chan, storage, 0, kind, 0.0, None, 0.0)
        else:
            bands = [RSCAN_ACT] if nLevels == 2 else [RSCAN_ACT, RSCAN_ACT + 2 * bsx]
            lv = g.levels_from_points(bands)
            if nLevels > 2:
                lv = g.regularizeRefinementMap(lv)
            g.setRefinementMap(lv)
            g.alloc_demo(chan, storage)
        g.splat(chan, 2.0, 2.0)
        g.finalize(chan, 2.0, 2.0)
    except Exception:
        g.destroy()
        raise
    return g, nObj


def batched_study(nGrids=256, sizes=(128, 256, 512, 1024), blockSizes=(16, 32), levels=(1, 2, 3),
                  fracRange=(0.01, 0.20), objRange=(8, 128), sweeps=50, dt=0.1, seed=2024,
                  storage=L.F16, check=True, log=None):
    """Timings (ms) of `sweeps` sweeps over all grids: one grid at a time,
    batched eager launches, one batched graph launch, and the graph launch
    after the autotuner. With check=True every batched result is compared
    bitwise with the one-grid-at-a-time result on a copy of the channel."""
    CH, CHB = 50, 51
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    grids, meta = [], []
    try:
        for k in range(nGrids):
            S = int(rng.choice(sizes))
            bsx = int(rng.choice(blockSizes))
            nl = int(rng.choice(levels))
            frac = float(rng.uniform(*fracRange))
            nObj = int(rng.integers(objRange[0], objRange[1] + 1))
            g, nObj = _scene_grid(rng, S, bsx, nl, frac, nObj, CH, storage)
            grids.append(g)
            _, npg = g.pagemap(CH)
            meta.append(dict(S=S, bsx=bsx, nLevels=nl, frac_target=frac, nObj=nObj, pages=npg,
                             frac=npg / g.nBlocks(), voxels=npg * bsx ** 3))
            if log and (k + 1) % 32 == 0:
                log(json.dumps({"built": k + 1, "seconds": time.perf_counter() - t0}))
        vox = sum(m["voxels"] for m in meta)
        out = {"grids": nGrids, "sweeps": sweeps, "dt": dt, "seed": seed, "storage": storage,
               "voxels": vox, "setup_seconds": time.perf_counter() - t0,
               "frac": [float(np.min([m["frac"] for m in meta])),
                        float(np.mean([m["frac"] for m in meta])),
                        float(np.max([m["frac"] for m in meta]))],
               "by_size": {str(S): sum(m["S"] == S for m in meta) for S in sizes}}

        def same(tag):
            if not check:
                return
            for g in grids:
                mx, _ = g.diff(CH, CHB)
                assert mx == 0.0, (tag, mx)

        def timed(fn):
            for g in grids:
                g.synchronize()
            t = time.perf_counter()
            fn()
            for g in grids:
                g.synchronize()
            return (time.perf_counter() - t) * 1e3

        def sequential(chan):
            for g in grids:
                g.applyChannelPdeFast(chan, 0, 0, None, MCF_8COLOR, 1, 0, 0, 0, sweeps, 1.0, dt)

        def rate(ms):
            return vox * sweeps / (ms * 1e-3)

        for g in grids:
            g.copyDataTo(CH, CHB, storage)
        with Batch() as ba, Batch() as bb:
            for g in grids:
                ba.add(g, CH, dt, 1)
                bb.add(g, CHB, dt, 1)
            res = {}
            res["sequential_ms"] = timed(lambda: sequential(CHB))
            res["batch_eager_ms"] = timed(lambda: ba.run(sweeps, useGraph=False))
            same("eager")
            res["batch_graph_first_ms"] = timed(lambda: ba.run(sweeps, useGraph=True))
            bb.run(sweeps, useGraph=False)
            same("graph capture")
            res["batch_graph_ms"] = timed(lambda: ba.run(sweeps, useGraph=True))
            bb.run(sweeps, useGraph=False)
            same("graph replay")
            # one persistent launch per kernel group, per-grid dependency waits
            res["batch_persistent_first_ms"] = timed(
                lambda: ba.run(sweeps, useGraph=Batch.PERSISTENT))
            bb.run(sweeps, useGraph=False)
            same("persistent capture")
            res["batch_persistent_ms"] = timed(lambda: ba.run(sweeps, useGraph=Batch.PERSISTENT))
            bb.run(sweeps, useGraph=False)
            same("persistent replay")
            t = time.perf_counter()
            ba.autotune(passes=4)
            res["autotune_seconds"] = time.perf_counter() - t
            res["tiles"] = ba.tiles()
            res["persistent_by_group"] = ba.modes()
            res["batch_graph_tuned_ms"] = timed(lambda: ba.run(sweeps, useGraph=True))
            res["sequential_tuned_ms"] = timed(lambda: sequential(CH))
            for key in ("sequential_ms", "batch_eager_ms", "batch_graph_ms", "batch_persistent_ms",
                        "batch_graph_tuned_ms", "sequential_tuned_ms"):
                res[key.replace("_ms", "_updates_per_s")] = rate(res[key])
            res["speedup_graph_vs_sequential"] = res["sequential_ms"] / res["batch_graph_ms"]
            res["speedup_tuned_vs_sequential"] = res["sequential_ms"] / res["batch_graph_tuned_ms"]
            out.update(res)
        out["seconds"] = time.perf_counter() - t0
        out["grid_list"] = meta
        return out
    finally:
        for g in grids:
            g.destroy()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("study", nargs="?", default="precision", choices=("precision", "batched"))
    ap.add_argument("--grids", type=int, default=256)
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--objects", type=int, default=512)
    ap.add_argument("--points", type=int, default=16000)
    ap.add_argument("--sweeps", type=int, default=None,
                    help="default: 200 (precision), 50 (batched)")
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--exact", action="store_true")
    a = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    if a.study == "batched":
        print(json.dumps(batched_study(nGrids=a.grids, sweeps=a.sweeps or 50,
                                       check=not a.no_check, log=log)))
        return
    res = precision_study(S=a.size, nObj=a.objects, nPts=a.points, sweeps=a.sweeps or 200,
                          every=a.every, arith=L.ARITH_EXACT if a.exact else L.ARITH_FAST, log=log)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
