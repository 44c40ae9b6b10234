"""Precision study of the MCF sweep over storage formats (FP32 / FP16 / BF16
storage with FP32 arithmetic, UNORM16 as the reference format) against an
FP64 run of the same sweep on the GPU.

Every run starts from the same FP32 density field (splat + finalize into an
FP32 channel, then converted on the device). Every `every` sweeps the study
records, per storage: max and RMS difference to the FP64 field (accumulated
in FP64 on the device) and the FP64 RMS of the last sweep's update.

Run: python -m msbg_amd.studies [--size 1024 --sweeps 200 ...]; prints JSON.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import _lib as L
from .msbg import MCF_8COLOR, MultiresSparseGrid, scene_spheres

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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--objects", type=int, default=512)
    ap.add_argument("--points", type=int, default=16000)
    ap.add_argument("--sweeps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--exact", action="store_true")
    a = ap.parse_args(argv)
    res = precision_study(S=a.size, nObj=a.objects, nPts=a.points, sweeps=a.sweeps, every=a.every,
                          arith=L.ARITH_EXACT if a.exact else L.ARITH_FAST,
                          log=lambda s: print(s, file=sys.stderr, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
