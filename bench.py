"""Benchmark of the sparse-block MCF sweep on one MI355X.

Workload (default): effective 4096^3 domain, 32^3 blocks, 3 resolution levels
assigned by distance to the particles (<= 8 voxels -> 0, <= 32 -> 1, else 2,
then regularized), 2048 seeded spheres and ellipsoids with radii/semi-axes
uniform in [30, 200] and 32k jittered surface points each (~65.5M particles),
FP16 storage with FP32 arithmetic. Setup = upload, activation, refinement
map, splat, finalize, ghost blocks. A step is one call of the sweep driver
with numIter = 2, i.e. 2 sweeps of 8 colour passes over every level-0 block
(the reference runs sweeps in pairs: odd numIter rounds up).

Prints one JSON line. value = active level-0 voxel updates per second over
the timed steps (inputs resident in HBM). The roofline object prices the
colour-pass kernel at 2 * sizeof(storage) algorithmic bytes per voxel update
against MI355X's 8 TB/s, from HIP events recorded on the kernel's stream.
cpu_baseline times the same sweep on the host cores (CPU oracle built with
the reference's flags + OpenMP) on a bounded sub-box sample.

Beside the rate, the line carries the rest of the metric and the run's
precision statement: the other arithmetic mode's kernel rate on the same
pages (exact beside fast), error_vs_fp64 (the FP16 fast sweep against an FP64
exact run of the same field on the 512^3 centre sub-box, same sweep count)
and time_to_tol (a held-out 2048^3 scene solved to FP64 residual 1e-4 by
plain sweeps and by the learned warm start + plain sweeps).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (S, bsx, nLevels, nObj, rmin, rmax, nPts, ellipsoidFrac, bands)
    "cfg2": dict(S=4096, bsx=32, nLevels=3, nObj=2048, rmin=30.0, rmax=200.0, nPts=32000,
                 ell=0.5, bands=[8.0, 32.0],
                 desc="4096^3 effective, 32^3 blocks, 3 levels (<=8 / <=32 / else), 2048 spheres+"
                      "ellipsoids r in [30,200], 32k jittered points each"),
    "cfg1": dict(S=512, bsx=16, nLevels=3, nObj=64, rmin=20.0, rmax=60.0, nPts=20000, ell=0.0,
                 bands=[4.0], cmin=64.0, cmax=448.0,
                 desc="512^3 effective, 16^3 blocks, single level, 64 spheres r in [20,60], "
                      "20k points each"),
    "cfg3": dict(S=1024, bsx=32, nLevels=3, nObj=512, rmin=10.0, rmax=80.0, nPts=16000, ell=0.0,
                 bands=[4.0],
                 desc="1024^3 effective, 32^3 blocks, single level, 512 spheres r in [10,80], "
                      "16k points each"),
}

STORAGES = {"u16": 1, "f32": 2, "f16": 3, "bf16": 4, "f64": 5}

# HBM bytes per colour-pass launch / algorithmic bytes, measured with
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE on this workload, each
# scaled by the same counter's reading for k_copy16 (a copy of known size:
# FETCH_SIZE x 2.000, WRITE_SIZE x 1.000); profiles/k_sweep_cfg2_f16.md
TRAFFIC_RATIO = {("cfg2", "f16", "fast"): 1.168}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_scene(cfg, seed):
    import msbg_amd as M
    S = cfg["S"]
    cmin = cfg.get("cmin", cfg["rmax"] + 8.0)
    cmax = cfg.get("cmax", S - cfg["rmax"] - 8.0)
    return M.scene_spheres(seed, cfg["nObj"], cmin, cmax, cfg["rmin"], cfg["rmax"], cfg["nPts"],
                           0.5, cfg["ell"])


def setup_grid(cfg, xyz, storage, chan=30):
    """Demo phases on the GPU; returns the grid and setup timings (s)."""
    import msbg_amd as M
    S, bsx, nL = cfg["S"], cfg["bsx"], cfg["nLevels"]
    multi = len(cfg["bands"]) > 1
    t = {}
    g = M.MultiresSparseGrid.create("bench", S, S, S, bsx,
                                    options=0 if multi else M.OPT_SINGLE_LEVEL, nLevels=nL)
    t0 = time.perf_counter()
    nk = g.points_upload(xyz)
    g.synchronize()
    t["upload_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    lv = g.levels_from_points(cfg["bands"])
    lv = g.regularizeRefinementMap(lv)
    g.setRefinementMap(lv)
    t["levels_s"] = time.perf_counter() - t0
    levels, _ = g.getBlockInfo()
    if multi:
        # multiresolution channel: level-L blocks live on level L; coarse
        # levels far from the particles hold the empty value
        for level in range(nL):
            kind = np.zeros(g.nBlocks(), np.uint8)
            kind[levels == level] = 3 if level == 0 else 1
            g.channel_define(chan, storage, level, kind, 0.0, None, None)
        g.set_exists(None)
    else:
        g.alloc_demo(chan, storage)
    g.synchronize()
    t0 = time.perf_counter()
    g.splat(chan, 2.0, 2.0)
    g.synchronize()
    t["splat_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.finalize(chan, 2.0, 2.0)
    g.synchronize()
    t["finalize_s"] = time.perf_counter() - t0
    if multi and storage in (2, 3, 4):
        # the reference's prepareGhostBlocks handles float channels only
        t0 = time.perf_counter()
        g.prepareGhostBlocks(chan)
        g.synchronize()
        t["ghost_s"] = time.perf_counter() - t0
        # fine-coarse blocks whose ghost pages the pass computed
        t["ghost_blocks"] = [len(g.blocksFineCoarse(level)) for level in range(nL - 1)]
    n0 = int((levels == 0).sum())
    return g, nk, n0, t


def cpu_baseline(cfg, xyz, sweeps_budget_s=15.0):
    """The CPU oracle (reference flags + OpenMP) on a 512^3 sub-box sample."""
    from oracle import liboracle as O
    S, bsx = cfg["S"], cfg["bsx"]
    sub = min(512, S)
    lo = np.float32((S - sub) // 2 // bsx * bsx)
    p = xyz - lo
    m = np.all((p >= 0) & (p < sub), axis=1)
    p = np.ascontiguousarray(p[m], np.float32)
    nb = sub // bsx
    keep = O.domain_filter(sub, sub, sub, p)
    p = np.ascontiguousarray(p[keep == 1])
    cnt, _ = O.activate(bsx, nb, nb, nb, p, cfg["bands"][0])
    lv = O.levels_from_activity(cnt, 3)
    lv, _ = O.regularize(nb, nb, nb, 3, lv)
    blv, _ = O.block_flags(nb, nb, nb, 3, lv)
    active = np.nonzero(blv == 0)[0].astype(np.int32)
    pm = np.full(nb ** 3, O.EMPTY, np.int32)
    pm[active] = np.arange(active.size, dtype=np.int32)
    L = O.Level(sub, sub, sub, bsx, pm, np.full((active.size, bsx, bsx, bsx), 65535, np.uint16),
                O.U16)
    O.splat_u16(L, p, 4.0)
    O.finalize_u16(L.pages)
    threads = os.cpu_count() or 1
    nvox = active.size * bsx ** 3
    t0 = time.perf_counter()
    O.mcf_sweeps(L, active, blv, 2, 0.1, 1, nthreads=threads, kind="fast")
    t1 = time.perf_counter() - t0
    n_more = max(0, int(sweeps_budget_s / max(t1, 1e-3) / 2) * 2)
    n_more = min(n_more, 200)
    total_sweeps, total_t = 2, t1
    if n_more:
        t0 = time.perf_counter()
        O.mcf_sweeps(L, active, blv, n_more, 0.1, 1, nthreads=threads, kind="fast")
        total_t = time.perf_counter() - t0
        total_sweeps = n_more
    return {
        "value": nvox * total_sweeps / total_t,
        "unit": "voxel-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sub}^3 centre sub-box of the same scene: {active.size} active {bsx}^3 blocks "
                  f"({nvox} voxels), UNORM16 storage, {total_sweeps} 8-colour MCF sweeps dt=0.1, "
                  f"{total_t:.1f} s on {threads} threads (oracle built -O3 -ffast-math "
                  f"-march=x86-64-v3 -fopenmp)",
    }


def error_vs_fp64(cfg, xyz, sweeps, seed_note=""):
    """Error of the FP16-storage fast sweep against an FP64 exact run of the
    same field on the 512^3 centre sub-box (same level bands, 32^3 blocks):
    max and RMS difference after `sweeps` sweeps, accumulated in FP64 on the
    device, with both runs' FP64 residual of the last sweep."""
    import msbg_amd as M
    S, bsx = cfg["S"], cfg["bsx"]
    sub = min(512, S)
    lo = np.float32((S - sub) // 2 // bsx * bsx)
    p = xyz - lo
    p = np.ascontiguousarray(p[np.all((p >= 0) & (p < sub), axis=1)], np.float32)
    CH, CH64, CH16 = 30, 31, 32
    with M.MultiresSparseGrid.create("err", sub, sub, sub, bsx, options=0,
                                     nLevels=cfg["nLevels"]) as g:
        g.points_upload(p)
        lv = g.regularizeRefinementMap(g.levels_from_points(cfg["bands"]))
        g.setRefinementMap(lv)
        blv, _ = g.getBlockInfo()
        kind = np.where(blv == 0, 3, 1).astype(np.uint8)  # level 0 values, EMPTY elsewhere
        g.channel_define(CH, M.U16, 0, kind, 0.0, None, 65535.0)
        g.splat(CH, 2.0, 2.0)
        g.finalize(CH, 2.0, 2.0)
        g.copyDataTo(CH, CH64, M.F64)
        g.copyDataTo(CH, CH16, M.F16)
        q_max, q_rms = g.diff(CH16, CH64)
        g.set_arith(M.ARITH_EXACT)
        r64 = g.applyChannelPdeFast(CH64, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, sweeps, 1.0, 0.1,
                                    residual=True)
        g.set_arith(M.ARITH_FAST)
        r16 = g.applyChannelPdeFast(CH16, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, sweeps, 1.0, 0.1,
                                    residual=True)
        e_max, e_rms = g.diff(CH16, CH64)
        n0 = int((blv == 0).sum())
    return {"sample": f"{sub}^3 centre sub-box of the same scene, {n0} level-0 {bsx}^3 blocks, "
                      f"UNORM16 splat converted to FP16 (fast arithmetic) and FP64 (exact)",
            "sweeps": int(len(r16)), "max_abs": e_max, "rms": e_rms,
            "quantisation_max_abs": q_max, "quantisation_rms": q_rms,
            "residual_last_fp16": float(r16[-1]), "residual_last_fp64": float(r64[-1])}


def time_to_tol(params_path, seeds=(777,), S=2048, nObj=128):
    """The metric's second part: a held-out scene solved to FP64 residual 1e-4
    by plain sweeps and by the learned warm start + plain sweeps (the
    committed K = 1024 network), each with its device time, and the two
    solutions' distance beside the plain field's own drift in 300 more
    sweeps."""
    import msbg_amd as M
    from msbg_amd import warmstart as WS
    out = []
    with M.UNet(seed=0) as net:
        net.set_params(np.load(params_path, allow_pickle=False))
        for sd in seeds:
            r = WS.time_to_tolerance(net, S, nObj, sd)
            out.append({"seed": sd, "S": S, "objects": nObj, "level0_blocks": r["level0_blocks"],
                        "tol": r["tol"], "speedup": r["speedup"],
                        "plain_ms": 0.5 * (r["plain"]["ms"] + r["plain"]["ms_repeat"]),
                        "plain_sweeps": r["plain"]["sweeps"],
                        "learned_ms": r["learned"]["total_ms"],
                        "warm_start_ms": r["learned"]["warm_start_ms"],
                        "learned_sweeps": r["learned"]["sweeps"],
                        "fallback": r["learned"]["fallback"] is not None,
                        "learned_vs_plain": r["learned_vs_plain_solution"],
                        "plain_drift_300_sweeps": {k: r["plain_drift"][k]
                                                   for k in ("max_abs", "rms")}})
    return out


# RaymarchRenderer as msbg_demo.cpp:122-166 sets it up
RENDER_CAM = dict(camPos=(.5, .8, .7), lookAt=(.5, .5, .5), zoom=1.0, sunPos=(.5, 5, 5),
                  color=(0.8, 0.6, 0.4))
RENDER_W, RENDER_H = 3840, 2160  # camRes (msbg_demo.cpp:126)
FP64_PEAK_TFLOPS = 78.6


def render_check_and_cpu(g, chan, cfg, img, budget_s):
    """Oracle leg of the render path: bit-exact check of sampled rows of the
    GPU image (strict oracle), then the fast oracle timed on a centre band of
    rows sized to about budget_s seconds."""
    from oracle import liboracle as O
    S, bsx = cfg["S"], cfg["bsx"]
    pm, _ = g.pagemap(chan)
    L = O.Level(S, S, S, bsx, pm, g.download(chan), O.U16)
    cam = O.camera(RENDER_CAM["camPos"], RENDER_CAM["lookAt"], RENDER_CAM["zoom"],
                   RENDER_CAM["sunPos"], RENDER_CAM["color"])
    W, H = RENDER_W, RENDER_H
    rows = np.linspace(0, H - 1, 24).astype(int)
    bad = 0
    for y in rows:
        bad += int((O.render(L, cam, W, H, int(y), int(y) + 1) != img[y:y + 1]).sum())
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    y0, n = H // 2, 8
    t0 = time.perf_counter()
    O.render(L, cam, W, H, y0, y0 + n, nthreads=threads, kind="fast")
    t = time.perf_counter() - t0
    n2 = int(min(H // 2, max(n, n * budget_s / max(t, 1e-3))))
    y0 = (H - n2) // 2
    t0 = time.perf_counter()
    O.render(L, cam, W, H, y0, y0 + n2, nthreads=threads, kind="fast")
    t = time.perf_counter() - t0
    cpu = {"value": W * n2 / t, "unit": "pixels/s", "cores": threads, "kind": "port",
           "sample": f"rows {y0}..{y0 + n2 - 1} of the {W}x{H} frame ({W * n2} rays), {t:.1f} s on "
                     f"{threads} threads (oracle built -O3 -ffast-math -march=x86-64-v3 -fopenmp)"}
    return cpu, bad, int(rows.size * W)


def bench_render(args):
    """Field-consumer path: the demo's final 4K ray-march render of the
    smoothed UNORM16 field (msbg_demo.cpp:701-747)."""
    import msbg_amd as M
    cfg = CONFIGS[args.config or "cfg1"]
    t0 = time.perf_counter()
    xyz = build_scene(cfg, args.seed)
    log(f"scene {xyz.shape[0]} points in {time.perf_counter() - t0:.1f}s")
    g, nk, n0, tsetup = setup_grid(cfg, xyz, 1)  # RenderDensity = UNORM16
    chan = 30
    # the demo's PDE step: nIterPde = 5, dt 0.1 (msbg_demo.cpp:708-722)
    g.applyChannelPdeFast(chan, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 5, 1.0, 0.1)
    W, H = RENDER_W, RENDER_H
    img, cnt = g.render_counts(chan, W, H, **RENDER_CAM)
    for _ in range(args.warmup):
        g.render(chan, W, H, **RENDER_CAM)
    g.synchronize()
    kern_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, ms = g.render(chan, W, H, timing=True, **RENDER_CAM)
        kern_ms.append(ms)
    g.synchronize()
    wall = time.perf_counter() - t0
    assert np.array_equal(out, img)
    avg_ms = float(np.mean(kern_ms))
    value = W * H / (avg_ms * 1e-3)
    # per-launch work: FP64 march (3 FMAs per advanced position, 6 flops)
    # and the voxel gathers of the interpolated samples (8 per sample) and
    # hit normals (64)
    flops = 6.0 * cnt["advanced"]
    gather_bytes = 2.0 * (8 * cnt["interp"] + 64 * cnt["hits"]) + 4.0 * W * H
    tflops = flops / (avg_ms * 1e-3) / 1e12
    gbs = gather_bytes / (avg_ms * 1e-3) / 1e9
    cpu, bad, checked = None, None, None
    if not args.no_cpu:
        try:
            cpu, bad, checked = render_check_and_cpu(g, chan, cfg, out, args.cpu_seconds)
        except OSError as e:
            log(f"cpu baseline unavailable: {e}")
    res = {
        "metric": "pixels/s",
        "value": value,
        "unit": "pixels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config or 'cfg1'}: {cfg['desc']}; splat + finalize + 5 MCF "
                        f"iterations, then one {W}x{H} ray-march render with the demo camera "
                        f"(step = one frame)",
            "storage": "u16",
            "level0_blocks": n0,
            "rays": W * H,
            "march_positions": cnt["advanced"],
            "density_samples": cnt["samples"],
            "interpolated_samples": cnt["interp"],
            "hits": cnt["hits"],
            "frame_ms_with_copy": wall / args.steps * 1e3,
            "rows_checked_bitexact": None if checked is None else f"{checked - bad}/{checked} "
                                                                  f"pixels equal to the oracle",
        },
        "roofline": {
            "bound": "fp64",
            "achieved": tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": tflops / FP64_PEAK_TFLOPS,
            "traffic": None,
            "kernel": "k_render<u16> frame",
            "avg_launch_ms": avg_ms,
            "launches": args.steps,
            "note": "FP64 march FMAs against the 78.6 TFLOP/s FP64 peak; gathers "
                    f"{gbs:.0f} GB/s of 8000",
        },
        "cpu_baseline": cpu,
    }
    g.destroy()
    print(json.dumps(res))


def levels_cpu_baseline(g, chan, cfg, levels, level, opt, budget_s):
    """The fast oracle's or_mcf_sweeps_level on a sample of the same level's
    blocks (the level and the next coarser one are downloaded whole)."""
    from oracle import liboracle as O
    S, bsx = cfg["S"], cfg["bsx"]
    ost = {1: O.U16, 2: O.F32, 3: O.F16, 4: O.BF16}[g._storage[chan]]
    L = []
    for lv in (level, level + 1):
        if lv >= cfg["nLevels"]:
            L.append(None)
            continue
        pm, _ = g.pagemap(chan, lv)
        pages = g.download(chan, lv)
        if ost == O.F16:
            pages = pages.view(np.uint16)
        s = S >> lv
        L.append(O.Level(s, s, s, bsx >> lv, pm, pages, ost, 0.0, None))
    blocks = np.nonzero(levels == level)[0].astype(np.int32)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nv = (bsx >> level) ** 3
    take = blocks[::max(1, blocks.size // 256)]
    t0 = time.perf_counter()
    O.mcf_sweeps_level(L[0], L[1], take, levels, level, opt, 2, 0.1, 1, nthreads=threads,
                       kind="fast")
    t = time.perf_counter() - t0
    n = int(min(blocks.size, max(take.size, take.size * budget_s / max(t, 1e-3))))
    take = blocks[::max(1, blocks.size // n)]
    t0 = time.perf_counter()
    O.mcf_sweeps_level(L[0], L[1], take, levels, level, opt, 2, 0.1, 1, nthreads=threads,
                       kind="fast")
    t = time.perf_counter() - t0
    return {"value": take.size * nv * 2 / t, "unit": "voxel-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"level {level}: {take.size} of its {blocks.size} blocks ({bsx >> level}^3), "
                      f"2 sweeps with the same halo rule, {t:.1f} s on {threads} threads "
                      f"(oracle built -O3 -ffast-math -march=x86-64-v3 -fopenmp)"}


def bench_levels(args):
    """Coarse-level sweeps of the multiresolution scene: after the splat and
    prepareGhostBlocks, levels 1..n-1 are swept on their own grids with the
    COARSE_LEVEL|INTERPOLATE halo (the coarsest level with NEUMANN)."""
    import msbg_amd as M
    cfg = CONFIGS[args.config or "cfg2"]
    storage = STORAGES[args.storage]
    S, bsx, nL = cfg["S"], cfg["bsx"], cfg["nLevels"]
    xyz = build_scene(cfg, args.seed)
    chan = 30
    g = M.MultiresSparseGrid.create("levels", S, S, S, bsx, options=0, nLevels=nL)
    g.points_upload(xyz)
    lv = g.regularizeRefinementMap(g.levels_from_points(cfg["bands"]))
    g.setRefinementMap(lv)
    levels, _ = g.getBlockInfo()
    for level in range(nL):
        kind = np.where(levels == level, 3, 0).astype(np.uint8)
        g.channel_define(chan, storage, level, kind, 0.0, None, 0.0)
    g.set_exists(None)
    g.splat(chan, 2.0, 2.0)
    g.finalize(chan, 2.0, 2.0)
    g.prepareGhostBlocks(chan)
    g.set_arith(M.ARITH_FAST if args.arith == "fast" else M.ARITH_EXACT)
    esz = {1: 2, 2: 4, 3: 2, 4: 2}[storage]
    per, tot_vox, tot_ms, tot_wall = {}, 0, 0.0, 0.0
    for level in range(1, nL):
        opt = (M.OPT_BC_COARSE_LEVEL | M.OPT_INTERPOLATE) if level + 1 < nL else M.OPT_BC_NEUMANN
        nb = int((levels == level).sum())
        nvox = nb * (bsx >> level) ** 3
        for _ in range(args.warmup):
            g.applyChannelPdeLevel(chan, level, opt, None, 2, 0.1, 1)
        g.synchronize()
        kern, passes = 0.0, 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.applyChannelPdeLevel(chan, level, opt, None, 2, 0.1, 1)
            t, a, n = g.last_sweep_timing()
            kern += t
            passes += n
        g.synchronize()
        wall = time.perf_counter() - t0
        per[level] = {"blocks": nb, "block_edge": bsx >> level, "voxels": nvox,
                      "halo": "COARSE_LEVEL|INTERPOLATE" if level + 1 < nL else "NEUMANN",
                      "sweep_rate": nvox * 2 * args.steps / (kern * 1e-3),
                      "call_rate_incl_setup": nvox * 2 * args.steps / wall,
                      "avg_pass_ms": kern / max(passes, 1),
                      "hbm_frac": nvox * 2 * esz / 8 / (kern / max(passes, 1) * 1e-3) / 1e9 /
                                  HBM_PEAK_GBS}
        log(f"level {level}: {per[level]}")
        tot_vox += nvox * 2 * args.steps
        tot_ms += kern
        tot_wall += wall
    value = tot_vox / (tot_ms * 1e-3)
    lvl1 = per[1]
    achieved = lvl1["voxels"] * 2 * esz / 8 / (lvl1["avg_pass_ms"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        try:
            cpu = levels_cpu_baseline(g, chan, cfg, levels, 1,
                                      M.OPT_BC_COARSE_LEVEL | M.OPT_INTERPOLATE,
                                      args.cpu_seconds)
        except OSError as e:
            log(f"cpu baseline unavailable: {e}")
    res = {
        "metric": "voxel-updates/s/sweep",
        "value": value,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tot_wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config or 'cfg2'}: {cfg['desc']}; coarse-level sweeps of levels "
                        f"1..{nL - 1} after splat + prepareGhostBlocks; step = 2 sweeps per level",
            "storage": args.storage, "arith": args.arith,
            "levels": {str(k): v for k, v in per.items()},
            "rate_incl_call_setup": tot_vox / tot_wall,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"k_sweep<{bsx >> 1},{args.storage}> level-1 colour pass",
                     "avg_launch_ms": lvl1["avg_pass_ms"]},
        "cpu_baseline": cpu,
    }
    g.destroy()
    print(json.dumps(res))


def pressure_cpu_baseline(G, budget_s):
    """The oracle's relax (single-threaded C, ps_oracle.c) on the same tank
    scene at 128^3 (8^3 blocks of 16^3): relax iterations timed within the
    budget."""
    from oracle import ps as P
    nb, bsx0 = (8, 8, 8), 16
    lev, typ = G.tank_scene(nb, bsx0, seed=7)
    o = P.PsOracle(nb, bsx0, lev, typ, dbc=(0, 0, 1, 1, 1, 0), ndense=2)
    fl = o.flags(0)
    rng = np.random.default_rng(3)
    x = np.where((fl & (P.CELL_SOLID | P.CELL_VOID)) == 0,
                 rng.standard_normal(o.size(0)), 0).astype(np.float32)
    o.set(P.CH_X, 0, x)
    o.multiply(0, 0, chan_dst=P.CH_B)
    o.fill(P.CH_X, 0, 0.0)
    bf = o.block_flags(0)
    L = P.block_levels_at(lev, 0)
    cells = int(sum((bsx0 >> int(L[b])) ** 3 for b in range(L.size) if not bf[b] & P.BLK_NO_FLUID))
    t0 = time.perf_counter()
    o.relax(0, 1)
    t1 = time.perf_counter() - t0
    n = max(1, int(budget_s / max(t1, 1e-3)))
    t0 = time.perf_counter()
    for _ in range(n):
        o.relax(0, 1)
    t = time.perf_counter() - t0
    o.close()
    return {"value": cells * n / t, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"tank scene 128^3 (8^3 blocks of 16^3, levels {np.bincount(lev, minlength=4).tolist()}),"
                      f" {cells} cells per relax iteration, {n} iterations in {t:.1f} s on 1 thread "
                      f"(oracle built -O2 -ffp-contract=off)"}


def bench_pressure(args):
    """Adaptive pressure relax (msbg3.cpp relax()): one step = one relax
    iteration at multigrid level 0 (red then black blocks, every block with
    fluid relaxed by 5 lexicographic sub-iterations), on a 512^3 tank."""
    from msbg_amd import pressure as G
    nb, bsx0 = (32, 32, 32), 16
    t0 = time.perf_counter()
    lev, typ = G.tank_scene(nb, bsx0, seed=args.seed)
    g = G.PressureGrid(nb, bsx0, lev, typ, dbc=(0, 0, 1, 1, 1, 0), dense_levels=3)
    tsetup = time.perf_counter() - t0
    fl = g.flags(0)
    fluid = (fl & (G.CELL_SOLID | G.CELL_VOID)) == 0
    rng = np.random.default_rng(args.seed)
    xs = np.where(fluid, rng.standard_normal(g.size(0)), 0).astype(np.float32)
    g.upload(G.CH_X, 0, xs)
    g.fill(G.CH_B, 0, 0.0)
    g.multiplyLaplacianMatrixOpt(0, 0, G.CH_X, G.CH_B, 0.0, G.CH_B)  # b = A x*
    g.fill(G.CH_X, 0, 0.0)
    cells = g.relaxed_cells(0)
    log(f"pressure setup {tsetup:.1f}s, levels {np.bincount(lev, minlength=4).tolist()}, "
        f"{cells} cells per relax iteration, {g.nmg} multigrid levels")
    for _ in range(args.warmup):
        g.relax(0, 0, 0, G.CH_X, G.CH_B, G.CH_T, 1)
    g.synchronize()
    kern = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.relax(0, 0, 0, G.CH_X, G.CH_B, G.CH_T, 1)
        kern += g.last_timing()
    g.synchronize()
    wall = time.perf_counter() - t0
    relax_ms = kern / args.steps
    # Laplacian multiply with residual, and V-cycles on the manufactured problem
    mul = 0.0
    for _ in range(5):
        g.multiplyLaplacianMatrixOpt(G.OPT_CALC_RESIDUAL, 0, G.CH_X, G.CH_B, 0.0, G.CH_R)
        mul += g.last_timing()
    g.fill(G.CH_X, 0, 0.0)
    b = g.download(G.CH_B, 0).astype(np.float64)
    res, vms = [], 0.0
    for _ in range(4):
        g.vcycle(2, 2, 30)
        vms += g.last_timing()
        g.fill(G.CH_R, 0, 0.0)
        g.multiplyLaplacianMatrixOpt(G.OPT_CALC_RESIDUAL, 0, G.CH_X, G.CH_B, 0.0, G.CH_R)
        res.append(float(np.linalg.norm(g.download(G.CH_R, 0)) / np.linalg.norm(b)))
    err = float(np.abs(g.download(G.CH_X, 0) - xs).max() / np.abs(xs).max())
    # algorithmic HBM bytes of a relax iteration per cell: X read + write,
    # B, D (4 bytes each), flags (2)
    achieved = cells * 18 / (relax_ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        try:
            cpu = pressure_cpu_baseline(G, args.cpu_seconds)
        except OSError as e:
            log(f"cpu baseline unavailable: {e}")
    res_out = {
        "metric": "cell-updates/s/relax-iteration",
        "value": cells / (relax_ms * 1e-3),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "pressure relax: 512^3 tank (32^3 blocks of 16^3; fluid to 78% height, "
                        "void above, seeded solid sphere r = 0.17*512), block levels 0-3 by "
                        "distance to the fluid/solid/void band, + 3 dense multigrid levels; "
                        "solid floor and y walls, open x ends and top; step = relax() at level 0 "
                        "with 5 block sub-iterations, omega 6/7",
            "levels": np.bincount(lev, minlength=4).tolist(),
            "cells_per_iteration": cells,
            "relax_ms": relax_ms,
            "multiply_residual_ms": mul / 5,
            "multiply_rate": cells / (mul / 5 * 1e-3),
            "vcycle_ms": vms / 4,
            # |b - A x| / |b| after each V-cycle (2+2 relax iterations, 30 on the
            # coarsest level) from x = 0; FP32 storage floors it near 1e-5
            "vcycle_residual": res,
            "vcycle_max_err_after_4": err,
            "setup_s": tsetup,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_ps_block<16,relax> (+ the 8^3/4^3/2^3 launches of the same "
                               "iteration): 18 B per cell per iteration; the 5 sub-iterations "
                               "run in LDS, the bound is the 4x32-step wavefront",
                     "avg_launch_ms": relax_ms},
        "cpu_baseline": cpu,
    }
    g.close()
    print(json.dumps(res_out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--path", default="sweep", choices=["sweep", "render", "levels", "pressure"],
                    help="sweep: the 8-colour MCF sweep (default); render: the demo's 4K "
                         "ray-march render of the smoothed field; levels: coarse-level sweeps "
                         "of the multiresolution scene; pressure: the adaptive pressure relax")
    ap.add_argument("--config", default=None, choices=sorted(CONFIGS),
                    help="default cfg2 (sweep), cfg1 (render)")
    ap.add_argument("--storage", default="f16", choices=sorted(STORAGES))
    ap.add_argument("--arith", default="fast", choices=["fast", "exact"])
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--other-steps", type=int, default=3,
                    help="steps of the other arithmetic mode timed after the run")
    ap.add_argument("--no-error", action="store_true", help="skip the FP64 error run")
    ap.add_argument("--no-tol", action="store_true", help="skip the time-to-tolerance run")
    ap.add_argument("--calib", action="store_true",
                    help="more copy-kernel launches for profiler byte-counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.path == "render":
        bench_render(args)
        return
    if args.path == "levels":
        bench_levels(args)
        return
    if args.path == "pressure":
        bench_pressure(args)
        return
    args.config = args.config or "cfg2"
    import msbg_amd as M
    cfg = CONFIGS[args.config]
    storage = STORAGES[args.storage]
    t0 = time.perf_counter()
    xyz = build_scene(cfg, args.seed)
    log(f"scene {xyz.shape[0]} points in {time.perf_counter() - t0:.1f}s")
    g, nk, n0, tsetup = setup_grid(cfg, xyz, storage)
    g.set_arith(M.ARITH_FAST if args.arith == "fast" else M.ARITH_EXACT)
    bsx = cfg["bsx"]
    nvox = n0 * bsx ** 3
    log(f"setup {tsetup}, kept {nk} particles, {n0} level-0 blocks, {nvox} voxels")

    chan = 30
    for _ in range(args.warmup):
        g.applyChannelPdeFast(chan, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 2, 1.0, 0.1)
    g.synchronize()
    kern_ms, npasses = 0.0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.applyChannelPdeFast(chan, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 2, 1.0, 0.1)
        tot, avg, n = g.last_sweep_timing()
        kern_ms += tot
        npasses += n
    g.synchronize()
    wall = time.perf_counter() - t0
    sweeps = 2 * args.steps
    value = nvox * sweeps / wall
    esz = {1: 2, 2: 4, 3: 2, 4: 2, 5: 8}[storage]
    bytes_per_pass_avg = nvox * 2 * esz / 8.0
    avg_pass_ms = kern_ms / max(npasses, 1)
    achieved = bytes_per_pass_avg / (avg_pass_ms * 1e-3) / 1e9
    res = g.applyChannelPdeFast(chan, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 2, 1.0, 0.1,
                                residual=True)
    log(f"residual after {args.warmup * 2 + sweeps + 2} sweeps: {res}")
    # the other arithmetic's kernel rate on the same pages (device time)
    other = "exact" if args.arith == "fast" else "fast"
    g.set_arith(M.ARITH_EXACT if other == "exact" else M.ARITH_FAST)
    o_ms, o_n = 0.0, 0
    for _ in range(args.other_steps):
        g.applyChannelPdeFast(chan, 0, 0, None, M.MCF_8COLOR, 1, 0, 0, 0, 2, 1.0, 0.1)
        tot, _, n = g.last_sweep_timing()
        o_ms += tot
        o_n += n
    other_rate = nvox * 2 * args.other_steps / (o_ms * 1e-3) if o_ms > 0 else None
    log(f"{other} arithmetic: {other_rate:.4g} voxel-updates/s, "
        f"{o_ms / max(o_n, 1):.3f} ms per colour pass")
    g.destroy()
    err = None
    if not args.no_error and storage != 5:
        t0 = time.perf_counter()
        err = error_vs_fp64(cfg, xyz, args.warmup * 2 + sweeps + 2)
        log(f"error vs FP64 ({time.perf_counter() - t0:.1f}s): {err}")
    ttt = None
    if not args.no_tol:
        t0 = time.perf_counter()
        ttt = time_to_tol(os.path.join(ROOT, "msbg_amd", "data", "unet_k1024.npy"))
        log(f"time to tolerance ({time.perf_counter() - t0:.1f}s): {ttt}")
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(cfg, xyz, args.cpu_seconds)
        except OSError as e:
            log(f"cpu baseline unavailable: {e}")
    ST_NAME = {1: "u16", 2: "f32", 3: "f16", 4: "bf16", 5: "f64"}
    out = {
        "metric": "voxel-updates/s/sweep",
        "value": value,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if storage == 5 else "f32",
        "storage_dtype": ST_NAME[storage],
        "compute_dtype": "f64" if storage == 5 else "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}: {cfg['desc']}; step = 2 sweeps of 8 colour passes",
            "storage": args.storage,
            "arith": args.arith,
            f"{other}_arith_rate": other_rate,
            f"{other}_arith_avg_pass_ms": o_ms / max(o_n, 1) if o_n else None,
            "particles": int(xyz.shape[0]),
            "particles_kept": int(nk),
            "level0_blocks": n0,
            "active_voxels": int(nvox),
            "sweeps_timed": sweeps,
            "hbm_fraction": value * 2 * esz / (HBM_PEAK_GBS * 1e9),
            "setup_s": {k: round(v, 4) for k, v in tsetup.items() if k.endswith("_s")},
            "splat_particles_per_s": nk / tsetup["splat_s"] if tsetup.get("splat_s") else None,
            "ghost_blocks": tsetup.get("ghost_blocks"),
            "warm_start_network": "3-stage U-Net of the learned warm start trained with horizon "
                                  "K = 1024 on 2000 scenes (cfg5 as written, K = 64 on 10k "
                                  "scenes, gives 0.91x)",
        },
        "error_vs_fp64": err,
        "time_to_tol": ttt,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (TRAFFIC_RATIO[(args.config, args.storage, args.arith)] * bytes_per_pass_avg
                        if (args.config, args.storage, args.arith) in TRAFFIC_RATIO else None),
            "kernel": f"k_sweep<{bsx},{args.storage}> colour pass",
            "avg_launch_ms": avg_pass_ms,
            "launches": npasses,
        },
        "cpu_baseline": cpu,
    }
    # measured copy bandwidth on this box (1 GiB read + 1 GiB written per
    # launch): the reference for the profiler's byte counters
    copy_gbs = M.copy_bandwidth(1 << 30, 5 if args.calib else 3)
    out["roofline"]["copy_peak_measured"] = copy_gbs
    out["roofline"]["frac_of_copy"] = achieved / copy_gbs
    print(json.dumps(out))


if __name__ == "__main__":
    main()
