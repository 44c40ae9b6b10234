"""Learned warm start of the MCF solve: training data from the plain solver,
training on the GPU, and time to tolerance of plain versus learned + plain.

- Scenes: seeded spheres plus Perlin-displaced ellipsoids (radii [10, 150]),
  splatted and finalized into an FP32 channel of 32^3 blocks, 3 levels by
  distance to the particles (level 0 within 4 voxels).
- Training pairs: a level-0 block's 36^3 window of the splatted field (BF16)
  and the block's change over K plain sweeps (FP32). The command line
  defaults to K = 1024 on 2000 scenes; make_training_set defaults to K = 64.
  Scenes are swept in groups through one batched launch.
- Tolerance: the FP64 RMS of one sweep's update over the swept voxels (the
  residual applyChannelPdeFast returns), checked after every sweep.
- Learned + plain: msbg_warm_start adds the predicted change to every
  level-0 block, then plain sweeps run to the tolerance. If the residual
  rises between two checks, the field is restored from the splat and the
  plain solve runs from there (the time spent counts). The two solutions
  are compared (max and RMS difference) at the end.

Run: python -m msbg_amd.warmstart [--scenes N ...]; prints JSON.
The sample sets (windows, targets) live in library-owned device memory
(DeviceArray); training gathers each batch's rows on the device and queues
its steps without waiting, reading the loss once per epoch.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import _lib as L
from .msbg import MCF_8COLOR, Batch, DeviceArray, MultiresSparseGrid, UNet

CH_IN, CH_K, CH_RUN, CH_PLAIN = 60, 61, 62, 63
RSCAN = 4.0
LAYERS = [(1, 16, 36), (16, 16, 34), (16, 32, 16), (32, 32, 16), (32, 64, 8), (64, 64, 8),
          (96, 32, 16), (32, 32, 16), (48, 16, 32), (16, 16, 32)]  # (cin, cout, din)
PAD = [0, 0, 1, 1, 1, 1, 1, 1, 1, 1]


def macs_per_block():
    """Multiply-adds of one forward pass (convolutions and the head)."""
    m = 0
    for (cin, cout, din), pad in zip(LAYERS, PAD):
        dout = din + 2 * pad - 2
        m += dout ** 3 * 27 * cin * cout
    return m + 32 ** 3 * 16


def build_scene(seed, S, nObj, storage=L.F32, chan=CH_IN, rmin=10.0, rmax=150.0, nPts=40000,
                perlinAmp=0.2, perlinFreq=1.0 / 48):
    """Half spheres, half Perlin-displaced ellipsoids (points generated on
    the device), splatted and finalized into `chan`."""
    g = MultiresSparseGrid.create("warmstart", S, S, S, 32, options=L.OPT_SINGLE_LEVEL, nLevels=3)
    try:
        ns = nObj // 2
        m = rmax * (1 + perlinAmp) + 8
        g.points_generate(seed, ns, m, S - m, rmin, rmax, nPts, 0.5, 0.0)
        g.points_generate(seed + 7919, nObj - ns, m, S - m, rmin, rmax, nPts, 0.5, 1.0, perlinAmp,
                          perlinFreq, append=True)
        lv = g.regularizeRefinementMap(g.levels_from_points([RSCAN, RSCAN + 32]))
        g.setRefinementMap(lv)
        g.alloc_demo(chan, storage)
        g.splat(chan, 2.0, 2.0)
        g.finalize(chan, 2.0, 2.0)
    except Exception:
        g.destroy()
        raise
    return g


def level0_blocks(g, chan):
    pm, _ = g.pagemap(chan)
    levels, _ = g.getBlockInfo()
    return np.nonzero((pm >= 0) & (levels == 0))[0].astype(np.int32)


def make_training_set(nScenes=10000, perScene=8, K=64, S=1024, seed=100, group=16,
                      objRange=(8, 40), dt=0.1, log=None):
    """(X, T, stats): device sets of BF16 windows [N,36,36,36] (uint16 bits)
    and FP32 changes [N,32,32,32]."""
    N = nScenes * perScene
    X = DeviceArray((N, 36, 36, 36), np.uint16)
    T = DeviceArray((N, 32, 32, 32), np.float32)
    after = DeviceArray((perScene, 32, 32, 32), np.float32)
    base = DeviceArray((perScene, 32, 32, 32), np.float32)
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    n, blocks, sweep_ms = 0, 0, 0.0
    for s0 in range(0, nScenes, group):
        grids = []
        try:
            for s in range(s0, min(s0 + group, nScenes)):
                grids.append(build_scene(seed * 1000003 + s, S, int(rng.integers(*objRange))))
            for g in grids:
                g.copyDataTo(CH_IN, CH_K, L.F32)
            with Batch() as b:
                for g in grids:
                    b.add(g, CH_K, dt, 1)
                sweep_ms += b.run(K, useGraph=False)
            for g in grids:
                cand = level0_blocks(g, CH_IN)
                blocks += cand.size
                pick = np.sort(rng.choice(cand, perScene, replace=cand.size < perScene))
                UNet.gather(g, CH_IN, pick, X.rows(n, perScene))
                UNet.extract(g, CH_K, pick, after)
                UNet.extract(g, CH_IN, pick, base)
                # the change over K sweeps, an FP32 difference
                T.rows(n, perScene).upload(after.numpy() - base.numpy())
                n += perScene
        finally:
            for g in grids:
                g.destroy()
        if log and (s0 // group) % 16 == 0:
            log(json.dumps({"scenes": s0 + len(grids), "seconds": time.perf_counter() - t0}))
    after.free()
    base.free()
    return X, T, {"scenes": nScenes, "samples": N, "K": K, "S": S,
                  "mean_level0_blocks": blocks / nScenes, "sweep_ms": sweep_ms,
                  "seconds": time.perf_counter() - t0}


def train(net, X, T, epochs=4, batch=32, lr=1e-3, val_frac=0.05, seed=0, log=None):
    """Adam on the resident sets: every batch's rows are gathered on the
    device and its step is queued; the loss is read once per epoch."""
    N = X.shape[0]
    nval = max(batch, int(N * val_frac))
    rng = np.random.default_rng(seed)
    perm = rng.permutation(N).astype(np.int32)
    val, tr = perm[:nval], perm[nval:]
    ntr = (tr.size // batch) * batch
    steps = epochs * (ntr // batch)
    hist, step = [], 0
    nv = (nval // batch) * batch
    tval = np.stack([T.rows(int(i)).numpy()[0] for i in val[:nv]])
    y = DeviceArray((batch, 32, 32, 32), np.float32)
    t0 = time.perf_counter()

    def val_loss():
        se, zero = 0.0, 0.0
        for i in range(0, nv, batch):
            net.forward_rows(X, val[i:i + batch], y)
            d = y.numpy().astype(np.float64) - tval[i:i + batch]
            se += float((d * d).sum())
            zero += float((tval[i:i + batch].astype(np.float64) ** 2).sum())
        return se / max(zero, 1e-30)

    rel0 = val_loss()
    net.take_loss()
    for ep in range(epochs):
        order = rng.permutation(tr)[:ntr]
        for i in range(0, ntr, batch):
            # cosine decay to lr/10
            a = lr * (0.55 + 0.45 * np.cos(np.pi * step / max(steps, 1)))
            net.train_step_rows(X, T, order[i:i + batch], float(a))
            step += 1
        lsum, nsteps = net.take_loss()
        rel = val_loss()
        hist.append({"epoch": ep + 1, "train_mse": lsum / max(nsteps, 1), "val_rel_mse": rel,
                     "seconds": time.perf_counter() - t0})
        if log:
            log(json.dumps(hist[-1]))
    secs = time.perf_counter() - t0
    y.free()
    return {"samples": N, "val": nval, "epochs": epochs, "batch": batch, "lr": lr, "steps": steps,
            "val_rel_mse_initial": rel0, "history": hist, "seconds": secs,
            "train_tflops": 3 * 2 * macs_per_block() * steps * batch / secs / 1e12}


def solve_to_tol(g, chan, tol, dt, maxSweeps, chunk=20, watch_rise=False):
    """Plain sweeps until the per-sweep residual drops below tol.
    Returns (sweeps, device ms, residual history, rose)."""
    sweeps, ms, hist = 0, 0.0, []
    prev = None
    while sweeps < maxSweeps:
        r = g.applyChannelPdeFast(chan, 0, 0, None, MCF_8COLOR, 1, 0, 0, 0, chunk, 1.0, dt,
                                  residual=True)
        t, _, _ = g.last_sweep_timing()
        hit = np.nonzero(r < tol)[0]
        if hit.size:
            k = int(hit[0]) + 1
            return sweeps + k, ms + t * k / len(r), hist + [float(r[k - 1])], False
        sweeps += len(r)
        ms += t
        hist.append(float(r[-1]))
        m = float(np.mean(r))
        if watch_rise and prev is not None and m > prev:
            return sweeps, ms, hist, True
        prev = m
    return sweeps, ms, hist, False


def time_to_tolerance(net, S=2048, nObj=128, seed=777, tol=1e-4, dt=0.1, maxSweeps=20000,
                      batch=256, drift=300, log=None):
    g = build_scene(seed, S, nObj)
    try:
        nblk = int(level0_blocks(g, CH_IN).size)
        g.set_arith(L.ARITH_FAST)
        g.copyDataTo(CH_IN, CH_RUN, L.F32)
        p_sweeps, p_ms, p_hist, _ = solve_to_tol(g, CH_RUN, tol, dt, maxSweeps)
        if log:
            log(json.dumps({"plain_sweeps": p_sweeps, "plain_ms": p_ms}))
        g.copyDataTo(CH_RUN, CH_PLAIN, L.F32)  # the plain solution, for comparison
        # the scale of "equal accuracy": how far the plain field itself still
        # moves in `drift` more sweeps once it is below the tolerance (the
        # flow has no fixed point; the residual criterion stops it)
        g.applyChannelPdeFast(CH_RUN, 0, 0, None, MCF_8COLOR, 1, 0, 0, 0, drift, 1.0, dt)
        drift_max, drift_rms = g.diff(CH_RUN, CH_PLAIN)
        g.copyDataTo(CH_IN, CH_RUN, L.F32)
        ws_ms = net.warm_start(g, CH_RUN, batch=batch)
        l_sweeps, l_ms, l_hist, rose = solve_to_tol(g, CH_RUN, tol, dt, maxSweeps, watch_rise=True)
        fb = None
        if rose:  # fallback: restore the splat and solve plainly
            g.copyDataTo(CH_IN, CH_RUN, L.F32)
            fb_sweeps, fb_ms, _, _ = solve_to_tol(g, CH_RUN, tol, dt, maxSweeps)
            fb = {"sweeps": fb_sweeps, "ms": fb_ms}
        learned_ms = ws_ms + l_ms + (fb["ms"] if fb else 0.0)
        # both solutions stop at the same FP64 residual; how far apart are they?
        d_max, d_rms = g.diff(CH_RUN, CH_PLAIN)
        # the plain solve again, so clock or cache drift shows up in the pair
        g.copyDataTo(CH_IN, CH_RUN, L.F32)
        p2_sweeps, p2_ms, _, _ = solve_to_tol(g, CH_RUN, tol, dt, maxSweeps)
        plain_ms = 0.5 * (p_ms + p2_ms)
        flops = 2.0 * macs_per_block() * nblk
        out = {"S": S, "nObj": nObj, "seed": seed, "tol": tol, "level0_blocks": nblk,
               "plain": {"sweeps": p_sweeps, "ms": p_ms, "ms_repeat": p2_ms,
                         "residuals_every_20": p_hist[:50]},
               "learned": {"warm_start_ms": ws_ms, "sweeps": l_sweeps, "sweep_ms": l_ms,
                           "total_ms": learned_ms, "fallback": fb,
                           "residuals_every_20": l_hist[:50]},
               "sweeps_saved": p_sweeps - l_sweeps,
               "learned_vs_plain_solution": {"max_abs": d_max, "rms": d_rms},
               "plain_drift": {"sweeps": drift, "max_abs": drift_max, "rms": drift_rms},
               "speedup": plain_ms / learned_ms,
               "cnn": {"tflops": flops / (ws_ms * 1e-3) / 1e12,
                       "frac_of_bf16_peak": flops / (ws_ms * 1e-3) / 2.5e15,
                       "ms_per_block": ws_ms / nblk,
                       "sweep_equivalents": ws_ms / (plain_ms / p_sweeps)}}
        if log:
            log(json.dumps({k: out[k] for k in ("speedup", "sweeps_saved",
                                                "learned_vs_plain_solution", "plain_drift",
                                                "cnn")}))
        return out
    finally:
        g.destroy()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--per-scene", type=int, default=8)
    ap.add_argument("--K", type=int, default=1024)
    ap.add_argument("--train-size", type=int, default=1024, help="edge of the training scenes")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-size", type=int, default=2048)
    ap.add_argument("--eval-objects", type=int, default=128)
    ap.add_argument("--eval-seeds", type=int, nargs="*", default=[777, 778])
    ap.add_argument("--seed", type=int, default=100)
    ap.add_argument("--save", default=None, help="write the trained parameters (.npy, FP32)")
    ap.add_argument("--load", default=None,
                    help="evaluate parameters from a .npy file instead of training")
    a = ap.parse_args(argv)

    def log(s):
        print(s, file=sys.stderr, flush=True)

    ds = tr = None
    with UNet(seed=a.seed) as net:
        if a.load:
            net.set_params(np.load(a.load, allow_pickle=False))
        else:
            X, T, ds = make_training_set(a.scenes, a.per_scene, a.K, S=a.train_size, seed=a.seed,
                                         log=log)
            log(json.dumps({"dataset": ds}))
            tr = train(net, X, T, a.epochs, a.batch, a.lr, seed=a.seed, log=log)
            X.free()
            T.free()
            if a.save:
                np.save(a.save, net.get_params())
        evals = [time_to_tolerance(net, a.eval_size, a.eval_objects, s, log=log)
                 for s in a.eval_seeds]
    print(json.dumps({"dataset": ds, "training": tr, "eval": evals,
                      "macs_per_block": macs_per_block()}))


if __name__ == "__main__":
    main()
