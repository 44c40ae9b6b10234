"""Shared scene builders and pipelines for the parity tests.

`oracle_pipeline` runs the demo's hot path (msbg_demo.cpp:339-727) on the CPU
oracle; `gpu_pipeline` runs the same steps through the C ABI. Both start from
the same host particle array.
"""
import numpy as np

from oracle import liboracle as O

RSCAN = 4.0  # rParticle + nbDist of the demo (msbg_demo.cpp:235-236,317)


def small_scene(S, nObj=6, nPts=3000, seed=7, rmin=None, rmax=None, margin=None):
    """Jittered sphere points inside an S^3 box (numpy, float32)."""
    rng = np.random.default_rng(seed)
    rmin = S / 10 if rmin is None else rmin
    rmax = S / 5 if rmax is None else rmax
    margin = rmax + 2 if margin is None else margin
    pts = []
    for _ in range(nObj):
        c = rng.uniform(margin, S - margin, 3)
        r = rng.uniform(rmin, rmax)
        n = rng.normal(size=(nPts, 3))
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        pts.append(c + r * n + rng.normal(scale=0.5, size=(nPts, 3)))
    return np.concatenate(pts).astype(np.float32)


def edge_points(S, n=500, seed=3):
    """Points near and beyond the domain faces (filter and clamp edge cases)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(-2, S + 2, size=(n, 3)).astype(np.float32)
    k = rng.integers(0, 3, n)
    side = rng.integers(0, 2, n)
    eps = np.float32([0, 1e-4, 2e-4, 5e-5, 0.49, 0.5, 0.51, 1.0])
    vals = np.where(side == 1, S - eps[rng.integers(0, 8, n)], eps[rng.integers(0, 8, n)])
    p[np.arange(n), k] = vals.astype(np.float32)
    return p


def oracle_pipeline(S, bsx, xyz, nLevels=3, rScan=RSCAN, storage=O.U16, numIter=2, dt=0.1,
                    residual=False, do_sweep=True):
    nb = S // bsx
    keep = O.domain_filter(S, S, S, xyz)
    xk = np.ascontiguousarray(xyz[keep == 1])
    cnt, home = O.activate(bsx, nb, nb, nb, xk, rScan)
    lv0 = O.levels_from_activity(cnt, nLevels)
    lv, rc = O.regularize(nb, nb, nb, nLevels, lv0)
    assert rc == 0
    blv, bfl = O.block_flags(nb, nb, nb, nLevels, lv)
    active = np.nonzero(blv == 0)[0].astype(np.int32)
    pagemap = np.full(nb ** 3, O.EMPTY, np.int32)
    pagemap[active] = np.arange(active.size, dtype=np.int32)
    pages = np.full((active.size, bsx, bsx, bsx), 65535, np.uint16)
    L = O.Level(S, S, S, bsx, pagemap, pages, O.U16)
    O.splat_u16(L, xk, rScan)
    O.finalize_u16(L.pages)
    out = dict(keep=keep, xk=xk, cnt=cnt, lv0=lv0, lv=lv, blv=blv, bfl=bfl, active=active,
               pagemap=pagemap, splat=L.pages.copy())
    if storage != O.U16:
        vals = L.pages.astype(np.float32) * np.float32(1.0 / 65535.0)
        L = to_storage_level(S, bsx, pagemap, vals, storage)
    res = None
    if do_sweep:
        res = O.mcf_sweeps(L, active, blv, numIter, dt, 1, residual=residual)
    out.update(level=L, res=res)
    return out


def to_storage_level(S, bsx, pagemap, vals_f32, storage):
    L = O.load()
    v = np.ascontiguousarray(vals_f32, np.float32)
    if storage == O.F32:
        pages = v.copy()
    elif storage == O.F64:
        pages = v.astype(np.float64)
    elif storage == O.F16:
        pages = v.astype(np.float16).view(np.uint16)
    elif storage == O.BF16:
        pages = np.vectorize(L.or_bf16_from_float, otypes=[np.uint16])(v)
    elif storage == O.U8:  # round_to_int(v * 255), half to even (sbg.h:319-323)
        pages = np.rint(v * np.float32(255)).astype(np.uint8)
    else:
        raise ValueError(storage)
    full = 255.0 if storage == O.U8 else 1.0
    return O.Level(S, S, S, bsx, pagemap, pages, storage, 0.0, full)


def gpu_pipeline(S, bsx, xyz, nLevels=3, rScan=RSCAN, storage=1, numIter=2, dt=0.1,
                 residual=False, arith=0, do_sweep=True, chan=30, do_splat=True):
    import msbg_amd as M
    g = M.MultiresSparseGrid.create("test", S, S, S, bsx, options=M.OPT_SINGLE_LEVEL,
                                    nLevels=nLevels)
    g.set_arith(arith)
    nk = g.points_upload(xyz)
    cnt = g.activate(rScan)
    lv0 = g.levels_from_points([rScan])
    lv = g.regularizeRefinementMap(lv0)
    g.setRefinementMap(lv)
    blv, bfl = g.getBlockInfo()
    active = g.alloc_demo(chan, storage)
    splat = None
    if do_splat:
        g.splat(chan, rScan / 2, rScan / 2)
        g.finalize(chan, rScan / 2, rScan / 2)
        splat = g.download(chan)
    res = None
    if do_sweep:
        res = g.applyChannelPdeFast(chan, 0, 0, active, M.MCF_8COLOR, 1, 0, 0, 0, numIter, 1.0,
                                    dt, residual=residual)
    out = dict(grid=g, nk=nk, cnt=cnt, lv0=lv0, lv=lv, blv=blv, bfl=bfl, active=active,
               splat=splat, pages=g.download(chan), res=res)
    return out


def dense_halo(pagemap, pages_f32, S, bsx, bid, empty=0.0, full=1.0):
    """Independent restatement of the Neumann halo of one block through a
    dense embedding: every voxel of the domain is looked up in a dense array,
    NULL blocks and out-of-domain voxels take the block's own clamped voxel."""
    nb = S // bsx
    bz, rem = divmod(bid, nb * nb)
    by, bx = divmod(rem, nb)
    o = np.array([bx, by, bz]) * bsx
    h = np.zeros((bsx + 2,) * 3, np.float32)
    for hz in range(bsx + 2):
        for hy in range(bsx + 2):
            for hx in range(bsx + 2):
                g = o + np.array([hx, hy, hz]) - 1
                if np.all(g >= 0) and np.all(g < S):
                    b = g // bsx
                    code = pagemap[(b[2] * nb + b[1]) * nb + b[0]]
                    if code >= 0:
                        v = g % bsx
                        h[hz, hy, hx] = pages_f32[code, v[2], v[1], v[0]]
                        continue
                    if code == O.EMPTY:
                        h[hz, hy, hx] = empty
                        continue
                    if code == O.FULL:
                        h[hz, hy, hx] = full
                        continue
                c = np.clip(g, o, o + bsx - 1) - o
                h[hz, hy, hx] = pages_f32[pagemap[bid], c[2], c[1], c[0]]
    return h
