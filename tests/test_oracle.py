"""Pins of the CPU oracle against closed forms, brute force and library
routines (no GPU). The reference ships no golden vectors for this path, so
every oracle function is checked here against something that does not
restate its formula.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import liboracle as O
from tests.helpers import dense_halo, small_scene

L = O.load()


# --------------------------------------------------------------------------
# conversions (sbg.h:225-366)
# --------------------------------------------------------------------------
def test_u16_roundings_differ_only_at_even_ties():
    """The splat/finalize path rounds half away from zero (roundf), the sweep
    store rounds half to even (cvtps2dq): find exact .5 products and check
    both modes."""
    seen_even = 0
    for k in range(0, 4000):
        f = np.float32(k + 0.5) / np.float32(65535)
        if np.float32(f) * np.float32(65535) != np.float32(k + 0.5):
            continue
        assert L.or_u16_from_float_roundf(float(f)) == k + 1
        assert L.or_u16_from_float_rne(float(f)) == (k if k % 2 == 0 else k + 1)
        seen_even += k % 2 == 0
    assert seen_even > 10
    assert L.or_u16_to_float(0) == 0.0
    assert abs(L.or_u16_to_float(65535) - 1.0) <= 2 ** -23


def test_f16_matches_numpy_rne():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-2, 2, 20000), rng.uniform(0, 1, 20000),
                        np.float32(2.0) ** rng.uniform(-30, 17, 5000) * rng.choice([-1, 1], 5000),
                        [0.0, -0.0, 65504.0, 65520.0, 65519.99, 6e-8, 3e-8, 2.98e-8]]
                       ).astype(np.float32)
    ref = x.astype(np.float16).view(np.uint16)
    got = np.array([L.or_f16_from_float(float(v)) for v in x], np.uint16)
    assert np.array_equal(got, ref)
    allh = np.arange(65536, dtype=np.uint16)
    finite = np.isfinite(allh.view(np.float16))
    dec = np.array([L.or_f16_to_float(int(h)) for h in allh[finite]], np.float32)
    assert np.array_equal(dec, allh[finite].view(np.float16).astype(np.float32))


def test_bf16_matches_torch_rne():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-2, 2, 20000), rng.uniform(0, 1, 20000)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    got = np.array([L.or_bf16_from_float(float(v)) for v in x], np.uint16)
    assert np.array_equal(got, ref)


# --------------------------------------------------------------------------
# A1 domain bounds (sbg.cpp:262-316)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("S,beps", [(512, 1e-4), (4096, 2e-4), (64, 1e-4), (2048, 1e-4)])
def test_domain_bounds_closed_form(S, beps):
    lo, hi = O.domain_bounds(S, S, S)
    b = np.float32(beps)
    # beps doubles until S - beps rounds below S in float32
    assert np.all(lo == np.float32(0) + b)
    assert np.all(hi == np.float32(S) - b)
    assert np.all(hi < S) and np.all(hi == np.float32(S) - b)


def test_domain_filter_truncation():
    S = 64
    lo, hi = O.domain_bounds(S, S, S)
    p = np.array([[lo[0], 1, 1], [np.nextafter(lo[0], 0), 1, 1], [hi[0], 1, 1],
                  [np.nextafter(hi[0], 100), 1, 1], [np.nan, 1, 1], [63.999, 0.5, 1]], np.float32)
    keep = O.domain_filter(S, S, S, p)
    assert keep.tolist() == [1, 0, 1, 0, 0, 1]


# --------------------------------------------------------------------------
# A2 activation (msbg_demo.cpp:368-389)
# --------------------------------------------------------------------------
def test_activation_hand_kat():
    # bpos = p/16 +- 0.25 truncated toward zero, clamped to [0, nb-1]
    p = np.array([[15.9, 0.2, 31.99]], np.float32)
    cnt, home = O.activate(16, 2, 4, 3, p, 4.0)
    c = cnt.reshape(3, 4, 2)  # z, y, x
    want = np.zeros((3, 4, 2), np.uint32)
    want[1:3, 0, 0:2] = 1  # x: trunc(0.74375)..trunc(1.24375) = 0..1; y: 0..0; z: 1..2
    assert np.array_equal(c, want)
    assert home[0] == (1 * 4 + 0) * 2 + 0  # (15>>4, 0>>4, 31>>4) = (0, 0, 1)


def test_activation_covers_footprint():
    """Invariant: every in-sphere voxel of a point's splat footprint lies in
    a block activated by that point (same rScan)."""
    S, bsx = 64, 16
    nb = S // bsx
    rng = np.random.default_rng(5)
    pts = rng.uniform(0.01, S - 0.01, size=(300, 3)).astype(np.float32)
    for p in pts:
        cnt, _ = O.activate(bsx, nb, nb, nb, p[None], 4.0)
        act = cnt.reshape(nb, nb, nb) > 0
        lo = np.maximum(np.ceil(p - 4.5).astype(int), 0)
        hi = np.minimum(np.floor(p + 3.5).astype(int), S - 1)
        zz, yy, xx = np.meshgrid(*[np.arange(lo[k], hi[k] + 1) for k in (2, 1, 0)], indexing="ij")
        d2 = (xx + 0.5 - p[0]) ** 2 + (yy + 0.5 - p[1]) ** 2 + (zz + 0.5 - p[2]) ** 2
        m = d2 <= 16 - 1e-3
        assert act[zz[m] // bsx, yy[m] // bsx, xx[m] // bsx].all()


# --------------------------------------------------------------------------
# A4 regularizeRefinementMap (msbg.cpp:1202-1423)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("nLevels", [3, 4, 5])
def test_regularize_is_clamped_chebyshev_distance(nLevels):
    """For a {0, nLevels-1} input map, pass ll sets blocks at Chebyshev
    distance ll from level 0 to ll, so the result is
    min(nLevels-1, chessboard distance to the nearest level-0 block)."""
    rng = np.random.default_rng(nLevels)
    nb = 12
    m = np.where(rng.random((nb, nb, nb)) < 0.02, 0, nLevels - 1).astype(np.int32)
    out, rc = O.regularize(nb, nb, nb, nLevels, m.ravel())
    assert rc == 0
    dist = ndi.distance_transform_cdt(m != 0, metric="chessboard")
    want = np.minimum(dist, nLevels - 1)
    assert np.array_equal(out.reshape(nb, nb, nb), want)
    out2, _ = O.regularize(nb, nb, nb, nLevels, out)
    assert np.array_equal(out2, out)  # idempotent


def test_regularize_two_levels_untouched():
    m = np.array([0, 1, 1, 0, 1, 1, 1, 1], np.int32)
    out, rc = O.regularize(2, 2, 2, 2, m)
    assert rc == 0 and np.array_equal(out, m)


# --------------------------------------------------------------------------
# A5 block flags (msbg.cpp:1649-1745)
# --------------------------------------------------------------------------
def test_block_flags_vs_filters():
    rng = np.random.default_rng(2)
    nb = 10
    m = np.where(rng.random((nb, nb, nb)) < 0.05, 0, 2).astype(np.int32)
    lv, _ = O.regularize(nb, nb, nb, 3, m.ravel())
    lv = lv.reshape(nb, nb, nb)
    inflags = (rng.random(nb ** 3) < 0.3).astype(np.uint32) << 12
    blv, bfl = O.block_flags(nb, nb, nb, 3, lv.ravel(), inflags)
    fp = np.ones((3, 3, 3), bool)
    fp[1, 1, 1] = False
    nmax = ndi.maximum_filter(lv, footprint=fp, mode="constant", cval=-1)
    nmin = ndi.minimum_filter(lv, footprint=fp, mode="constant", cval=99)
    border = np.zeros((nb, nb, nb), bool)
    border[[0, -1], :, :] = border[:, [0, -1], :] = border[:, :, [0, -1]] = True
    fc = nmax > lv
    cf = nmin < lv
    ex = (inflags.reshape(nb, nb, nb) & (1 << 12)) != 0
    want = (ex << 12) | (border << 1) | (cf << 2) | (fc << 3) | ((~ex & ~fc & ~cf) << 13)
    assert np.array_equal(bfl.reshape(nb, nb, nb), want.astype(np.uint16))
    assert np.array_equal(blv.reshape(nb, nb, nb), lv)


# --------------------------------------------------------------------------
# A7 colours (msbg.h:642-646)
# --------------------------------------------------------------------------
def test_colours_disjoint_halos():
    nb = 6
    b = np.arange(nb ** 3, dtype=np.int32)
    out, off = O.color_lists(nb, nb, b)
    assert off[-1] == nb ** 3
    for c in range(8):
        blk = out[off[c]:off[c + 1]]
        assert np.all(np.diff(blk) > 0)  # list order kept (ascending)
        z, rem = np.divmod(blk, nb * nb)
        y, x = np.divmod(rem, nb)
        assert np.all((x & 1) + 2 * (y & 1) + 4 * (z & 1) == c)
        # same-colour blocks never share a face, edge or corner
        pts = np.stack([x, y, z], 1)
        d = np.abs(pts[:, None, :] - pts[None, :, :]).max(-1)
        np.fill_diagonal(d, 99)
        assert d.min() >= 2


# --------------------------------------------------------------------------
# A8 splat (msbg_demo.cpp:541-636)
# --------------------------------------------------------------------------
def _single_block_level(S, bsx, code=0):
    nb = S // bsx
    pm = np.full(nb ** 3, code, np.int32)
    npg = 0
    if code == 0:
        pm = np.arange(nb ** 3, dtype=np.int32)
        npg = nb ** 3
    pages = np.full((max(npg, 1), bsx, bsx, bsx), 65535, np.uint16)
    return O.Level(S, S, S, bsx, pm, pages, O.U16)


def _dense(level, S):
    bsx = level.bsx
    nb = S // bsx
    d = np.zeros((S, S, S), np.uint16)
    for bid in range(nb ** 3):
        bz, rem = divmod(bid, nb * nb)
        by, bx = divmod(rem, nb)
        c = level.page[bid]
        d[bz * bsx:(bz + 1) * bsx, by * bsx:(by + 1) * bsx, bx * bsx:(bx + 1) * bsx] = (
            level.pages[c] if c >= 0 else 0)
    return d


@pytest.mark.parametrize("p", [(10.25, 11.5, 9.75), (16.0, 16.0, 16.0), (3.125, 28.5, 17.375)])
def test_splat_single_point_closed_form(p):
    """Dyadic positions make every FP32 step exact, so the code equals the
    float64 closed form roundf(d^2/16*65535) voxel by voxel."""
    S, bsx = 32, 16
    lvl = _single_block_level(S, bsx)
    pt = np.array([p], np.float32)
    O.splat_u16(lvl, pt, 4.0)
    d = _dense(lvl, S)
    z, y, x = np.meshgrid(np.arange(S), np.arange(S), np.arange(S), indexing="ij")
    d2 = (x + 0.5 - p[0]) ** 2 + (y + 0.5 - p[1]) ** 2 + (z + 0.5 - p[2]) ** 2
    want = np.where(d2 <= 16, np.floor(d2 / 16 * 65535 + 0.5), 65535).astype(np.uint16)
    assert np.array_equal(d, want)


def test_splat_abandons_at_first_non_value_block():
    """goto loop_exit (msbg_demo.cpp:615-632): the first in-sphere voxel (z,y,x
    order) in a non-value block ends the particle; earlier voxels keep
    their update, later ones in value blocks are not touched."""
    S, bsx = 32, 16
    nb = 2
    pm = np.full(nb ** 3, O.EMPTY, np.int32)
    pm[0] = 0  # only block (0,0,0) holds values
    lvl = O.Level(S, S, S, bsx, pm, np.full((1, bsx, bsx, bsx), 65535, np.uint16), O.U16)
    p = np.array([[14.5, 10.5, 10.5]], np.float32)  # footprint x in 10..17 crosses into bx=1
    O.splat_u16(lvl, p, 4.0)
    got = lvl.pages[0]
    # walk the footprint in reference order and stop at the first voxel in bx=1
    want = np.full((bsx, bsx, bsx), 65535, np.uint16)
    lo = np.maximum(np.ceil(p[0] - 4.5).astype(int), 0)
    hi = np.minimum(np.floor(p[0] + 3.5).astype(int), S - 1)
    done = False
    for iz in range(lo[2], hi[2] + 1):
        for iy in range(lo[1], hi[1] + 1):
            for ix in range(lo[0], hi[0] + 1):
                d2 = (ix + 0.5 - p[0, 0]) ** 2 + (iy + 0.5 - p[0, 1]) ** 2 + (iz + 0.5 - p[0, 2]) ** 2
                if d2 > 16:
                    continue
                if ix >= 16:
                    done = True
                    break
                want[iz, iy, ix] = min(want[iz, iy, ix], int(np.floor(d2 / 16 * 65535 + 0.5)))
            if done:
                break
        if done:
            break
    assert np.array_equal(got, want)
    assert (got < 65535).sum() > 0 and done


def test_splat_order_independent():
    S, bsx = 64, 16
    pts = small_scene(S, nObj=3, nPts=2000, seed=11)
    keep = O.domain_filter(S, S, S, pts)
    pts = pts[keep == 1]
    a = _single_block_level(S, bsx)
    b = _single_block_level(S, bsx)
    O.splat_u16(a, pts, 4.0)
    O.splat_u16(b, pts[np.random.default_rng(0).permutation(len(pts))], 4.0)
    assert np.array_equal(a.pages, b.pages)


# --------------------------------------------------------------------------
# A9 finalize (msbg_demo.cpp:655-680), exhaustive
# --------------------------------------------------------------------------
def test_finalize_table_closed_form():
    t = O.finalize_table(2.0, 2.0).astype(np.int64)
    u = np.arange(65536)
    q = u / 65535.0
    d = np.sqrt(q * 16) - 2
    f = 1 - np.clip((d + 2) / 6, 0, 1)
    want = np.where(q < 0.999, np.floor(f * 65535 + 0.5), 0)
    assert np.abs(t - want).max() <= 1
    assert t[0] == 65535 and t[65535] == 0
    # q >= 0.999 gives 0; the cut sits where u/65535 crosses 0.999 in FP32
    cut = np.nonzero((u.astype(np.float32) * np.float32(1 / 65535)) >= np.float32(0.999))[0][0]
    assert np.all(t[cut:] == 0) and t[cut - 1] > 0


def test_finalize_density_float_closed_form():
    # the float variant behind the FP32/FP16/BF16 finalize: with rParticle =
    # nbDist = 2 the ramp never saturates, so f = 1 - sqrt(q) * 4 / 6 below
    # the 0.999 cut and 0 above it, for every code
    f = O.finalize_density_table(2.0, 2.0).astype(np.float64)
    u = np.arange(65536)
    q32 = u.astype(np.float32) * np.float32(1 / 65535)
    below = q32 < np.float32(0.999)
    want = 1 - np.sqrt(q32.astype(np.float64)) * 4 / 6
    assert np.abs(f[below] - want[below]).max() < 3e-7
    assert np.all(f[~below] == 0.0) and f[0] == 1.0
    assert np.all(np.diff(f[below]) <= 0)  # farther from a particle, less density
    # the UNORM16 finalize is the same value rounded half away from zero
    f32 = O.finalize_density_table(2.0, 2.0)
    prod = (f32 * np.float32(65535)).astype(np.float64)  # the FP32 product, then roundf
    assert np.array_equal(O.finalize_table(2.0, 2.0), np.floor(prod + 0.5).astype(np.uint16))
    # other radii move the ramp: rParticle 1, nbDist 3 (rScan 4, ramp [-1, 4])
    g = O.finalize_density_table(1.0, 3.0).astype(np.float64)
    d = np.sqrt(q32.astype(np.float64) * 16) - 1
    want = 1 - np.clip((d + 1) / 5, 0, 1)
    assert np.abs(g[below] - want[below]).max() < 3e-7


# --------------------------------------------------------------------------
# A11 halo (halo.cpp:267-479): tag test against a dense embedding
# --------------------------------------------------------------------------
def _random_grid(S, bsx, seed, storage=O.F32):
    rng = np.random.default_rng(seed)
    nb = S // bsx
    kinds = rng.choice([O.NULL, O.EMPTY, O.FULL, 0], size=nb ** 3, p=[0.25, 0.15, 0.1, 0.5])
    pm = np.full(nb ** 3, O.NULL, np.int32)
    pm[kinds == O.EMPTY] = O.EMPTY
    pm[kinds == O.FULL] = O.FULL
    vb = np.nonzero(kinds == 0)[0]
    pm[vb] = np.arange(vb.size)
    # tag each voxel with its global linear index (exact in float32 < 2^24)
    pages = np.zeros((vb.size, bsx, bsx, bsx), np.float32)
    for i, bid in enumerate(vb):
        bz, rem = divmod(bid, nb * nb)
        by, bx = divmod(rem, nb)
        z, y, x = np.meshgrid(*[np.arange(bsx)] * 3, indexing="ij")
        pages[i] = ((bz * bsx + z) * S + by * bsx + y) * S + bx * bsx + x + 7
    return O.Level(S, S, S, bsx, pm, pages, O.F32, 0.5, 0.25), vb


def test_halo_neumann_tag_dense_embedding():
    S, bsx = 32, 8
    lvl, vb = _random_grid(S, bsx, 4)
    for bid in vb:
        h = O.halo_fill(lvl, int(bid), 1, O.OPT_BC_NEUMANN)
        want = dense_halo(lvl.page, lvl.pages, S, bsx, int(bid), empty=0.5, full=0.25)
        assert np.array_equal(h, want), bid


# --------------------------------------------------------------------------
# A12 MCF voxel closed forms (msbg4.cpp:559-637)
# --------------------------------------------------------------------------
def _halo_of(fn, n=5):
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64)] * 3, indexing="ij")
    return fn(x, y, z).astype(np.float32)


def _centre(n=5):
    c = n // 2
    return (c * n + c) * n + c


def test_mcf_constant_and_linear_fields_unchanged():
    for fn in [lambda x, y, z: 0 * x + 0.375, lambda x, y, z: 0.5 + 0.01 * x - 0.02 * y + 0.005 * z]:
        h = _halo_of(fn)
        assert O.mcf_voxel(h, _centre(), 0.1) == h.ravel()[_centre()]


def test_mcf_quadratic_radial_exact():
    """phi = s*r^2 + c: central differences are exact for quadratics and
    H/|grad|^2 = |grad phi| * (2/r) = 4s, so the update is dt*4s."""
    s, c = 0.001, 0.25
    cx, cy, cz = 0.5, -1.25, 0.75
    h = _halo_of(lambda x, y, z: s * ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) + c)
    f0 = h.ravel()[_centre()]
    got = O.mcf_voxel(h, _centre(), 0.1)
    assert abs((got - f0) - 0.1 * 4 * s) < 2e-6


@pytest.mark.parametrize("axes", [(0, 1), (0, 2), (1, 2)])
def test_mcf_cross_terms(axes):
    """phi = a + s*u*v (u, v two axes): H/|grad|^2 = -2 s^2 u v / (s^2 (u^2+v^2))
    * s = -2 s u v/(u^2+v^2), exact for the quadratic."""
    s, a = 0.01, 0.5
    off = np.array([3.0, -2.0, 5.0])

    def fn(x, y, z):
        g = [x + off[0], y + off[1], z + off[2]]
        return a + s * g[axes[0]] * g[axes[1]]

    h = _halo_of(fn)
    u, v = off[axes[0]] + 2, off[axes[1]] + 2
    want = 0.1 * (-2 * s * u * v / (u * u + v * v))
    got = O.mcf_voxel(h, _centre(), 0.1) - h.ravel()[_centre()]
    assert abs(got - want) < 2e-6


def test_mcf_threshold_limiter_clamp():
    # |grad|^2 <= 1e-7 -> no update
    h = _halo_of(lambda x, y, z: 0.5 + 1e-4 * x * x)
    assert O.mcf_voxel(h, _centre(), 0.1) == h.ravel()[_centre()]
    # limiter: |D| <= 0.1 and the clamp keeps [0,1]
    h = _halo_of(lambda x, y, z: np.where((x == 2) & (y == 2) & (z == 2), 0.95, 0.0) + 0.001 * x)
    got = O.mcf_voxel(h, _centre(), 10.0)
    assert abs(got - 0.95) <= 0.1 + 1e-7
    h2 = _halo_of(lambda x, y, z: 0.02 + 0.001 * x * x + 0.0 * y)
    assert O.mcf_voxel(h2, _centre(), -100.0, 1) >= 0.0


def test_mcf_axis_permutation_symmetry():
    rng = np.random.default_rng(9)
    h = rng.random((5, 5, 5)).astype(np.float32)
    a = O.mcf_voxel(h, _centre(), 0.1)
    for perm in [(0, 2, 1), (1, 0, 2), (2, 1, 0)]:
        b = O.mcf_voxel(np.ascontiguousarray(h.transpose(perm)), _centre(), 0.1)
        assert abs(a - b) <= 4e-7


# --------------------------------------------------------------------------
# A10 sweep driver vs a dense numpy Gauss-Seidel-across-colours restatement
# --------------------------------------------------------------------------
def _np_mcf(h, dt):
    """Vectorised MCF of all interior voxels of a halo block, float32,
    reference operation order."""
    f = lambda dx, dy, dz: h[1 + dz:h.shape[0] - 1 + dz, 1 + dy:h.shape[1] - 1 + dy,
                             1 + dx:h.shape[2] - 1 + dx]
    hf = np.float32(0.5)
    f0, f1, f2, f3, f4, f5, f6 = (f(0, 0, 0), f(-1, 0, 0), f(1, 0, 0), f(0, -1, 0), f(0, 1, 0),
                                  f(0, 0, -1), f(0, 0, 1))
    fx, fy, fz = hf * (f2 - f1), hf * (f4 - f3), hf * (f6 - f5)
    two = np.float32(2)
    fxx, fyy, fzz = f1 + f2 - two * f0, f3 + f4 - two * f0, f5 + f6 - two * f0
    fyx = ((f(1, 1, 0) - f(1, -1, 0)) * hf - (f(-1, 1, 0) - f(-1, -1, 0)) * hf) * hf
    fzx = ((f(1, 0, 1) - f(1, 0, -1)) * hf - (f(-1, 0, 1) - f(-1, 0, -1)) * hf) * hf
    fzy = ((f(0, 1, 1) - f(0, 1, -1)) * hf - (f(0, -1, 1) - f(0, -1, -1)) * hf) * hf
    H = ((fy * fy + fz * fz) * fxx + (fx * fx + fz * fz) * fyy + (fx * fx + fy * fy) * fzz
         - two * (fx * fy * fyx + fx * fz * fzx + fy * fz * fzy))
    g = fx * fx + fy * fy + fz * fz
    with np.errstate(divide="ignore", invalid="ignore"):
        H = np.where(g > np.float32(1e-7), H / g, np.float32(0))
    D = np.float32(dt) * H
    D = np.where(D > np.float32(-0.1), D, np.float32(-0.1))
    D = np.where(D < np.float32(0.1), D, np.float32(0.1))
    p = f0 + D
    p = np.where(p > 0, p, np.float32(0))
    return np.where(p < 1, p, np.float32(1)).astype(np.float32)


def test_sweep_driver_matches_dense_colour_restatement():
    S, bsx = 32, 8
    nb = S // bsx
    rng = np.random.default_rng(3)
    kinds = rng.random(nb ** 3)
    pm = np.full(nb ** 3, O.EMPTY, np.int32)
    pm[kinds < 0.15] = O.NULL
    vb = np.nonzero(kinds >= 0.4)[0].astype(np.int32)
    pm[vb] = np.arange(vb.size)
    z, y, x = np.meshgrid(*[np.arange(S)] * 3, indexing="ij")
    field = (0.5 + 0.45 * np.sin(x / 5.0) * np.cos(y / 7.0 + z / 9.0)).astype(np.float32)
    pages = np.zeros((vb.size, bsx, bsx, bsx), np.float32)
    for i, bid in enumerate(vb):
        bz, rem = divmod(bid, nb * nb)
        by, bx = divmod(rem, nb)
        pages[i] = field[bz * bsx:(bz + 1) * bsx, by * bsx:(by + 1) * bsx, bx * bsx:(bx + 1) * bsx]
    levels = np.zeros(nb ** 3, np.uint16)
    lvl = O.Level(S, S, S, bsx, pm, pages.copy(), O.F32)
    n = O.mcf_sweeps(lvl, vb, levels, 3, 0.1, 1, residual=True)
    assert n.size == 4  # numIter 3 rounds up to 4 sweeps
    ref = pages.copy()
    rms = []
    for _ in range(4):
        before = ref.astype(np.float64)
        for c in range(8):
            for bid in vb:
                bz, rem = divmod(int(bid), nb * nb)
                by, bx = divmod(rem, nb)
                if (bx & 1) + 2 * (by & 1) + 4 * (bz & 1) != c:
                    continue
                h = dense_halo(pm, ref, S, bsx, int(bid))
                ref[pm[bid]] = _np_mcf(h, 0.1)
        # the residual: RMS over the swept voxels of this sweep's change
        rms.append(np.sqrt(np.mean((ref.astype(np.float64) - before) ** 2)))
    assert np.array_equal(lvl.pages, ref)
    np.testing.assert_allclose(n, rms, rtol=1e-12, atol=0)
    assert np.all(np.diff(n) < 0)  # the flow settles


def test_residual_zero_for_fields_the_flow_keeps():
    # constant and linear fields are fixed points of the update (the linear
    # one away from the domain faces, where the Neumann halo bends it): every
    # sweep's residual over the interior blocks is exactly 0
    S, bsx = 32, 8
    z, y, x = np.meshgrid(*[np.arange(S, dtype=np.float64)] * 3, indexing="ij")
    nb = S // bsx
    for f in (np.full((S,) * 3, 0.375), 0.25 + x / 128 + y / 256):
        for st in (O.F32, O.F64):
            dt_ = np.float64 if st == O.F64 else np.float32
            pages = np.stack([f[(b // (nb * nb)) * bsx:(b // (nb * nb) + 1) * bsx,
                                ((b // nb) % nb) * bsx:((b // nb) % nb + 1) * bsx,
                                (b % nb) * bsx:(b % nb + 1) * bsx] for b in range(nb ** 3)])
            lvl = O.Level(S, S, S, bsx, np.arange(nb ** 3, dtype=np.int32), pages.astype(dt_), st)
            inner = np.array([b for b in range(nb ** 3)
                              if all(0 < c < nb - 1 for c in (b % nb, (b // nb) % nb, b // nb ** 2))],
                             np.int32)
            r = O.mcf_sweeps(lvl, inner, np.zeros(nb ** 3, np.uint16), 2, 0.1, 1, residual=True)
            assert r.size == 2 and np.all(r == 0.0), (st, r)
            # all blocks: the faces move, and the residual sees it
            r = O.mcf_sweeps(lvl, np.arange(nb ** 3, dtype=np.int32), np.zeros(nb ** 3, np.uint16),
                             2, 0.1, 1, residual=True)
            assert np.all(r == 0.0) == (f.std() == 0), (st, r)


def test_ghost_downsample_closed_forms():
    bsx = 8
    z, y, x = np.meshgrid(*[np.arange(bsx + 2, dtype=np.float64)] * 3, indexing="ij")
    const = np.full((bsx + 2,) * 3, 0.625, np.float32)
    assert np.all(O.ghost_downsample(const, bsx) == np.float32(0.625))
    a, b, c, d = 1 / 64, -1 / 128, 1 / 32, 0.25
    lin = (a * x + b * y + c * z + d).astype(np.float32)
    lo = O.ghost_downsample(lin, bsx)
    v = np.arange(bsx // 2)
    cz, cy, cx = np.meshgrid(v, v, v, indexing="ij")
    # window of halo indices 2v..2v+3 is centred at halo coordinate 2v+1.5
    want = a * (2 * cx + 1.5) + b * (2 * cy + 1.5) + c * (2 * cz + 1.5) + d
    assert np.abs(lo - want).max() < 1e-6


# --------------------------------------------------------------------------
# coarse-level sweeps (an extension: the reference sweeps level 0 only)
# vs the dense restatement with the COARSE_LEVEL halo rule
# --------------------------------------------------------------------------
def _coarse_halo(pm, pages, S, bsx, bid, lo_dense, empty=0.0, full=1.0):
    """getOutOfBlockValue with OPT_BC_COARSE_LEVEL (halo.h:288-352), no
    interpolation: in-domain non-NULL blocks give their value; NULL blocks
    read the coarser level at (x/2, y/2, z/2); out-of-domain voxels are
    clipped to the domain first and read like in-domain ones."""
    nb = S // bsx
    bz, rem = divmod(bid, nb * nb)
    by, bx = divmod(rem, nb)
    o = np.array([bx, by, bz]) * bsx
    h = np.zeros((bsx + 2,) * 3, np.float32)
    for hz in range(bsx + 2):
        for hy in range(bsx + 2):
            for hx in range(bsx + 2):
                g = np.clip(o + np.array([hx, hy, hz]) - 1, 0, S - 1)
                b = g // bsx
                code = pm[(b[2] * nb + b[1]) * nb + b[0]]
                if code >= 0:
                    v = g % bsx
                    h[hz, hy, hx] = pages[code, v[2], v[1], v[0]]
                elif code == O.EMPTY:
                    h[hz, hy, hx] = empty
                elif code == O.FULL:
                    h[hz, hy, hx] = full
                else:
                    gl = g // 2
                    h[hz, hy, hx] = lo_dense[gl[2], gl[1], gl[0]]
    return h


def test_coarse_level_sweep_matches_dense_restatement():
    S0, bsx0 = 64, 16
    S1, b1 = S0 // 2, bsx0 // 2  # level 1: 32^3 voxels in 8^3 blocks
    nb = S0 // bsx0
    rng = np.random.default_rng(21)
    levels = rng.choice([0, 1, 2], nb ** 3, p=[0.3, 0.45, 0.25]).astype(np.uint16)
    # level 1 holds its own blocks and (as prepareGhostBlocks leaves it)
    # ghost copies at some finer blocks; the rest is NULL
    pm1 = np.full(nb ** 3, O.NULL, np.int32)
    own = np.nonzero(levels == 1)[0]
    ghost = np.nonzero((levels == 0) & (rng.random(nb ** 3) < 0.6))[0]
    vals = np.sort(np.concatenate([own, ghost])).astype(np.int32)
    pm1[vals] = np.arange(vals.size)
    pm1[np.nonzero((levels == 0) & (pm1 == O.NULL))[0][:1]] = O.FULL
    z, y, x = np.meshgrid(*[np.arange(S1)] * 3, indexing="ij")
    f1 = (0.5 + 0.4 * np.sin(x / 4.0) * np.cos(y / 5.0 + z / 6.0)).astype(np.float32)
    pages1 = np.stack([f1[(b // (nb * nb)) * b1:(b // (nb * nb) + 1) * b1,
                          ((b // nb) % nb) * b1:((b // nb) % nb + 1) * b1,
                          (b % nb) * b1:(b % nb + 1) * b1] for b in vals])
    # level 2: 16^3 voxels in 4^3 blocks; its value blocks everywhere but one
    S2, b2 = S0 // 4, bsx0 // 4
    pm2 = np.arange(nb ** 3, dtype=np.int32)
    pm2[5] = O.NULL
    f2 = rng.random((S2, S2, S2)).astype(np.float32)
    pages2 = np.stack([f2[(b // (nb * nb)) * b2:(b // (nb * nb) + 1) * b2,
                          ((b // nb) % nb) * b2:((b // nb) % nb + 1) * b2,
                          (b % nb) * b2:(b % nb + 1) * b2] for b in range(nb ** 3)])
    lo_dense = f2.copy()
    bz, rem = divmod(5, nb * nb)
    by, bx = divmod(rem, nb)
    lo_dense[bz * b2:(bz + 1) * b2, by * b2:(by + 1) * b2, bx * b2:(bx + 1) * b2] = 0.0
    L1 = O.Level(S1, S1, S1, b1, pm1, pages1.copy(), O.F32)
    L2 = O.Level(S2, S2, S2, b2, pm2, pages2, O.F32)
    res = O.mcf_sweeps_level(L1, L2, vals, levels, 1, O.OPT_BC_COARSE_LEVEL, 2, 0.1, 1,
                             residual=True)
    assert res.size == 2 and np.all(res > 0)
    ref = pages1.copy()
    rms = []
    for _ in range(2):
        before = ref[pm1[own]].astype(np.float64)
        for c in range(8):
            for bid in own:
                bz, rem = divmod(int(bid), nb * nb)
                by, bx = divmod(rem, nb)
                if (bx & 1) + 2 * (by & 1) + 4 * (bz & 1) != c:
                    continue
                h = _coarse_halo(pm1, ref, S1, b1, int(bid), lo_dense)
                ref[pm1[bid]] = _np_mcf(h, 0.1)
        # RMS over the level's own (swept) blocks only: ghost pages do not count
        rms.append(np.sqrt(np.mean((ref[pm1[own]].astype(np.float64) - before) ** 2)))
    assert np.array_equal(L1.pages, ref)
    np.testing.assert_allclose(res, rms, rtol=1e-12, atol=0)
    # ghost pages of finer blocks are read, never written
    gi = pm1[ghost]
    assert np.array_equal(L1.pages[gi], pages1[gi])
    # the level-0 rule at level 1 is a different result
    L1n = O.Level(S1, S1, S1, b1, pm1, pages1.copy(), O.F32)
    O.mcf_sweeps_level(L1n, None, vals, levels, 1, O.OPT_BC_NEUMANN, 2, 0.1, 1)
    assert not np.array_equal(L1n.pages, L1.pages)


# --------------------------------------------------------------------------
# A13 interpolateGhost (sbg.h:3002-3095): closed forms and hand KATs
# --------------------------------------------------------------------------
def _level_from_dense(f, bsx, storage=O.F32, null=(), empty_raw=0.0):
    """A level whose value blocks hold the dense array f[z, y, x], with the
    blocks in `null` left NULL."""
    S = f.shape[0]
    nb = S // bsx
    pm = np.full(nb ** 3, O.NULL, np.int32)
    vb = [b for b in range(nb ** 3) if b not in set(null)]
    pm[vb] = np.arange(len(vb))
    pages = np.stack([f[(b // (nb * nb)) * bsx:(b // (nb * nb) + 1) * bsx,
                        ((b // nb) % nb) * bsx:((b // nb) % nb + 1) * bsx,
                        (b % nb) * bsx:(b % nb + 1) * bsx] for b in vb])
    return O.Level(S, S, S, bsx, pm, pages, storage, empty_raw)


def _grid(S):
    return np.meshgrid(*[np.arange(S, dtype=np.float64)] * 3, indexing="ij")  # z, y, x


def test_interpolate_ghost_linear_field_is_exact():
    # a linear field sampled at the cell centres is reproduced exactly at any
    # position whose eight corners lie in the domain (dyadic values, so FP32
    # is exact for positions on the half-voxel lattice)
    S, bsx = 16, 4
    a, bx, by, bz = 0.25, 1 / 32, -1 / 64, 1 / 16
    z, y, x = _grid(S)
    f = (a + bx * (x + 0.5) + by * (y + 0.5) + bz * (z + 0.5)).astype(np.float32)
    lvl = _level_from_dense(f, bsx)
    for g in [(1, 1, 1), (5, 20, 9), (2 * S - 2, 3, 17), (14, 15, 2 * S - 2)]:
        p = [(c + 0.5) / 2 for c in g]  # the fine-voxel centre in coarse coordinates
        want = a + bx * p[0] + by * p[1] + bz * p[2]
        assert O.interpolate_ghost(lvl, p) == want, g
    rng = np.random.default_rng(2)
    for p in rng.uniform(0.5, S - 0.5, (200, 3)):
        p32 = p.astype(np.float32).astype(np.float64)
        want = a + bx * p32[0] + by * p32[1] + bz * p32[2]
        assert abs(O.interpolate_ghost(lvl, p32) - want) < 2e-6


def test_interpolate_ghost_constant_and_clipping():
    S, bsx = 8, 4
    z, y, x = _grid(S)
    const = _level_from_dense(np.full((S,) * 3, 0.625, np.float32), bsx)
    for p in [(0.1, 0.2, 0.3), (7.9, 7.75, 0.0), (4.0, 4.0, 4.0), (-0.5, 9.0, 3.3)]:
        assert O.interpolate_ghost(const, p) == 0.625
    # clamp to edge: outside the cell centres of the first and last voxel the
    # value is that voxel's, whatever the fraction
    ramp = _level_from_dense((0.25 + x / 8).astype(np.float32), bsx)
    assert O.interpolate_ghost(ramp, (0.25, 3.5, 3.5)) == 0.25
    assert O.interpolate_ghost(ramp, (0.0, 0.3, 7.9)) == 0.25
    assert O.interpolate_ghost(ramp, (S - 0.25, 3.5, 3.5)) == 0.25 + (S - 1) / 8
    # between the centres it is the linear ramp
    assert O.interpolate_ghost(ramp, (2.25, 3.5, 3.5)) == 0.25 + 1.75 / 8


def test_interpolate_ghost_null_corners_take_the_value_at_p():
    # Corners in NULL blocks take getValue(p) (floor(p), clamped) once; if
    # that voxel's block is NULL too, the empty value. Hand-evaluated:
    # block 0 holds 1 except voxel (3,1,1) = 5; block 1 (x = 4..7) is NULL.
    S, bsx = 8, 4
    f = np.ones((S,) * 3, np.float32)
    f[1, 1, 3] = 5.0
    lvl = _level_from_dense(f, bsx, null=(1,), empty_raw=0.5)
    # p = (3.75, 1.75, 1.75): u = v = w = 1/4, getValue(p) = f(3,1,1) = 5
    #   x: (5,5)->5, (1,5)->2, (1,5)->2, (1,5)->2;  y: (5,2)->4.25, (2,2)->2;
    #   z: (4.25, 2) -> 3.6875
    assert O.interpolate_ghost(lvl, (3.75, 1.75, 1.75)) == 3.6875
    # p = (4.25, 1.75, 1.75): u = 3/4, floor(p) = (4,1,1) lies in the NULL
    # block, so the NULL corners read the empty value 0.5
    #   x: (5,.5)->1.625, (1,.5)->0.625 (x3);  y: (1.625,.625)->1.375, ->.625;
    #   z: (1.375, .625) -> 1.1875
    assert O.interpolate_ghost(lvl, (4.25, 1.75, 1.75)) == 1.1875


def test_interpolate_ghost_unorm16_truncates_each_stage():
    # UNORM16 intermediates are uint16: each stage truncates toward zero
    S, bsx = 8, 4
    z, y, x = _grid(S)
    lvl = _level_from_dense((3 * (x.astype(np.int64) % 2)).astype(np.uint16), bsx, O.U16)
    assert O.interpolate_ghost(lvl, (0.75, 2.5, 2.5)) == 0.0  # 3 * 0.25 = 0.75 -> 0
    assert O.interpolate_ghost(lvl, (1.25, 2.5, 2.5)) == 2.0  # 3 * 0.75 = 2.25 -> 2
    flt = _level_from_dense((3 * (x.astype(np.int64) % 2)).astype(np.float32), bsx)
    assert O.interpolate_ghost(flt, (0.75, 2.5, 2.5)) == 0.75


def test_coarse_interpolate_halo_is_the_linear_field():
    # OPT_BC_COARSE_LEVEL|OPT_INTERPOLATE (halo.h:288-310): a NULL neighbour's
    # halo voxel g reads the coarser level at (g + .5)/2, after clipping g
    # into the domain (a clipped voxel in a non-NULL block reads that block).
    # With a linear field on both levels (fine voxel g holds
    # F(g) = f((g + .5)/2), coarse voxel X holds f(X + .5)) the halo is F,
    # except that the coarse interpolation clamps to its edge cell centres
    # (0.5 and S/2 - 0.5 per axis).
    S, bsx = 32, 8
    nb = S // bsx
    a, bx, by, bz = 0.25, 1 / 16, 1 / 32, -1 / 64
    fz = lambda z_, y_, x_: a + bx * x_ + by * y_ + bz * z_  # noqa: E731
    z, y, x = _grid(S // 2)
    lo = _level_from_dense(fz(z + 0.5, y + 0.5, x + 0.5).astype(np.float32), bsx // 2)
    z, y, x = _grid(S)
    Fd = fz((z + 0.5) / 2, (y + 0.5) / 2, (x + 0.5) / 2).astype(np.float32)
    for b in (((1 * nb) + 1) * nb + 2, 0, nb ** 3 - 1):
        fine = _level_from_dense(Fd, bsx, null=[i for i in range(nb ** 3) if i != b])
        h = O.halo_fill(fine, b, 1, O.OPT_BC_COARSE_LEVEL | O.OPT_INTERPOLATE, lo)
        h = h.reshape((bsx + 2,) * 3)
        bz_, rem = divmod(b, nb * nb)
        by_, bx_ = divmod(rem, nb)
        o = np.array([bz_, by_, bx_]) * bsx
        for k in np.ndindex(*h.shape):
            gc = np.clip(o + np.array(k) - 1, 0, S - 1)
            if (gc // bsx == o // bsx).all():
                want = Fd[tuple(gc)]  # the block itself (in the domain or clipped into it)
            else:
                pc = np.clip((gc + 0.5) / 2, 0.5, S / 2 - 0.5)
                want = np.float32(fz(*pc))
            assert h[k] == want, (b, k)


# --------------------------------------------------------------------------
# prepareGhostBlocks' block selection (msbg.cpp:3100-3149)
# --------------------------------------------------------------------------
def test_prepare_ghost_blocks_selects_fine_coarse_existing_blocks():
    # Only blocks of the fine-coarse lists of levels 0..nLevels-2 that carry
    # BLK_EXISTS are downsampled into their level+1 slot; the others (no
    # BLK_EXISTS: the reference sets the EMPTY block there) and every block
    # outside the lists keep what the slot held, and the list of the last
    # level is not used.
    S0, b0 = 32, 8
    nb = S0 // b0
    pages0 = np.full((nb ** 3, b0, b0, b0), 0.5, np.float32)
    L0 = O.Level(S0, S0, S0, b0, np.arange(nb ** 3, dtype=np.int32), pages0, O.F32)
    pm1 = np.full(nb ** 3, O.NULL, np.int32)
    slots = [3, 17, 40, 41]
    pm1[slots] = np.arange(len(slots))
    L1 = O.Level(S0 // 2, S0 // 2, S0 // 2, b0 // 2, pm1,
                 np.full((len(slots), b0 // 2, b0 // 2, b0 // 2), 0.125, np.float32), O.F32)
    flags = np.zeros(nb ** 3, np.uint16)
    EXISTS, FC = 1 << 12, 1 << 3
    flags[[3, 40, 41]] |= EXISTS
    flags[[3, 17, 40]] |= FC
    O.prepare_ghost_blocks([L0, L1], [[3, 17, 40], [41]], flags)
    got = {b: L1.pages[pm1[b]] for b in slots}
    assert np.all(got[3] == 0.5) and np.all(got[40] == 0.5)  # downsampled (constant field)
    assert np.all(got[17] == 0.125)  # in the list but without BLK_EXISTS
    assert np.all(got[41] == 0.125)  # only in the last level's list
    assert np.all(L0.pages == 0.5)  # the fine level is only read


# --------------------------------------------------------------------------
# UNORM8 (applyChannelPdeFast<uint8_t>): decode and the stochastic store
# (sbg.h:228-231, 324-340; msbg4.cpp:829-837)
# --------------------------------------------------------------------------
def test_u8_decode_and_stochastic_rounding():
    u = np.arange(256)
    d = np.array([L.or_u8_to_float(int(x)) for x in u], np.float32)
    assert np.array_equal(d, u.astype(np.float32) * np.float32(1 / 255))
    # the uniform stream: [0, 1) on the 2^-23 lattice, mean 1/2, variance
    # 1/12, and no correlation between neighbouring voxels or steps
    r = np.array([[[L.or_sr_uniform(s, p, v) for v in range(256)] for p in range(8)]
                  for s in range(8)], np.float64)
    assert r.min() >= 0 and r.max() < 1
    assert np.all(r * 2 ** 23 == np.floor(r * 2 ** 23))
    assert abs(r.mean() - 0.5) < 0.01 and abs(r.var() - 1 / 12) < 0.004
    c = lambda a, b: np.corrcoef(a.ravel(), b.ravel())[0, 1]  # noqa: E731
    assert abs(c(r[:, :, 1:], r[:, :, :-1])) < 0.03
    assert abs(c(r[1:], r[:-1])) < 0.03 and abs(c(r[:, 1:], r[:, :-1])) < 0.03
    # the store: floor(255 f) or one above, up with probability frac(255 f),
    # so its mean over the stream is 255 f (unbiased)
    us = r.ravel()
    for f in (0.0, 1.0, 0.5, 0.3, 0.123456, 0.9987, 1 / 510):
        q = np.array([L.or_u8_from_float_sr(f, float(x)) for x in us], np.float64)
        v = float(np.float32(255) * np.float32(f))
        assert set(np.unique(q)) <= {np.floor(v), np.floor(v) + 1}
        assert abs(q.mean() - v) < 4 * 0.5 / np.sqrt(us.size) + 1e-6, (f, q.mean(), v)
    assert L.or_u8_from_float_sr(0.5, 0.0) == 128 and L.or_u8_from_float_sr(0.5, 0.5) == 127


def test_u8_sweep_is_unbiased_against_fp32():
    # The UNORM8 sweep tracks the FP32 sweep of the same start within a few
    # steps of 1/255, and its mean error stays far below one step: stochastic
    # rounding carries updates smaller than half a step on average, where
    # round-to-nearest would drop them.
    S, bsx = 32, 8
    nb = S // bsx
    z, y, x = np.meshgrid(*[np.arange(S)] * 3, indexing="ij")
    f = (0.5 + 0.45 * np.sin(x / 4.0) * np.cos(y / 5.0 + z / 7.0)).astype(np.float32)
    q = np.rint(f * np.float32(255)).astype(np.uint8)
    f0 = q.astype(np.float32) * np.float32(1 / 255)  # the same start for both

    def level(vals, st):
        pages = np.stack([vals[(b // (nb * nb)) * bsx:(b // (nb * nb) + 1) * bsx,
                               ((b // nb) % nb) * bsx:((b // nb) % nb + 1) * bsx,
                               (b % nb) * bsx:(b % nb + 1) * bsx] for b in range(nb ** 3)])
        return O.Level(S, S, S, bsx, np.arange(nb ** 3, dtype=np.int32), pages, st)

    blocks = np.arange(nb ** 3, dtype=np.int32)
    lv = np.zeros(nb ** 3, np.uint16)
    L8, L32 = level(q, O.U8), level(f0, O.F32)
    for _ in range(5):
        O.mcf_sweeps(L8, blocks, lv, 2, 0.1, 1)
        O.mcf_sweeps(L32, blocks, lv, 2, 0.1, 1)
    assert L8.sr_step == 10  # the step advances with every sweep
    e = L8.pages.astype(np.float64) / 255 - L32.pages.astype(np.float64)
    assert np.abs(e).max() < 3 / 255
    assert abs(e.mean()) < 0.05 / 255
    moved = np.abs(L32.pages - level(f0, O.F32).pages) > 0.5 / 255
    assert moved.sum() > 1000 and abs(e[moved].mean()) < 0.1 / 255
