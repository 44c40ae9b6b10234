"""Pressure hierarchies shared by the oracle pins and the GPU parity tests:
block levels, cell types and fields on the finest grid."""
import numpy as np

from oracle import ps as P


def levels_random(nb, nlev, seed):
    """Random block levels, regularized so face neighbours differ by <= 1."""
    rng = np.random.default_rng(seed)
    return P.regularize_levels(rng.integers(0, nlev, int(np.prod(nb))), nb)


def levels_by_x(nb, nlev):
    """Levels that change along x only (0 at the left, coarser to the right)."""
    nbx, nby, nbz = nb
    lx = np.minimum(np.arange(nbx) * nlev // max(nbx, 1), nlev - 1)
    return np.broadcast_to(lx[None, None, :], (nbz, nby, nbx)).ravel().astype(np.int8)


def cell_types(nb, bsx0, seed, solid=True, void=True, partial_rows=False):
    """Finest-grid cell types: fluid with a solid sphere and a void slab at
    the top (the free surface of a tank), both optional."""
    nbx, nby, nbz = nb
    sx, sy, sz = nbx * bsx0, nby * bsx0, nbz * bsx0
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(sz) + 0.5, np.arange(sy) + 0.5, np.arange(sx) + 0.5,
                          indexing="ij")
    t = np.zeros((sz, sy, sx), np.uint8)
    if solid:
        c = rng.uniform(0.3, 0.7, 3) * np.array([sx, sy, sz])
        r = rng.uniform(0.12, 0.22) * min(sx, sy, sz)
        t[(x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 < r * r] = P.SOLID
    if void:
        top = int(sz * rng.uniform(0.7, 0.85))
        t[top:] = np.where(t[top:] == P.SOLID, P.SOLID, P.VOID)
    if partial_rows:  # scattered single solid cells
        m = rng.random((sz, sy, sx)) < 0.01
        t[m] = P.SOLID
    return t


def pages_random(size, seed, scale=1.0):
    return (np.random.default_rng(seed).standard_normal(size) * scale).astype(np.float32)


def zero_nonfluid(values, flags):
    """The solver's invariant: non-fluid cells carry 0."""
    v = values.copy()
    v[(flags & (P.CELL_SOLID | P.CELL_VOID)) != 0] = 0
    return v


# configurations: (name, nb, bsx0, levels kind, seed, dbc, dense levels, solids/voids)
CASES = [
    ("uniform16", (3, 2, 2), 16, "zero", 1, (0,) * 6, 0, True, True),
    ("mixed16", (4, 3, 3), 16, "random", 2, (0, 0, 1, 1, 1, 0), 0, True, True),
    ("mixed8_dense", (4, 4, 4), 8, "random", 3, (0,) * 6, 2, True, True),
    ("mixed4_closed", (4, 4, 2), 4, "random", 4, (1,) * 6, 1, True, False),
    ("byx16_partial", (4, 2, 2), 16, "byx", 5, (0, 1, 0, 1, 1, 0), 1, True, True),
]


def build_inputs(case):
    name, nb, bsx0, lk, seed, dbc, ndense, solid, void = case
    nlev = int(np.log2(bsx0))
    if lk == "zero":
        lev = np.zeros(int(np.prod(nb)), np.int8)
    elif lk == "byx":
        lev = levels_by_x(nb, nlev)
    else:
        lev = levels_random(nb, nlev, seed)
    typ = cell_types(nb, bsx0, seed, solid=solid, void=void, partial_rows=(lk == "byx"))
    return dict(nb=nb, bsx0=bsx0, level=lev, cell_type=typ, dbc=dbc, ndense=ndense)
