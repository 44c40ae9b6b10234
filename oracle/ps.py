"""CPU oracle of the adaptive pressure relax (test infrastructure only).

ctypes view of libps_oracle.so (ps_oracle.c): set-up of cell flags,
diagonal and face areas on a block grid with per-block resolution levels,
the reference's block relax, Laplacian multiply and dense-level Jacobi, and
the library's own constant restriction/prolongation and V-cycle.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

FLUID, SOLID, VOID = 0, 1, 2
CELL_SOLID, CELL_PARTIAL_SOLID, CELL_VOID = 1 << 0, 1 << 6, 1 << 12
BLK_NO_FLUID, BLK_CUTS_SOLID = 1 << 5, 1 << 15
DBC_OPEN, DBC_SOLID = 0, 1
OPT_CALC_RESIDUAL, OPT_REVERSE_ORDER = 1 << 12, 1 << 18
CH_X, CH_B, CH_R, CH_T, CH_D = 0, 1, 2, 3, 4
OMEGA = np.float32(6.0 / 7.0)  # _mgSmOmega, msbg.cpp:819
BLOCK_ITER = 5                 # _mgSmBlockIter, msbg.cpp:837

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libps_oracle.so")
        if not os.path.exists(path):
            raise OSError(f"{path} not built; run `make -C oracle`")
        L = C.CDLL(path)
        P, I, I64, D = C.c_void_p, C.c_int, C.c_int64, C.c_double
        sig = {
            "or_ps_create": (P, [I, I, I, I, D, P, C.c_float, I, I, P, P, P]),
            "or_ps_destroy": (None, [P]),
            "or_ps_num_levels": (I, [P, P, P]),
            "or_ps_level_size": (I64, [P, I]),
            "or_ps_get": (I, [P, I, I, P]),
            "or_ps_set": (I, [P, I, I, P]),
            "or_ps_get_flags": (I, [P, I, P]),
            "or_ps_get_block_flags": (I, [P, I, P]),
            "or_ps_relax": (I, [P, I, I, I, I, I, I]),
            "or_ps_multiply": (I, [P, C.c_uint, I, I, I, I, P]),
            "or_ps_restrict": (I, [P, I, I, I]),
            "or_ps_prolongate": (I, [P, I, I, I]),
            "or_ps_fill": (I, [P, I, I, C.c_float]),
            "or_ps_vcycle": (I, [P, I, I, I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class PsOracle:
    """One pressure hierarchy: block grid (nbx, nby, nbz) of bsx0^3 blocks,
    block levels `level` (int8, block-id order, x fastest) and the finest
    grid's cell types `cell_type` (uint8, shape (sz, sy, sx))."""

    def __init__(self, nb, bsx0, level, cell_type, dx0=None, dbc=(DBC_OPEN,) * 6,
                 omega=OMEGA, block_iter=BLOCK_ITER, ndense=0):
        L = lib()
        self.nb = tuple(nb)
        self.bsx0 = bsx0
        nbx, nby, nbz = self.nb
        self.dx0 = 1.0 / (nbx * bsx0) if dx0 is None else dx0
        self._level = np.ascontiguousarray(level, dtype=np.int8).ravel()
        self._type = np.ascontiguousarray(cell_type, dtype=np.uint8)
        assert self._type.shape == (nbz * bsx0, nby * bsx0, nbx * bsx0)
        self._dbc = np.ascontiguousarray(dbc, dtype=np.int32)
        err = C.c_int(0)
        self.h = L.or_ps_create(nbx, nby, nbz, bsx0, self.dx0, _p(self._dbc), float(omega),
                                block_iter, ndense, _p(self._level), _p(self._type),
                                C.byref(err))
        if not self.h or err.value:
            raise ValueError("invalid pressure hierarchy")
        nmg, nlev = C.c_int(), C.c_int()
        L.or_ps_num_levels(self.h, C.byref(nmg), C.byref(nlev))
        self.nmg, self.nlev = nmg.value, nlev.value

    def close(self):
        if self.h:
            lib().or_ps_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def size(self, m):
        return lib().or_ps_level_size(self.h, m)

    def get(self, chan, m):
        out = np.empty(self.size(m), np.float32)
        assert lib().or_ps_get(self.h, chan, m, _p(out)) == 0
        return out

    def set(self, chan, m, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.size == self.size(m)
        assert lib().or_ps_set(self.h, chan, m, _p(a)) == 0

    def flags(self, m):
        out = np.empty(self.size(m), np.uint16)
        assert lib().or_ps_get_flags(self.h, m, _p(out)) == 0
        return out

    def block_flags(self, m):
        out = np.empty(int(np.prod(self.nb)), np.uint16)
        assert lib().or_ps_get_block_flags(self.h, m, _p(out)) == 0
        return out

    def relax(self, m, n_iter, reverse=0, chan_x=CH_X, chan_b=CH_B, chan_tmp=CH_T):
        assert lib().or_ps_relax(self.h, reverse, m, chan_x, chan_b, chan_tmp, n_iter) == 0

    def multiply(self, m, options=0, chan_x=CH_X, chan_b=CH_B, chan_dst=CH_R):
        dot = C.c_double(0)
        assert lib().or_ps_multiply(self.h, options, m, chan_x, chan_b, chan_dst,
                                    C.byref(dot)) == 0
        return dot.value

    def restrict(self, m, chan_src=CH_R, chan_dst=CH_B):
        assert lib().or_ps_restrict(self.h, m, chan_src, chan_dst) == 0

    def prolongate(self, m, chan_src=CH_X, chan_dst=CH_X):
        assert lib().or_ps_prolongate(self.h, m, chan_src, chan_dst) == 0

    def fill(self, chan, m, v=0.0):
        assert lib().or_ps_fill(self.h, chan, m, v) == 0

    def vcycle(self, nu1=2, nu2=2, coarse_iter=50):
        assert lib().or_ps_vcycle(self.h, nu1, nu2, coarse_iter) == 0


# ------------------------------------------------------------ layout helpers

def block_levels_at(level, m):
    return np.maximum(np.asarray(level, np.int64).ravel(), m)


def page_offsets(level, bsx0, m):
    """Offsets of every block's page in a block-level channel of mg level m."""
    L = block_levels_at(level, m)
    sizes = (bsx0 >> L) ** 3
    return np.concatenate([[0], np.cumsum(sizes)[:-1]]), L


def pages_from_field(f, nb, bsx0, level, m):
    """Block-level channel (mg level m) sampled from per-level global fields:
    f(l) returns the global array of level l, shape (sz_l, sy_l, sx_l)."""
    nbx, nby, nbz = nb
    L = block_levels_at(level, m)
    out = []
    cache = {}
    for b in range(nbx * nby * nbz):
        l = int(L[b])
        if l not in cache:
            cache[l] = np.asarray(f(l), np.float32)
        g = cache[l]
        bs = bsx0 >> l
        bx, by, bz = b % nbx, (b // nbx) % nby, b // (nbx * nby)
        out.append(g[bz * bs:(bz + 1) * bs, by * bs:(by + 1) * bs, bx * bs:(bx + 1) * bs].ravel())
    return np.concatenate(out).astype(np.float32)


def field_from_pages(pages, nb, bsx0, level, m, l):
    """Global array at level l of the blocks whose page at mg level m is at
    level l (NaN elsewhere)."""
    nbx, nby, nbz = nb
    off, L = page_offsets(level, bsx0, m)
    bs = bsx0 >> l
    g = np.full((nbz * bs, nby * bs, nbx * bs), np.nan, np.float64)
    for b in range(nbx * nby * nbz):
        if L[b] != l:
            continue
        bx, by, bz = b % nbx, (b // nbx) % nby, b // (nbx * nby)
        g[bz * bs:(bz + 1) * bs, by * bs:(by + 1) * bs, bx * bs:(bx + 1) * bs] = \
            pages[off[b]:off[b] + bs ** 3].reshape(bs, bs, bs)
    return g


def regularize_levels(level, nb):
    """Lower block levels until face neighbours differ by at most one."""
    nbx, nby, nbz = nb
    lv = np.asarray(level, np.int64).reshape(nbz, nby, nbx).copy()
    changed = True
    while changed:
        changed = False
        for ax in range(3):
            for sh in (1, -1):
                n = np.roll(lv, sh, axis=ax)
                # edges wrap in roll: mask the wrapped slab
                sl = [slice(None)] * 3
                sl[ax] = 0 if sh == 1 else -1
                n[tuple(sl)] = lv[tuple(sl)]
                bad = lv > n + 1
                if bad.any():
                    lv[bad] = n[bad] + 1
                    changed = True
    return lv.ravel().astype(np.int8)
