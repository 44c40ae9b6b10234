"""Host mirror of MultiresSparseGrid's pressure relax (msbg3.cpp) over the
msbg_ps_* C ABI.

PressureGrid holds one hierarchy: a block grid whose blocks each sit at their
own resolution level, cell flags / diagonal / face areas per multigrid level,
and the channels X (unknown), B (right-hand side), R (residual) and T
(Jacobi temporary). Method names and argument meanings follow the
reference: relax(boundaryZoneOnly, reverseOrder, levelMg, chanX, chanB,
chanTmp, nIter) (msbg3.cpp:2244) and multiplyLaplacianMatrixOpt(options,
levelMg, chanX, chanB, kap, chanDstY, pDotprod) (msbg3.cpp:1756).
"""
import ctypes as C

import numpy as np

from . import _lib

CH_X, CH_B, CH_R, CH_T, CH_D = 0, 1, 2, 3, 4
FLUID, SOLID, VOID = 0, 1, 2
DBC_OPEN, DBC_SOLID = 0, 1
OPT_BOUNDZONE_ONLY, OPT_CALC_RESIDUAL, OPT_REVERSE_ORDER = 1 << 4, 1 << 12, 1 << 18
CELL_SOLID, CELL_PARTIAL_SOLID, CELL_VOID = 1 << 0, 1 << 6, 1 << 12
BLK_NO_FLUID, BLK_CUTS_SOLID = 1 << 5, 1 << 15
OMEGA = np.float32(6.0 / 7.0)  # _mgSmOmega (msbg.cpp:819)
BLOCK_ITER = 5                 # _mgSmBlockIter (msbg.cpp:837)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class PressureGrid:
    """One pressure hierarchy on the device.

    nb = (nbx, nby, nbz) blocks of bsx0^3 finest cells; level[nb] (block-id
    order, x fastest) the blocks' resolution levels; cell_type the finest
    grid's FLUID/SOLID/VOID, shape (nbz*bsx0, nby*bsx0, nbx*bsx0)."""

    def __init__(self, nb, bsx0, level, cell_type, dx0=None, dbc=(DBC_OPEN,) * 6,
                 omega=OMEGA, block_iter=BLOCK_ITER, dense_levels=0):
        L = _lib.lib()
        self.nb = tuple(int(v) for v in nb)
        self.bsx0 = int(bsx0)
        nbx, nby, nbz = self.nb
        d = _lib.PsDesc()
        d.nbx, d.nby, d.nbz, d.bsx0 = nbx, nby, nbz, self.bsx0
        d.dx0 = 1.0 / (nbx * self.bsx0) if dx0 is None else float(dx0)
        for a in range(3):
            for s in range(2):
                d.domainBc[a][s] = int(dbc[2 * a + s])
        d.omega = float(omega)
        d.blockIter = int(block_iter)
        d.denseLevels = int(dense_levels)
        self.dx0 = d.dx0
        lv = np.ascontiguousarray(level, dtype=np.int8).ravel()
        ct = np.ascontiguousarray(cell_type, dtype=np.uint8)
        if lv.size != nbx * nby * nbz or ct.shape != (nbz * bsx0, nby * bsx0, nbx * bsx0):
            raise ValueError("level / cell_type shape does not match the block grid")
        self.h = C.c_void_p()
        _lib.check(L.msbg_ps_create(C.byref(d), _ptr(lv), _ptr(ct), C.byref(self.h)))
        nmg, nlev = C.c_int(), C.c_int()
        _lib.check(L.msbg_ps_get_levels(self.h, C.byref(nmg), C.byref(nlev)))
        self.nmg, self.nlev = nmg.value, nlev.value

    def close(self):
        if self.h:
            _lib.lib().msbg_ps_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self, levelMg):
        n = C.c_int64()
        _lib.check(_lib.lib().msbg_ps_level_size(self.h, levelMg, C.byref(n)))
        return n.value

    def upload(self, chan, levelMg, values):
        a = np.ascontiguousarray(values, dtype=np.float32)
        if a.size != self.size(levelMg):
            raise ValueError("channel size mismatch")
        _lib.check(_lib.lib().msbg_ps_upload(self.h, chan, levelMg, _ptr(a)))

    def download(self, chan, levelMg):
        out = np.empty(self.size(levelMg), np.float32)
        _lib.check(_lib.lib().msbg_ps_download(self.h, chan, levelMg, _ptr(out)))
        return out

    def flags(self, levelMg):
        out = np.empty(self.size(levelMg), np.uint16)
        _lib.check(_lib.lib().msbg_ps_download_flags(self.h, levelMg, _ptr(out), None))
        return out

    def block_flags(self, levelMg):
        out = np.empty(int(np.prod(self.nb)), np.uint16)
        _lib.check(_lib.lib().msbg_ps_download_flags(self.h, levelMg, None, _ptr(out)))
        return out

    def fill(self, chan, levelMg, v=0.0):
        _lib.check(_lib.lib().msbg_ps_fill(self.h, chan, levelMg, float(v)))

    def relax(self, boundaryZoneOnly, reverseOrder, levelMg, chanX, chanB, chanTmp, nIter):
        _lib.check(_lib.lib().msbg_ps_relax(self.h, int(boundaryZoneOnly), int(reverseOrder),
                                            levelMg, chanX, chanB, chanTmp, nIter))

    def multiplyLaplacianMatrixOpt(self, options, levelMg, chanX, chanB, kap, chanDstY,
                                   want_dot=False):
        dot = C.c_double(0)
        _lib.check(_lib.lib().msbg_ps_multiply(self.h, int(options), levelMg, chanX, chanB,
                                               float(kap), chanDstY,
                                               C.byref(dot) if want_dot else None))
        return dot.value if want_dot else None

    def restrict(self, levelMg, chanSrc, chanDst):
        _lib.check(_lib.lib().msbg_ps_restrict(self.h, levelMg, chanSrc, chanDst))

    def prolongate(self, levelMg, chanSrc, chanDst):
        _lib.check(_lib.lib().msbg_ps_prolongate(self.h, levelMg, chanSrc, chanDst))

    def vcycle(self, nu1=2, nu2=2, coarseIter=50):
        _lib.check(_lib.lib().msbg_ps_vcycle(self.h, nu1, nu2, coarseIter))

    def synchronize(self):
        _lib.check(_lib.lib().msbg_ps_synchronize(self.h))

    def last_timing(self):
        """Device milliseconds of the last relax / multiply / vcycle call."""
        ms = C.c_double()
        _lib.check(_lib.lib().msbg_ps_last_timing(self.h, C.byref(ms)))
        return ms.value

    def relaxed_cells(self, levelMg):
        """Cells one relax iteration updates at levelMg (relax()'s nTotalCells)."""
        n = C.c_int64()
        _lib.check(_lib.lib().msbg_ps_relaxed_cells(self.h, levelMg, C.byref(n)))
        return n.value


def tank_scene(nb, bsx0, seed=1, fill=0.78, radius=0.17):
    """Synthetic pressure workload: a tank filled with fluid up to fill * the
    domain height (void above), a solid sphere of radius * the smallest
    extent at a seeded position in the fluid. Block levels follow the
    chessboard block distance d to the blocks that hold fluid next to solid
    or void: level max(0, min(d - 1, 3)), so face neighbours differ by one
    level at most. Returns (levels int8 [nb], cell types uint8 [z][y][x])."""
    nbx, nby, nbz = nb
    sx, sy, sz = nbx * bsx0, nby * bsx0, nbz * bsx0
    rng = np.random.default_rng(seed)
    t = np.zeros((sz, sy, sx), np.uint8)
    t[int(fill * sz):] = VOID
    c = np.array([rng.uniform(0.3, 0.7) * sx, rng.uniform(0.3, 0.7) * sy,
                  rng.uniform(0.25, 0.5) * sz])
    r = radius * min(sx, sy, sz)
    zz = (np.arange(sz) + 0.5 - c[2]) ** 2
    yy = (np.arange(sy) + 0.5 - c[1]) ** 2
    xx = (np.arange(sx) + 0.5 - c[0]) ** 2
    for z in range(sz):
        if zz[z] < r * r:
            m = (yy[:, None] + xx[None, :] + zz[z]) < r * r
            t[z][m] = SOLID
    tb = t.reshape(nbz, bsx0, nby, bsx0, nbx, bsx0)
    fl = (tb == FLUID).any(axis=(1, 3, 5))
    band = fl & ((tb == SOLID).any(axis=(1, 3, 5)) | (tb == VOID).any(axis=(1, 3, 5)))
    nlev = int(np.log2(bsx0))
    d = np.where(band, 0, nlev + 1).astype(np.int64)
    for _ in range(nlev + 1):  # chessboard distance by repeated 3x3x3 minimum
        p = np.pad(d, 1, constant_values=nlev + 1)
        m = d.copy()
        for dz in range(3):
            for dy in range(3):
                for dx in range(3):
                    m = np.minimum(m, p[dz:dz + nbz, dy:dy + nby, dx:dx + nbx] + 1)
        d = m
    lev = np.clip(d - 1, 0, nlev - 1).astype(np.int8)
    return lev.ravel(), t
