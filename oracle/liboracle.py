"""ctypes binding of the CPU oracle (liboracle_strict.so / liboracle_fast.so).

Test infrastructure only: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg. Each wrapper names the C function it calls;
the C sources cite the reference lines they restate.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

U16, F32, F16, BF16, F64, U8 = 1, 2, 3, 4, 5, 6
OPT_INTERPOLATE, OPT_BC_NEUMANN, OPT_BC_COARSE_LEVEL = 1 << 7, 1 << 10, 1 << 26
NULL, EMPTY, FULL = -1, -2, -3
_NP = {U16: np.uint16, F32: np.float32, F16: np.uint16, BF16: np.uint16, F64: np.float64,
       U8: np.uint8}


class OrLevel(C.Structure):
    _fields_ = [("sx", C.c_int), ("sy", C.c_int), ("sz", C.c_int), ("bsx", C.c_int),
                ("nbx", C.c_int), ("nby", C.c_int), ("nbz", C.c_int),
                ("page", C.c_void_p), ("pages", C.c_void_p), ("storage", C.c_int),
                ("empty_raw", C.c_double), ("full_raw", C.c_double), ("sr_step", C.c_uint32)]


_libs = {}


def load(kind="strict"):
    if kind not in _libs:
        path = os.path.join(_HERE, f"liboracle_{kind}.so")
        if not os.path.exists(path):
            raise OSError(f"{path} not built; run `make -C oracle`")
        L = C.CDLL(path)
        P, I, I64, F, D = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_double
        sig = {
            "or_u16_to_float": (F, [C.c_uint16]),
            "or_u16_from_float_roundf": (C.c_uint16, [F]),
            "or_u16_from_float_rne": (C.c_uint16, [F]),
            "or_f16_from_float": (C.c_uint16, [F]),
            "or_f16_to_float": (F, [C.c_uint16]),
            "or_bf16_from_float": (C.c_uint16, [F]),
            "or_bf16_to_float": (F, [C.c_uint16]),
            "or_domain_bounds": (None, [I, I, I, P, P]),
            "or_domain_filter": (I64, [I, I, I, P, I64, P]),
            "or_activate": (None, [I, I, I, I, P, P, I64, F, P, P]),
            "or_levels_from_activity": (None, [P, I64, I, P]),
            "or_regularize": (I, [I, I, I, I, P, P]),
            "or_block_flags": (None, [I, I, I, I, P, P, P, P]),
            "or_color8": (I, [I, I, I]),
            "or_color_lists": (None, [I, I, P, I, P, P]),
            "or_splat_u16": (None, [P, P, P, I64, F]),
            "or_finalize_value": (C.c_uint16, [C.c_uint16, F, F]),
            "or_finalize_u16": (None, [P, I64, F, F]),
            "or_finalize_density_f": (F, [C.c_uint16, F, F]),
            "or_halo_fill": (None, [P, P, I, I, C.c_uint, P]),
            "or_halo_fill_f64": (None, [P, I, I, C.c_uint, P]),
            "or_mcf_sweeps": (I, [P, P, I, P, I, F, I, P, I]),
            "or_mcf_sweeps_level": (I, [P, P, P, I, P, I, C.c_uint, I, F, I, P, I]),
            "or_mcf_voxel": (F, [P, I, I, F, I]),
            "or_pde_sweeps": (I, [P, P, I, P, I, P, P, I]),
            "or_pde_voxel": (F, [I, P, I, I, F, F, F, I]),
            "or_fast_exp": (F, [F]),
            "or_mcf_voxel_f64": (D, [P, I, I, D, I]),
            "or_ghost_downsample": (None, [P, I, P]),
            "or_prepare_ghost_blocks": (None, [P, I, P, P, P]),
            "or_interpolate_ghost": (C.c_double, [P, F, F, F]),
            "or_u8_to_float": (F, [C.c_uint8]),
            "or_sr_uniform": (F, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "or_u8_from_float_sr": (C.c_uint8, [F, F]),
            "or_interp_linear": (F, [P, P]),
            "or_bspline3_weights": (None, [F, P, P]),
            "or_interp_points": (None, [P, I, P, I64, P, P]),
            "or_camera_basis": (None, [P, P, P, P, P]),
            "or_render": (None, [P, P, I, I, I, I, P, I]),
            "or_version": (C.c_uint64, []),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _libs[kind] = L
    return _libs[kind]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Level:
    """One level of one channel: block map + pages (numpy), as or_level."""

    def __init__(self, sx, sy, sz, bsx, pagemap, pages, storage, empty_raw=0.0, full_raw=None):
        self.sx, self.sy, self.sz, self.bsx = sx, sy, sz, bsx
        self.nbx, self.nby, self.nbz = sx // bsx, sy // bsx, sz // bsx
        self.page = np.ascontiguousarray(pagemap, dtype=np.int32)
        self.pages = np.ascontiguousarray(pages, dtype=_NP[storage])
        self.storage = storage
        self.empty_raw = empty_raw
        self.full_raw = ({U16: 65535.0, U8: 255.0}.get(storage, 1.0) if full_raw is None
                         else full_raw)
        # UNORM8: the sweep step of the next call's first sweep (each call
        # advances it by its sweep count, like a grid of the library)
        self.sr_step = 0

    def struct(self):
        return OrLevel(self.sx, self.sy, self.sz, self.bsx, self.nbx, self.nby, self.nbz,
                       _p(self.page), _p(self.pages), self.storage, self.empty_raw,
                       self.full_raw, self.sr_step & 0xffffffff)


def domain_bounds(sx, sy, sz):
    a, b = np.zeros(3, np.float32), np.zeros(3, np.float32)
    load().or_domain_bounds(sx, sy, sz, _p(a), _p(b))
    return a, b


def domain_filter(sx, sy, sz, xyz):
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    keep = np.zeros(xyz.shape[0], np.uint8)
    load().or_domain_filter(sx, sy, sz, _p(xyz), xyz.shape[0], _p(keep))
    return keep


def activate(bsx, nbx, nby, nbz, xyz, rScan, keep=None):
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    cnt = np.zeros(nbx * nby * nbz, np.uint32)
    home = np.zeros(xyz.shape[0], np.int32)
    k = None if keep is None else np.ascontiguousarray(keep, dtype=np.uint8)
    load().or_activate(bsx, nbx, nby, nbz, _p(xyz), _p(k), xyz.shape[0], rScan, _p(cnt), _p(home))
    return cnt, home


def levels_from_activity(count, nLevels):
    count = np.ascontiguousarray(count, dtype=np.uint32)
    out = np.zeros(count.size, np.int32)
    load().or_levels_from_activity(_p(count), count.size, nLevels, _p(out))
    return out


def regularize(nbx, nby, nbz, nLevels, levels):
    a = np.ascontiguousarray(levels, dtype=np.int32)
    out = a.copy()
    rc = load().or_regularize(nbx, nby, nbz, nLevels, _p(a), _p(out))
    return out, rc


def block_flags(nbx, nby, nbz, nLevels, levels, inFlags=None):
    a = np.ascontiguousarray(levels, dtype=np.int32)
    f = None if inFlags is None else np.ascontiguousarray(inFlags, dtype=np.uint32)
    lv = np.zeros(a.size, np.uint16)
    fl = np.zeros(a.size, np.uint16)
    load().or_block_flags(nbx, nby, nbz, nLevels, _p(a), _p(f), _p(lv), _p(fl))
    return lv, fl


def color_lists(nbx, nby, blocks):
    b = np.ascontiguousarray(blocks, dtype=np.int32)
    out = np.zeros(b.size, np.int32)
    off = np.zeros(9, np.int32)
    load().or_color_lists(nbx, nby, _p(b), b.size, _p(out), _p(off))
    return out, off


def splat_u16(level, xyz, rScan, home=None):
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    s = level.struct()
    h = None if home is None else np.ascontiguousarray(home, dtype=np.int32)
    load().or_splat_u16(C.byref(s), _p(xyz), _p(h), xyz.shape[0], rScan)


def finalize_u16(pages, rParticle=2.0, nbDist=2.0):
    load().or_finalize_u16(_p(pages), pages.size, rParticle, nbDist)


def finalize_table(rParticle=2.0, nbDist=2.0):
    L = load()
    return np.array([L.or_finalize_value(u, rParticle, nbDist) for u in range(65536)], np.uint16)


def finalize_density_table(rParticle=2.0, nbDist=2.0):
    L = load()
    return np.array([L.or_finalize_density_f(u, rParticle, nbDist) for u in range(65536)],
                    np.float32)


def halo_fill(level, bid, dmax=1, opt=OPT_BC_NEUMANN, lo=None):
    hsx = level.bsx + 2 * dmax
    out = np.zeros((hsx, hsx, hsx), np.float32)
    s = level.struct()
    so = lo.struct() if lo is not None else None
    load().or_halo_fill(C.byref(s), C.byref(so) if so is not None else None, bid, dmax, opt,
                        _p(out))
    return out


MODE_COLOR8, MODE_REDBLACK, MODE_JACOBI = 0, 1, 2


class PdeParams(C.Structure):
    _fields_ = [("op", C.c_int), ("mode", C.c_int), ("dt", C.c_float),
                ("doConstrZeroOne", C.c_int), ("densThr", C.c_float), ("densThrTol1", C.c_float),
                ("densThrTol2", C.c_float), ("taubin", C.c_float), ("alpha", C.c_float)]


def pde_sweeps(level, blocks, levels, numIter, op, mode=MODE_COLOR8, dt=0.1, doConstrZeroOne=1,
               densThr=0.0, densThrTol1=0.0, densThrTol2=0.0, taubin=0.0, alpha=0.0,
               residual=False, nthreads=0):
    """or_pde_sweeps: applyChannelPdeFast's stencils 1/2/3/4/14/15/17 in
    colour, red-black or Jacobi mode; pages of `level` are updated."""
    b = np.ascontiguousarray(blocks, dtype=np.int32)
    lv = None if levels is None else np.ascontiguousarray(levels, dtype=np.uint16)
    n = numIter + (numIter % 2)
    res = np.zeros(max(n, 1), np.float64) if residual else None
    P = PdeParams(op, mode, dt, doConstrZeroOne, densThr, densThrTol1, densThrTol2, taubin, alpha)
    s = level.struct()
    rc = load().or_pde_sweeps(C.byref(s), _p(b), b.size, _p(lv), numIter, C.byref(P), _p(res),
                              nthreads)
    if rc < 0:
        raise ValueError("or_pde_sweeps: unsupported combination")
    level.sr_step += n
    return res[:n] if residual else None


def pde_voxel(op, h, hid, dt, taubin=0.0, alpha=0.0, it=0):
    h = np.ascontiguousarray(h, dtype=np.float32)
    return load().or_pde_voxel(op, _p(h), h.shape[-1], hid, dt, taubin, alpha, it)


def fast_exp(x):
    return load().or_fast_exp(float(x))


def mcf_sweeps(level, blocks, levels, numIter, dt, doConstrZeroOne=1, residual=False,
               nthreads=0, kind="strict"):
    b = np.ascontiguousarray(blocks, dtype=np.int32)
    lv = None if levels is None else np.ascontiguousarray(levels, dtype=np.uint16)
    n = numIter + (numIter % 2)
    res = np.zeros(max(n, 1), np.float64) if residual else None
    s = level.struct()
    load(kind).or_mcf_sweeps(C.byref(s), _p(b), b.size, _p(lv), numIter, dt, doConstrZeroOne,
                             _p(res), nthreads)
    level.sr_step += n
    return res[:n] if residual else None


def mcf_sweeps_level(level, lo, blocks, levels, sweepLevel, opt, numIter, dt,
                     doConstrZeroOne=1, residual=False, nthreads=0, kind="strict"):
    """or_mcf_sweeps_level: 8-colour MCF sweeps of the blocks at sweepLevel on
    that level's pages, halo rule `opt` (NEUMANN, or COARSE_LEVEL [| INTERPOLATE]
    reading `lo`, the next coarser level)."""
    b = np.ascontiguousarray(blocks, dtype=np.int32)
    lv = np.ascontiguousarray(levels, dtype=np.uint16)
    n = numIter + (numIter % 2)
    res = np.zeros(max(n, 1), np.float64) if residual else None
    s = level.struct()
    so = lo.struct() if lo is not None else None
    rc = load(kind).or_mcf_sweeps_level(C.byref(s), C.byref(so) if so is not None else None, _p(b),
                                    b.size, _p(lv), sweepLevel, opt, numIter, dt,
                                    doConstrZeroOne, _p(res), nthreads)
    if rc < 0:
        raise ValueError("or_mcf_sweeps_level: unsupported combination")
    level.sr_step += n
    return res[:n] if residual else None


def mcf_voxel(h, hid, dt, doConstrZeroOne=1):
    h = np.ascontiguousarray(h, dtype=np.float32)
    return load().or_mcf_voxel(_p(h), h.shape[-1], hid, dt, doConstrZeroOne)


def ghost_downsample(halo, bsxHi):
    halo = np.ascontiguousarray(halo, dtype=np.float32)
    out = np.zeros((bsxHi // 2,) * 3, np.float32)
    load().or_ghost_downsample(_p(halo), bsxHi, _p(out))
    return out


def interpolate_ghost(level, p):
    """or_interpolate_ghost: interpolateGhost of `level` at p (voxel
    coordinates, cell centres at i + 0.5), in the level's raw domain."""
    s = level.struct()
    return load().or_interpolate_ghost(C.byref(s), float(p[0]), float(p[1]), float(p[2]))


def prepare_ghost_blocks(levels, fc_lists, flags):
    """levels: list of Level (one per resolution level, all defined)."""
    arr = (OrLevel * len(levels))(*[lv.struct() for lv in levels])
    fc = np.concatenate([np.asarray(x, np.int32) for x in fc_lists] + [np.zeros(0, np.int32)])
    off = np.zeros(len(fc_lists) + 1, np.int32)
    off[1:] = np.cumsum([len(x) for x in fc_lists])
    fl = np.ascontiguousarray(flags, dtype=np.uint16)
    load().or_prepare_ghost_blocks(arr, len(levels), _p(fc), _p(off), _p(fl))


# ---- field consumers: interpolation and the ray-march renderer ----

def interp_linear(level, pts, kind="strict"):
    """or_interp_points kind 0: trilinear interpolation at voxel coordinates
    (raw UNORM16 units for UNORM16 levels)."""
    pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
    val = np.zeros(pts.shape[0], np.float32)
    s = level.struct()
    load(kind).or_interp_points(C.byref(s), 0, _p(pts), pts.shape[0], _p(val), None)
    return val


def interp_bspline3(level, pts, kind="strict"):
    """or_interp_points kind 1: cubic B-spline value and gradient."""
    pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
    val = np.zeros(pts.shape[0], np.float32)
    grad = np.zeros((pts.shape[0], 3), np.float32)
    s = level.struct()
    load(kind).or_interp_points(C.byref(s), 1, _p(pts), pts.shape[0], _p(val), _p(grad))
    return val, grad


def bspline3_weights(t):
    w, d = np.zeros(4, np.float32), np.zeros(4, np.float32)
    load().or_bspline3_weights(float(t), _p(w), _p(d))
    return w, d


class OrCamera(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("look", C.c_double * 3), ("focal", C.c_double),
                ("sun", C.c_double * 3), ("color", C.c_double * 3)]


def camera(pos, look, focal, sun, color):
    """Camera as the demo builds it: every input is a float widened to double."""
    f = lambda v: [float(np.float32(x)) for x in v]
    return OrCamera((C.c_double * 3)(*f(pos)), (C.c_double * 3)(*f(look)),
                    float(np.float32(focal)), (C.c_double * 3)(*f(sun)),
                    (C.c_double * 3)(*f(color)))


def camera_basis(pos, look):
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (pos, look)]
    out = [np.zeros(3, np.float64) for _ in range(3)]
    load().or_camera_basis(_p(a[0]), _p(a[1]), *[_p(o) for o in out])
    return out


def render(level, cam, W, H, y0=0, y1=None, nthreads=0, kind="strict"):
    """or_render: rows y0..y1-1 of the W x H image, uint32 BMP_MKRGB words
    (nthreads 0: OpenMP's default, OMP_NUM_THREADS)."""
    y1 = H if y1 is None else y1
    out = np.zeros((y1 - y0, W), np.uint32)
    s = level.struct()
    load(kind).or_render(C.byref(s), C.byref(cam), W, H, y0, y1, _p(out), nthreads)
    return out
