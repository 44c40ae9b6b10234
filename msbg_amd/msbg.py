"""Host-side mirror of the reference's MultiresSparseGrid interface for the
sparse-block stencil path, over the C ABI of libmsbg_gpu.so.

Method names and argument meanings follow the reference (msbg.h,
msbg_demo.cpp); reference paths are relative to the reference's src/
directory. All device memory is owned by the library; arrays crossing this
boundary are host numpy arrays.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib

_NP_DTYPE = {L.U16: np.uint16, L.F32: np.float32, L.F16: np.float16, L.BF16: np.uint16,
             L.F64: np.float64, L.U8: np.uint8}


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def storage_dtype(storage):
    return _NP_DTYPE[storage]


def getBlockColor8(bx, by, bz):
    """msbg.h:642-646."""
    return (bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2)


def scene_spheres(seed, nObj, cmin, cmax, rmin, rmax, nPts, sigma=0.5, ellipsoidFrac=0.0):
    """Seeded jittered sphere/ellipsoid surface points, (nObj*nPts, 3) float32."""
    out = np.empty((nObj * nPts, 3), dtype=np.float32)
    check(lib().msbg_scene_spheres(seed, nObj, cmin, cmax, rmin, rmax, nPts, sigma,
                                   ellipsoidFrac, _ptr(out)))
    return out


def scene_objects(seed, nObj, cmin, cmax, rmin, rmax, nPts, sigma=0.5, ellipsoidFrac=0.0,
                  perlinAmp=0.0, perlinFreq=0.0):
    """scene_spheres with Perlin-noise displaced surfaces, (nObj*nPts, 3) float32."""
    out = np.empty((nObj * nPts, 3), dtype=np.float32)
    check(lib().msbg_scene_objects(seed, nObj, cmin, cmax, rmin, rmax, nPts, sigma,
                                   ellipsoidFrac, perlinAmp, perlinFreq, _ptr(out)))
    return out


class DeviceArray:
    """Device memory held through the library (msbg_dev_alloc) with a numpy
    shape and dtype. data_ptr() is the device address the msbg_unet_* and
    msbg_cnn_* entry points take, so their callers need no other GPU runtime."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._p = C.c_void_p()
        check(lib().msbg_dev_alloc(self.nbytes, C.byref(self._p)))

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def data_ptr(self):
        return self._p.value or 0

    def rows(self, start, count=1):
        """View of rows [start, start + count) along the first axis."""
        return DeviceView(self, start, count)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(lib().msbg_dev_upload(self._p, _ptr(a), a.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().msbg_dev_download(_ptr(out), self._p, out.nbytes))
        return out

    def free(self):
        if self._p.value:
            check(lib().msbg_dev_free(self._p))
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


class DeviceView:
    """Rows [start, start + count) of a DeviceArray (no ownership)."""

    def __init__(self, base, start, count):
        assert 0 <= start and start + count <= base.shape[0]
        self.base = base
        self.shape = (int(count),) + base.shape[1:]
        self.dtype = base.dtype
        self._row = int(np.prod(base.shape[1:], dtype=np.int64)) * base.dtype.itemsize
        self._off = int(start) * self._row
        self.nbytes = int(count) * self._row

    def data_ptr(self):
        return self.base.data_ptr() + self._off

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(lib().msbg_dev_upload(C.c_void_p(self.data_ptr()), _ptr(a), a.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().msbg_dev_download(_ptr(out), C.c_void_p(self.data_ptr()), out.nbytes))
        return out


def _dptr(t):
    """Device address of a DeviceArray / DeviceView (or of any object with
    data_ptr(), whose pending writes the caller has completed)."""
    return C.c_void_p(t.data_ptr())


class UNet:
    """Learned warm start: a 3-stage 3D U-Net mapping 36^3 block windows to
    the block's change over K plain sweeps (no counterpart in the reference).
    Tensor arguments are DeviceArray / DeviceView objects (library-owned
    device memory): windows BF16 bits [n,36,36,36] (uint16), predictions /
    targets / output gradients FP32 [n,32,32,32]."""

    def __init__(self, seed=0):
        h = C.c_void_p()
        check(lib().msbg_unet_create(seed, C.byref(h)))
        self._h = h
        n = C.c_int64()
        check(lib().msbg_unet_num_params(h, C.byref(n)))
        self.nparam = n.value

    def get_params(self):
        out = np.empty(self.nparam, np.float32)
        check(lib().msbg_unet_get_params(self._h, _ptr(out), out.size))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().msbg_unet_set_params(self._h, _ptr(p), p.size))

    def get_grads(self):
        out = np.empty(self.nparam, np.float32)
        check(lib().msbg_unet_get_grads(self._h, _ptr(out), out.size))
        return out

    def forward(self, x, y):
        check(lib().msbg_unet_forward(self._h, _dptr(x), x.shape[0], _dptr(y)))

    def backward(self, x, dy):
        check(lib().msbg_unet_backward(self._h, _dptr(x), _dptr(dy), x.shape[0]))

    def train_step(self, x, target, lr, wait=True):
        """One Adam step; returns the loss (wait=True) or queues the step and
        adds its loss to the running sum of take_loss() (wait=False)."""
        loss = C.c_double()
        check(lib().msbg_unet_train_step(self._h, _dptr(x), _dptr(target), x.shape[0], lr,
                                         C.byref(loss) if wait else None))
        return loss.value if wait else None

    def train_step_rows(self, X, T, rows, lr, wait=False):
        """One Adam step on rows of resident sets X [N,36,36,36] (BF16 bits)
        and T [N,32,32,32]; queued unless wait=True."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        assert r.size and r.max() < X.shape[0] and r.max() < T.shape[0]
        loss = C.c_double()
        check(lib().msbg_unet_train_step_rows(self._h, _dptr(X), _dptr(T), _ptr(r), r.size, lr,
                                              C.byref(loss) if wait else None))
        return loss.value if wait else None

    def forward_rows(self, X, rows, y):
        r = np.ascontiguousarray(rows, dtype=np.int32)
        assert r.size == y.shape[0] and r.max() < X.shape[0]
        check(lib().msbg_unet_forward_rows(self._h, _dptr(X), _ptr(r), r.size, _dptr(y)))

    def take_loss(self):
        """(sum of the mean losses, steps) of the queued steps since the last call."""
        s, n = C.c_double(), C.c_int64()
        check(lib().msbg_unet_take_loss(self._h, C.byref(s), C.byref(n)))
        return s.value, n.value

    def warm_start(self, grid, chan, batch=64):
        """Add the predicted change to every level-0 value block; device ms."""
        ms = C.c_double()
        check(lib().msbg_warm_start(grid._h, chan, self._h, batch, C.byref(ms)))
        return ms.value

    @staticmethod
    def gather(grid, chan, blocks, x):
        b = np.ascontiguousarray(blocks, dtype=np.int32)
        check(lib().msbg_unet_gather(grid._h, chan, _ptr(b), b.size, _dptr(x)))

    @staticmethod
    def extract(grid, chan, blocks, y):
        b = np.ascontiguousarray(blocks, dtype=np.int32)
        check(lib().msbg_unet_extract(grid._h, chan, _ptr(b), b.size, _dptr(y)))

    @staticmethod
    def scatter(grid, chan, blocks, y):
        b = np.ascontiguousarray(blocks, dtype=np.int32)
        check(lib().msbg_unet_scatter(grid._h, chan, _ptr(b), b.size, _dptr(y)))

    def destroy(self):
        if self._h:
            lib().msbg_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()


def cnn_conv3d(x, cin, cout, dout, pad, w, b, relu, y):
    """One network convolution on device tensors (layer tests)."""
    check(lib().msbg_cnn_conv3d(_dptr(x), x.shape[0], cin, cout, dout, pad, _dptr(w),
                                None if b is None else _dptr(b), int(relu), _dptr(y)))


def cnn_conv3d_bwd_data(dy, cin, cout, din, pad, w, dx):
    check(lib().msbg_cnn_conv3d_bwd_data(_dptr(dy), dy.shape[0], cin, cout, din, pad, _dptr(w),
                                         _dptr(dx)))


def cnn_conv3d_bwd_weights(x, dy, cin, cout, dout, pad, dw, db):
    check(lib().msbg_cnn_conv3d_bwd_weights(_dptr(x), _dptr(dy), x.shape[0], cin, cout, dout, pad,
                                            _dptr(dw), None if db is None else _dptr(db)))


def copy_bandwidth(nbytes=1 << 30, reps=5):
    """Measured read+write GB/s of a plain device copy of nbytes."""
    v = C.c_double()
    check(lib().msbg_copy_bandwidth(nbytes, reps, C.byref(v)))
    return v.value


class MultiresSparseGrid:
    """MultiresSparseGrid (msbg.h:834-3907) restricted to the stencil path."""

    def __init__(self, handle):
        self._h = handle
        dims = np.zeros(8, dtype=np.int32)
        nb = C.c_int64()
        check(lib().msbg_get_dims(self._h, _ptr(dims), C.byref(nb)))
        self.sx, self.sy, self.sz, self.bsx0, self.nbx, self.nby, self.nbz, self.nLevels = (
            int(v) for v in dims)
        self._nblocks = nb.value
        self._storage = {}

    # -- lifetime ----------------------------------------------------------
    @classmethod
    def create(cls, name, sx, sy, sz, blockSize0=16, dTransRes=-1, maxSizeMB=0,
               initialSizeMB=-1, options=0, miscConf=None, pSolverConf=None, nLevels=-1):
        """MultiresSparseGrid::create (msbg.h:834-850). dTransRes, maxSizeMB,
        initialSizeMB and the solver configs do not affect this path."""
        h = C.c_void_p()
        check(lib().msbg_create(name.encode(), sx, sy, sz, blockSize0, options, nLevels,
                                C.byref(h)))
        return cls(h)

    def destroy(self):
        if self._h:
            lib().msbg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()

    # -- geometry ----------------------------------------------------------
    def nBlocks(self):
        return self._nblocks

    def getNumLevels(self):
        return self.nLevels

    def bsx(self, level=0):
        return self.bsx0 >> level

    def getBlockCoordsById(self, bid):
        """sbg.h:1789-1798."""
        bz, rem = divmod(bid, self.nbx * self.nby)
        by, bx = divmod(rem, self.nbx)
        return bx, by, bz

    def getBlockIndex(self, bx, by, bz):
        """sbg.h:1744-1749."""
        return (bz * self.nby + by) * self.nbx + bx

    def set_arith(self, mode):
        check(lib().msbg_set_arith(self._h, mode))

    def synchronize(self):
        check(lib().msbg_synchronize(self._h))

    # -- particles / refinement map (msbg_demo.cpp:339-441) -----------------
    def points_upload(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        nk = C.c_int64()
        check(lib().msbg_points_upload(self._h, _ptr(xyz), xyz.shape[0], C.byref(nk)))
        self._npoints = nk.value
        return nk.value

    def points_generate(self, seed, nObj, cmin, cmax, rmin, rmax, nPts, sigma=0.5,
                        ellipsoidFrac=0.0, perlinAmp=0.0, perlinFreq=0.0, append=False):
        """Scene points generated on the device (msbg_points_generate); returns
        the grid's particle count."""
        n = C.c_int64()
        check(lib().msbg_points_generate(self._h, seed, nObj, cmin, cmax, rmin, rmax, nPts, sigma,
                                         ellipsoidFrac, perlinAmp, perlinFreq, int(append),
                                         C.byref(n)))
        self._npoints = n.value
        return n.value

    def points_download(self):
        out = np.empty((self._npoints, 3), dtype=np.float32)
        check(lib().msbg_points_download(self._h, _ptr(out), out.shape[0]))
        return out

    def activate(self, rScan):
        """blockActive[] counts of msbg_demo.cpp:368-382."""
        counts = np.empty(self._nblocks, dtype=np.uint32)
        na = C.c_int64()
        check(lib().msbg_activate_from_points(self._h, rScan, _ptr(counts), C.byref(na)))
        return counts

    def levels_from_points(self, rScans):
        r = np.ascontiguousarray(rScans, dtype=np.float32)
        out = np.empty(self._nblocks, dtype=np.int32)
        check(lib().msbg_levels_from_points(self._h, _ptr(r), r.size, _ptr(out)))
        return out

    def regularizeRefinementMap(self, blockmapIn, blockmapOut=None):
        """msbg.cpp:1202-1423. Returns blockmapOut (a copy when None is given)."""
        a = np.ascontiguousarray(blockmapIn, dtype=np.int32)
        out = np.array(a, copy=True) if blockmapOut is None else blockmapOut
        check(lib().msbg_regularize_refinement_map(self._h, _ptr(a), _ptr(out)))
        return out

    def setRefinementMap(self, blockLevels=None, blockFlags=None):
        """msbg.cpp:1428-1745 at levelMg 0."""
        lv = None if blockLevels is None else np.ascontiguousarray(blockLevels, dtype=np.int32)
        fl = None if blockFlags is None else np.ascontiguousarray(blockFlags, dtype=np.uint32)
        check(lib().msbg_set_refinement_map(self._h, _ptr(lv), _ptr(fl)))

    def getBlockInfo(self, bid=None):
        """(levels, flags) of all blocks, or of block bid (msbg.h:945-950)."""
        lv = np.empty(self._nblocks, dtype=np.uint16)
        fl = np.empty(self._nblocks, dtype=np.uint16)
        check(lib().msbg_get_block_info(self._h, _ptr(lv), _ptr(fl)))
        if bid is None:
            return lv, fl
        return int(lv[bid]), int(fl[bid])

    def set_exists(self, exists=None):
        e = None if exists is None else np.ascontiguousarray(exists, dtype=np.uint8)
        check(lib().msbg_set_exists(self._h, _ptr(e)))

    def blocksFineCoarse(self, level):
        n = C.c_int64()
        check(lib().msbg_get_fine_coarse_blocks(self._h, level, None, 0, C.byref(n)))
        out = np.empty(n.value, dtype=np.int32)
        check(lib().msbg_get_fine_coarse_blocks(self._h, level, _ptr(out), n.value, None))
        return out

    # -- channels ------------------------------------------------------------
    def channel_define(self, chan, storage, level, kind, empty_raw=0.0, full_raw=None,
                       fill_raw=None):
        if full_raw is None:
            full_raw = 65535.0 if storage == L.U16 else 1.0
        if fill_raw is None:
            fill_raw = full_raw
        k = np.ascontiguousarray(kind, dtype=np.uint8)
        check(lib().msbg_channel_define(self._h, chan, storage, level, _ptr(k), empty_raw,
                                        full_raw, fill_raw))
        self._storage[chan] = storage

    def alloc_demo(self, chan, storage):
        """Allocation step of msbg_demo.cpp:451-511; returns activeBlocks."""
        out = np.empty(self._nblocks, dtype=np.int32)
        n = C.c_int64()
        check(lib().msbg_channel_alloc_demo(self._h, chan, storage, _ptr(out), out.size,
                                            C.byref(n)))
        self._storage[chan] = storage
        return out[:n.value].copy()

    def pagemap(self, chan, level=0):
        pm = np.empty(self._nblocks, dtype=np.int32)
        n = C.c_int64()
        check(lib().msbg_channel_get_pagemap(self._h, chan, level, _ptr(pm), C.byref(n)))
        return pm, n.value

    def download(self, chan, level=0):
        """All pages of (chan, level) as [npages, bsx, bsx, bsx] (z, y, x)."""
        st = self._storage[chan]
        _, np_ = self.pagemap(chan, level)
        b = self.bsx(level)
        out = np.empty((np_, b, b, b), dtype=_NP_DTYPE[st])
        check(lib().msbg_channel_download(self._h, chan, level, _ptr(out), out.nbytes))
        return out

    def upload(self, chan, level, pages):
        st = self._storage[chan]
        a = np.ascontiguousarray(pages, dtype=_NP_DTYPE[st])
        check(lib().msbg_channel_upload(self._h, chan, level, _ptr(a), a.nbytes))

    def copyDataTo(self, chanSrc, chanDst, storage, level=0):
        """SparseGrid::copyDataTo (sbg.h:550-557) into another storage format."""
        check(lib().msbg_channel_copy(self._h, chanSrc, chanDst, storage, level))
        self._storage[chanDst] = storage

    def diff(self, chanA, chanB, level=0):
        """(max |a-b|, RMS(a-b)) over the value voxels, FP64 on the device."""
        mx, rms = C.c_double(), C.c_double()
        check(lib().msbg_channel_diff(self._h, chanA, chanB, level, C.byref(mx), C.byref(rms)))
        return mx.value, rms.value

    def halo_table(self, chan, level=0):
        _, np_ = self.pagemap(chan, level)
        out = np.empty((np_, 27), dtype=np.int32)
        check(lib().msbg_get_halo_table(self._h, chan, level, _ptr(out), out.size))
        return out

    # -- hot path --------------------------------------------------------------
    def splat(self, chan, rParticle=2.0, nbDist=2.0):
        check(lib().msbg_splat_points(self._h, chan, rParticle, nbDist))

    def finalize(self, chan, rParticle=2.0, nbDist=2.0):
        check(lib().msbg_finalize_density(self._h, chan, rParticle, nbDist))

    def fillHaloBlock(self, chan, bid, dmax=1, opt=L.OPT_BC_NEUMANN):
        """HaloBlockSet::fillHaloBlock_ (halo.cpp:267-479): (hsx,hsx,hsx) float32."""
        level = int(self.getBlockInfo(bid)[0])
        hsx = self.bsx(level) + 2 * dmax
        out = np.empty((hsx, hsx, hsx), dtype=np.float32)
        check(lib().msbg_halo_fill(self._h, chan, bid, dmax, opt, _ptr(out)))
        return out

    def prepareGhostBlocks(self, chan):
        check(lib().msbg_prepare_ghost_blocks(self._h, chan))

    def applyChannelPdeFast(self, chPhi, chTmp, chTmp2, blockList, laplTyp, doConstrZeroOne,
                            densThr, densThrTol1, densThrTol2, numIter, gridSpacing=1.0,
                            timeStep=0.025, residual=False, lambda_=0.0, taubinStrength=0.0,
                            anisDiffAlpha=0.0):
        """msbg4.cpp:60-957 (msbg_apply_channel_pde). Returns the per-sweep FP64
        residuals if asked."""
        bl = None if blockList is None else np.ascontiguousarray(blockList, dtype=np.int32)
        nsweeps = numIter + (numIter % 2)
        res = np.zeros(max(nsweeps, 1), dtype=np.float64) if residual else None
        check(lib().msbg_apply_channel_pde(
            self._h, chPhi, chTmp, chTmp2, _ptr(bl), 0 if bl is None else bl.size, laplTyp,
            doConstrZeroOne, densThr, densThrTol1, densThrTol2, numIter, gridSpacing, timeStep,
            lambda_, taubinStrength, anisDiffAlpha, _ptr(res)))
        return res[:nsweeps] if residual else None

    def applyChannelPdeLevel(self, chan, level, haloOpt, blockList=None, numIter=2,
                             timeStep=0.1, doConstrZeroOne=1, residual=False):
        """8-colour MCF sweeps of the level-`level` blocks on that level's grid
        (msbg_apply_channel_pde_level; the reference sweeps level 0 only).
        haloOpt: OPT_BC_NEUMANN or OPT_BC_COARSE_LEVEL [| OPT_INTERPOLATE]."""
        bl = None if blockList is None else np.ascontiguousarray(blockList, dtype=np.int32)
        nsweeps = numIter + (numIter % 2)
        res = np.zeros(max(nsweeps, 1), dtype=np.float64) if residual else None
        check(lib().msbg_apply_channel_pde_level(
            self._h, chan, level, haloOpt, _ptr(bl), 0 if bl is None else bl.size, numIter,
            timeStep, doConstrZeroOne, _ptr(res)))
        return res[:nsweeps] if residual else None

    def set_tile(self, tile):
        check(lib().msbg_set_tile(self._h, tile))

    # -- field consumers ---------------------------------------------------------
    def interpolate(self, chan, xyz, kind=0, level=0):
        """Trilinear (kind 0, interpolateUint16ToFloatFast / interpolateFloatFast,
        sbg.h:4623-5031) or cubic B-spline with gradient (kind 1,
        interpolateWithDerivs0, sbg.h:3903-4135) at voxel coordinates of the
        level. Returns values (raw units for UNORM16), and gradients for kind 1."""
        p = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        val = np.empty(p.shape[0], dtype=np.float32)
        grad = np.empty((p.shape[0], 3), dtype=np.float32) if kind == 1 else None
        check(lib().msbg_interpolate(self._h, chan, level, kind, _ptr(p), p.shape[0], _ptr(val),
                                     _ptr(grad)))
        return val if kind == 0 else (val, grad)

    def render(self, chan, width=3840, height=2160, camPos=(.5, .8, .7), lookAt=(.5, .5, .5),
               zoom=1.0, sunPos=(.5, 5, 5), color=(0.8, 0.6, 0.4), timing=False):
        """RDR::RaymarchRenderer::render of the level-0 grid (render.cpp:19-100)
        with the demo's camera (msbg_demo.cpp:122-166): [height, width] uint32
        BMP_MKRGB words; with timing=True also the kernel's device ms."""
        f3 = lambda v: np.ascontiguousarray(v, dtype=np.float32)
        out = np.empty((height, width), dtype=np.uint32)
        ms = C.c_double()
        check(lib().msbg_render(self._h, chan, width, height, _ptr(f3(camPos)), _ptr(f3(lookAt)),
                                float(zoom), _ptr(f3(sunPos)), _ptr(f3(color)), _ptr(out),
                                C.byref(ms)))
        return (out, ms.value) if timing else out

    def render_counts(self, chan, width=3840, height=2160, camPos=(.5, .8, .7),
                      lookAt=(.5, .5, .5), zoom=1.0, sunPos=(.5, 5, 5), color=(0.8, 0.6, 0.4)):
        """render() through the counting kernel: (image, {advanced, samples, interp, hits})."""
        f3 = lambda v: np.ascontiguousarray(v, dtype=np.float32)
        out = np.empty((height, width), dtype=np.uint32)
        cnt = np.zeros(4, dtype=np.int64)
        check(lib().msbg_render_counts(self._h, chan, width, height, _ptr(f3(camPos)),
                                       _ptr(f3(lookAt)), float(zoom), _ptr(f3(sunPos)),
                                       _ptr(f3(color)), _ptr(out), _ptr(cnt)))
        return out, dict(advanced=int(cnt[0]), samples=int(cnt[1]), interp=int(cnt[2]),
                         hits=int(cnt[3]))

    def last_sweep_timing(self):
        t, a, n = C.c_double(), C.c_double(), C.c_int64()
        check(lib().msbg_last_sweep_timing(self._h, C.byref(t), C.byref(a), C.byref(n)))
        return t.value, a.value, n.value


class Batch:
    """Batched 8-colour MCF sweeps of independent grids (one graph launch).

    run(mode): EAGER (one launch per colour pass and kernel group), GRAPH
    (the same captured into one graph, or each group as the autotuner chose)
    or PERSISTENT (one graph launch, every group of block size 16 or 32 as one
    persistent launch with per-grid dependency waits)."""

    EAGER, GRAPH, PERSISTENT = 0, 1, 2

    def __init__(self):
        h = C.c_void_p()
        check(lib().msbg_batch_create(C.byref(h)))
        self._h = h
        self._grids = []  # keep members alive

    def add(self, grid, chan, timeStep=0.1, doConstrZeroOne=1):
        check(lib().msbg_batch_add(self._h, grid._h, chan, timeStep, doConstrZeroOne))
        self._grids.append(grid)

    def autotune(self, passes=3):
        check(lib().msbg_batch_autotune(self._h, passes))

    def tiles(self):
        n = C.c_int()
        check(lib().msbg_batch_get_tiles(self._h, None, None, None, 0, C.byref(n)))
        b, s, t = (np.zeros(n.value, np.int32) for _ in range(3))
        check(lib().msbg_batch_get_tiles(self._h, _ptr(b), _ptr(s), _ptr(t), n.value, None))
        return list(zip(b.tolist(), s.tolist(), t.tolist()))

    def modes(self):
        """Per kernel group: 1 if the tuned run uses the persistent launch."""
        n = C.c_int()
        check(lib().msbg_batch_get_modes(self._h, None, 0, C.byref(n)))
        m = np.zeros(n.value, np.int32)
        check(lib().msbg_batch_get_modes(self._h, _ptr(m), n.value, None))
        return m.tolist()

    def set_wait_limit(self, microseconds):
        """Wall-clock bound of one dependency wait of a persistent run."""
        check(lib().msbg_batch_set_wait_limit(self._h, microseconds))

    def run(self, numIter, useGraph=True):
        """useGraph: True/False (GRAPH/EAGER) or one of EAGER, GRAPH, PERSISTENT."""
        mode = int(useGraph) if isinstance(useGraph, bool) else useGraph
        ms = C.c_double()
        check(lib().msbg_batch_run(self._h, numIter, mode, C.byref(ms)))
        return ms.value

    def destroy(self):
        if self._h:
            lib().msbg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()


MCF_8COLOR = -(L.PDE_MEAN_CURVATURE + L.OPT_8_COLOR_SCHEME)
