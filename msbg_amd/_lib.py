"""ctypes binding of libmsbg_gpu.so (the C ABI declared in include/msbg_gpu.h).

The library is loaded from the package directory only; there is no fallback:
if it is missing, importing the binding raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmsbg_gpu.so")

# storage formats
U16, F32, F16, BF16, F64, U8 = 1, 2, 3, 4, 5, 6
STORAGE_BYTES = {U16: 2, F32: 4, F16: 2, BF16: 2, F64: 8, U8: 1}
STORAGE_NAMES = {U16: "u16", F32: "f32", F16: "f16", BF16: "bf16", F64: "f64", U8: "u8"}

# reference option values (msbg.h:101-391)
OPT_INTERPOLATE = 1 << 7
OPT_BC_NEUMANN = 1 << 10
OPT_SINGLE_LEVEL = 1 << 14
OPT_BC_COARSE_LEVEL = 1 << 26
OPT_8_COLOR_SCHEME = 1 << 29
PDE_MEAN_CURVATURE = 4
IP_LINEAR, IP_BSPLINE3 = 0, 1  # msbg_interpolate kinds
CH_NULL, CH_FLOAT_1, CH_FLOAT_2, CH_FLOAT_3, CH_UINT16_1, CH_UINT8_1 = 0, 1, 6, 7, 30, 38

BLK_DOM_BORDER = 1 << 1
BLK_COARSE_FINE = 1 << 2
BLK_FINE_COARSE = 1 << 3
BLK_EXISTS = 1 << 12
BLK_UNIFORM_EMPTY = 1 << 13

BLOCK_NULL, BLOCK_EMPTY, BLOCK_FULL, BLOCK_OUT = -1, -2, -3, -4
KIND_NULL, KIND_EMPTY, KIND_FULL, KIND_VALUE = 0, 1, 2, 3
ARITH_EXACT, ARITH_FAST = 0, 1

EINVAL, ENOMEM, EDEVICE, ESTATE, EUNSUPP = -1, -2, -3, -4, -5
ERRORS = {EINVAL: "EINVAL", ENOMEM: "ENOMEM", EDEVICE: "EDEVICE", ESTATE: "ESTATE",
          EUNSUPP: "EUNSUPP"}


class MsbgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_U64 = C.c_uint64
_F = C.c_float
_D = C.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "msbg_last_error": (C.c_char_p, []),
    "msbg_create": (_I, [C.c_char_p, _I, _I, _I, _I, C.c_uint, _I, C.POINTER(_P)]),
    "msbg_destroy": (_I, [_P]),
    "msbg_dev_alloc": (_I, [_I64, C.POINTER(_P)]),
    "msbg_dev_free": (_I, [_P]),
    "msbg_dev_upload": (_I, [_P, _P, _I64]),
    "msbg_dev_download": (_I, [_P, _P, _I64]),
    "msbg_get_dims": (_I, [_P, _P, _P]),
    "msbg_set_arith": (_I, [_P, _I]),
    "msbg_synchronize": (_I, [_P]),
    "msbg_points_upload": (_I, [_P, _P, _I64, _P]),
    "msbg_points_download": (_I, [_P, _P, _I64]),
    "msbg_activate_from_points": (_I, [_P, _F, _P, _P]),
    "msbg_levels_from_points": (_I, [_P, _P, _I, _P]),
    "msbg_regularize_refinement_map": (_I, [_P, _P, _P]),
    "msbg_set_refinement_map": (_I, [_P, _P, _P]),
    "msbg_get_block_info": (_I, [_P, _P, _P]),
    "msbg_set_exists": (_I, [_P, _P]),
    "msbg_get_fine_coarse_blocks": (_I, [_P, _I, _P, _I64, _P]),
    "msbg_channel_define": (_I, [_P, _I, _I, _I, _P, _D, _D, _D]),
    "msbg_channel_alloc_demo": (_I, [_P, _I, _I, _P, _I64, _P]),
    "msbg_channel_storage": (_I, [_P, _I, C.POINTER(_I)]),
    "msbg_channel_get_pagemap": (_I, [_P, _I, _I, _P, _P]),
    "msbg_channel_download": (_I, [_P, _I, _I, _P, _I64]),
    "msbg_channel_upload": (_I, [_P, _I, _I, _P, _I64]),
    "msbg_download_block": (_I, [_P, _I, _I, _I, _P]),
    "msbg_upload_block": (_I, [_P, _I, _I, _I, _P]),
    "msbg_get_halo_table": (_I, [_P, _I, _I, _P, _I64]),
    "msbg_splat_points": (_I, [_P, _I, _F, _F]),
    "msbg_finalize_density": (_I, [_P, _I, _F, _F]),
    "msbg_halo_fill": (_I, [_P, _I, _I, _I, C.c_uint, _P]),
    "msbg_prepare_ghost_blocks": (_I, [_P, _I]),
    "msbg_apply_channel_pde_fast": (_I, [_P, _I, _I, _I, _P, _I64, _I, _I, _F, _F, _F, _I, _F,
                                         _F, _P]),
    "msbg_apply_channel_pde": (_I, [_P, _I, _I, _I, _P, _I64, _I, _I, _F, _F, _F, _I, _F, _F,
                                    _F, _F, _F, _P]),
    "msbg_last_sweep_timing": (_I, [_P, _P, _P, _P]),
    "msbg_apply_channel_pde_level": (_I, [_P, _I, _I, C.c_uint, _P, _I64, _I, _F, _I, _P]),
    "msbg_scene_spheres": (_I, [_U64, _I, _F, _F, _F, _F, _I64, _F, _F, _P]),
    "msbg_copy_bandwidth": (_I, [_I64, _I, _P]),
    "msbg_channel_copy": (_I, [_P, _I, _I, _I, _I]),
    "msbg_batch_create": (_I, [C.POINTER(_P)]),
    "msbg_batch_destroy": (_I, [_P]),
    "msbg_batch_add": (_I, [_P, _P, _I, _F, _I]),
    "msbg_batch_autotune": (_I, [_P, _I]),
    "msbg_batch_get_tiles": (_I, [_P, _P, _P, _P, _I, _P]),
    "msbg_batch_run": (_I, [_P, _I, _I, _P]),
    "msbg_batch_get_modes": (_I, [_P, _P, _I, _P]),
    "msbg_batch_set_wait_limit": (_I, [_P, C.c_uint]),
    "msbg_set_tile": (_I, [_P, _I]),
    "msbg_channel_diff": (_I, [_P, _I, _I, _I, _P, _P]),
    "msbg_interpolate": (_I, [_P, _I, _I, _I, _P, _I64, _P, _P]),
    "msbg_render": (_I, [_P, _I, _I, _I, _P, _P, _F, _P, _P, _P, _P]),
    "msbg_render_counts": (_I, [_P, _I, _I, _I, _P, _P, _F, _P, _P, _P, _P]),
    "msbg_scene_objects": (_I, [_U64, _I, _F, _F, _F, _F, _I64, _F, _F, _F, _F, _P]),
    "msbg_points_generate": (_I, [_P, _U64, _I, _F, _F, _F, _F, _I64, _F, _F, _F, _F, _I, _P]),
    "msbg_unet_create": (_I, [_U64, C.POINTER(_P)]),
    "msbg_unet_destroy": (_I, [_P]),
    "msbg_unet_num_params": (_I, [_P, _P]),
    "msbg_unet_get_params": (_I, [_P, _P, _I64]),
    "msbg_unet_set_params": (_I, [_P, _P, _I64]),
    "msbg_unet_get_grads": (_I, [_P, _P, _I64]),
    "msbg_unet_forward": (_I, [_P, _P, _I, _P]),
    "msbg_unet_backward": (_I, [_P, _P, _P, _I]),
    "msbg_unet_train_step": (_I, [_P, _P, _P, _I, _F, _P]),
    "msbg_unet_train_step_rows": (_I, [_P, _P, _P, _P, _I, _F, _P]),
    "msbg_unet_forward_rows": (_I, [_P, _P, _P, _I, _P]),
    "msbg_unet_take_loss": (_I, [_P, C.POINTER(_D), C.POINTER(_I64)]),
    "msbg_unet_gather": (_I, [_P, _I, _P, _I, _P]),
    "msbg_unet_scatter": (_I, [_P, _I, _P, _I, _P]),
    "msbg_unet_extract": (_I, [_P, _I, _P, _I, _P]),
    "msbg_warm_start": (_I, [_P, _I, _P, _I, _P]),
    "msbg_cnn_conv3d": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _I, _P]),
    "msbg_cnn_conv3d_noalloc": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _I, _P]),
    "msbg_cnn_conv3d_bwd_data": (_I, [_P, _I, _I, _I, _I, _I, _P, _P]),
    "msbg_cnn_conv3d_bwd_weights": (_I, [_P, _P, _I, _I, _I, _I, _I, _P, _P]),
    "msbg_ps_create": (_I, [_P, _P, _P, C.POINTER(_P)]),
    "msbg_ps_destroy": (_I, [_P]),
    "msbg_ps_get_levels": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "msbg_ps_level_size": (_I, [_P, _I, C.POINTER(_I64)]),
    "msbg_ps_upload": (_I, [_P, _I, _I, _P]),
    "msbg_ps_download": (_I, [_P, _I, _I, _P]),
    "msbg_ps_download_flags": (_I, [_P, _I, _P, _P]),
    "msbg_ps_fill": (_I, [_P, _I, _I, _F]),
    "msbg_ps_relax": (_I, [_P, _I, _I, _I, _I, _I, _I, _I]),
    "msbg_ps_multiply": (_I, [_P, C.c_uint, _I, _I, _I, _F, _I, C.POINTER(_D)]),
    "msbg_ps_restrict": (_I, [_P, _I, _I, _I]),
    "msbg_ps_prolongate": (_I, [_P, _I, _I, _I]),
    "msbg_ps_vcycle": (_I, [_P, _I, _I, _I]),
    "msbg_ps_synchronize": (_I, [_P]),
    "msbg_ps_last_timing": (_I, [_P, C.POINTER(_D)]),
    "msbg_ps_relaxed_cells": (_I, [_P, _I, C.POINTER(_I64)]),
}


class PsDesc(C.Structure):
    """msbg_ps_desc"""
    _fields_ = [("nbx", _I), ("nby", _I), ("nbz", _I), ("bsx0", _I), ("dx0", _D),
                ("domainBc", (_I * 2) * 3), ("omega", _F), ("blockIter", _I),
                ("denseLevels", _I)]

_lib = None


def lib():
    """Load libmsbg_gpu.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built; run `python -m msbg_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().msbg_last_error()
        raise MsbgError(rc, msg.decode() if msg else "")
    return rc
