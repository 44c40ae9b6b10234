"""MI355X-native sparse-block stencil path of MSBG (splat, halo, 8-colour MCF).

The compute path is libmsbg_gpu.so (hand-written HIP kernels for gfx950
behind a C ABI, include/msbg_gpu.h); this package is its host-side binding.
"""
from . import _lib  # noqa: F401
from ._lib import (ARITH_EXACT, ARITH_FAST, BF16, CH_FLOAT_1, CH_FLOAT_2, CH_FLOAT_3, CH_NULL,  # noqa: F401
                   CH_UINT8_1, CH_UINT16_1, F16, F32,
                   F64, IP_BSPLINE3, IP_LINEAR, OPT_8_COLOR_SCHEME, OPT_BC_COARSE_LEVEL, OPT_BC_NEUMANN, OPT_INTERPOLATE,
                   OPT_SINGLE_LEVEL, PDE_MEAN_CURVATURE, U8, U16, MsbgError)
from .msbg import (MCF_8COLOR, Batch, DeviceArray, DeviceView, MultiresSparseGrid, UNet, cnn_conv3d,  # noqa: F401
                   cnn_conv3d_bwd_data, cnn_conv3d_bwd_weights, copy_bandwidth, getBlockColor8,
                   scene_objects, scene_spheres)
from .pressure import PressureGrid  # noqa: F401
