"""MI355X-native HSS compressed direct solver (HSS2D compress / invert / apply path).

The numerical work lives in libhssc.so (HIP kernels for gfx950 behind the C ABI of
include/hssc.h).  This package only binds it and mirrors the reference's MATLAB entry
points (see hss2d.py).  Importing it does not require a GPU; calling a GPU entry point
without a device raises HsscError.
"""
from ._lib import HsscError, cached_memory, lib, release_cached_memory  # noqa: F401
from .hss2d import (  # noqa: F401
    HSS2D_apply_fsym, HSS2D_apply_TI, HSS2D_applyinv_fsym, HSS2D_applyinv_TI,
    HSS2D_build_index_sets, HSS2D_build_tree_fsym, HSS2D_compress_fsym, HSS2D_invert_fsym,
    GridOperator, HSS_tree_parameters, Kernel_Eval, apply_device, applyinv_device, bicgstab,
    applyinv_device_noalloc, dense_inv_device, levHSS, solve_device, solve_exact_device,
)
from .interp import ID, LRID_rand  # noqa: F401
from .export import a_hss, export_mat, info_tree  # noqa: F401
from .hss1d import (  # noqa: F401
    Curve, HSS1D_apply_nsym, HSS1D_applyinv_nsym, HSS1D_compress_batched, HSS1D_compress_nsym,
    HSS1D_invert_nsym, HSS1D_solve, split, star_contour,
)
