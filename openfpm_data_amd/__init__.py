"""MI355X-native SparseGridGpu (block-sparse grid) hot path of openfpm_data.

The compute path is libsg_hip.so (hand-written gfx950 HIP kernels behind a C ABI,
include/sparsegrid_gpu.h); this package is its Python host layer.
"""
from . import _capi as capi  # noqa: F401
from ._capi import (  # noqa: F401
    APPLY_EXISTING, APPLY_INTERIOR, CALCULATE_EXISTING_POINTS, EXIST, NN_FULL, NN_STAR,
    NO_CALCULATE_EXISTING_POINTS, PADDING, SgError,
)
from .sparsegrid import SparseGridGpu  # noqa: F401
