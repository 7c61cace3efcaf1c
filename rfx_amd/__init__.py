"""rfx_amd -- MI355X-native random-forest proximity path.

Leaf assignment, same-leaf co-occurrence counts (ALL / OOB_ASYM / OOB_UPPER),
finishprox, proximity factors and 3-D MDS as hand-written gfx950 HIP kernels
behind the C ABI in include/rfx_prox.h (librfx.so).  The forest input is grown
on the host exactly as the reference CPU grower would.
"""
from . import _lib
from ._lib import (ALGO_AUTO, ALGO_BUCKET, ALGO_DENSE, LAYOUT_FULL, LAYOUT_UPPER, PROX_ALL, PROX_OOB_ASYM,
                   PROX_OOB_UPPER, lib)
from .forest import Forest, default_mtry, make_dataset

__all__ = [
    "ALGO_AUTO", "ALGO_BUCKET", "ALGO_DENSE", "LAYOUT_FULL", "LAYOUT_UPPER", "PROX_ALL", "PROX_OOB_ASYM",
    "PROX_OOB_UPPER", "Forest", "default_mtry", "lib", "make_dataset",
]
