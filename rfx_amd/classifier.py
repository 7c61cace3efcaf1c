"""RandomForestClassifier mirror of the reference pybind class, proximity path.

Keeps the reference's constructor options and the four proximity/MDS methods
of python/randomforest_py.cpp:606-1420 (get_proximity_matrix,
get_lowrank_factors, compute_mds_3d_cpu, compute_mds_from_factors), with the
reference's return dtypes, backed by the HIP path in librfx.
"""
from __future__ import annotations

import numpy as np


def detect_categorical(x: np.ndarray, maxcat: int) -> np.ndarray:
    """Categorical auto-detection of the classifier's fit
    (python/randomforest_py.cpp:771-803): a column whose values are all within
    1e-5 of an integer, with 2..maxcat distinct values, min >= 0 and
    max < maxcat, gets cat = number of distinct values; else 1 (numeric)."""
    x = np.asarray(x, np.float32)
    n, p = x.shape
    cat = np.ones(p, np.int32)
    for m in range(p):
        col = x[:, m]
        rounded = np.round(col)
        all_int = bool(np.all(np.abs(col - rounded) <= np.float32(1e-5)))
        nu = len(np.unique(col))
        if all_int and 1 < nu <= maxcat and col.min() >= 0.0 and col.max() < np.float32(maxcat):
            cat[m] = nu
    return cat
