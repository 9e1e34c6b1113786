"""MI355X-native HPDBSCAN: the reference's single-node clustering path as HIP
kernels behind a C-ABI (include/hpdbscan_capi.h, libhpdbscan_amd.so).

Python entry points:

* :class:`HPDBSCAN` -- drop-in for the reference's SWIG module class.
* :func:`cluster` -- functional form that keeps the input dtype.
* :class:`Context` -- explicit device context (stream + workspace), with the
  device-resident entry point ``cluster_device`` for data already in HBM.
"""
from __future__ import annotations

import numpy as np

from ._lib import (BORDER_MIN_CLUSTER, BORDER_ONE_THREAD, Context, HPDError,  # noqa: F401
                   Server, concat_jobs, default_context, load)
from .hpdbscan import HPDBSCAN  # noqa: F401

__version__ = "0.1.0"


def cluster(data, eps: float, min_points: int, device: int | None = None) -> np.ndarray:
    """Reference labels (input order) of a 2-D array of any supported dtype."""
    if not eps > 0:
        raise ValueError("epsilon needs to be positive")
    a = np.ascontiguousarray(data)
    return default_context(device).cluster(a, float(np.float32(eps)), int(min_points))


def version() -> str:
    return load().hpd_version().decode()
