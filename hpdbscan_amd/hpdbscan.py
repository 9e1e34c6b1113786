"""Drop-in for the reference's SWIG module ``hpdbscan`` (swig/hpdbscan.i).

    import hpdbscan_amd.hpdbscan as hpdbscan      # or: from hpdbscan_amd import HPDBSCAN
    labels = hpdbscan.HPDBSCAN(epsilon=0.3, min_points=4).cluster(data)

Behaviour mirrored from the binding:

* ``HPDBSCAN(epsilon, min_points)`` -- ``float`` / ``size_t`` arguments
  (hpdbscan.i:31); ``epsilon <= 0`` raises ``ValueError("epsilon needs to be
  positive")`` (hpdbscan.h:295-297 surfaced through SWIG's std::invalid_argument
  mapping); a negative ``min_points`` raises ``OverflowError`` and a
  non-integer one ``TypeError``, as SWIG's size_t conversion does.
* ``cluster(data)`` -- the array is converted to a C-contiguous float64 2-D
  array (numpy.i ``IN_ARRAY2`` with conversion, hpdbscan.i:27), so the
  clustering runs in double precision exactly like the reference's
  ``cluster<double>`` (hpdbscan.h:484).  A non-2-D input raises ``TypeError``.
  The result is a tuple of ints (the ``ClusterVector`` of hpdbscan.i:20).
* ``cluster(path, dataset[, threads])`` -- the HDF5 entry point
  (hpdbscan.h:300-356): the 2-D dataset is clustered in its stored element
  type (no float64 conversion, as in the reference, where only the array
  overload goes through numpy.i); ``threads`` must be an int and is
  otherwise unused.  A missing file or dataset, or a dataset that is not a
  matrix, raises ``ValueError`` with the reference's message (io.h:49-66).

Extensions (not in the reference): ``cluster(data, as_array=True)`` returns
an ``int64`` ndarray instead of a tuple, and ``cluster_native(data)`` keeps the
input dtype (float32 runs the float32 path the C++ API uses).
"""
from __future__ import annotations

import math
import numbers

import numpy as np

from . import _lib, h5

__all__ = ["HPDBSCAN"]

_FLT_MAX = 3.4028234663852886e38
_SIZE_MAX = (1 << 64) - 1


def _as_float_arg(value, what: str) -> float:
    if isinstance(value, bool) or not isinstance(value, numbers.Real):
        raise TypeError(f"in method '{what}', argument of type 'float' expected")
    v = float(value)
    if math.isfinite(v) and abs(v) > _FLT_MAX:
        raise OverflowError(f"in method '{what}', argument out of range of 'float'")
    return v


def _as_size_t_arg(value, what: str) -> int:
    if isinstance(value, bool) or not isinstance(value, numbers.Integral):
        raise TypeError(f"in method '{what}', argument of type 'size_t' expected")
    v = int(value)
    if v < 0 or v > _SIZE_MAX:
        raise OverflowError(f"in method '{what}', argument out of range of 'size_t'")
    return v


class HPDBSCAN:
    """HPDBSCAN(epsilon, min_points) -- see the module docstring."""

    def __init__(self, epsilon, min_points, *, device: int | None = None):
        eps = _as_float_arg(epsilon, "new_HPDBSCAN")
        mp = _as_size_t_arg(min_points, "new_HPDBSCAN")
        if not eps > 0.0:
            raise ValueError("epsilon needs to be positive")
        self._epsilon = float(np.float32(eps))
        self._min_points = mp
        self._device = device

    @property
    def epsilon(self) -> float:
        return self._epsilon

    @property
    def min_points(self) -> int:
        return self._min_points

    def _ctx(self) -> _lib.Context:
        return _lib.default_context(self._device)

    def cluster(self, *args, as_array: bool = False):
        """cluster(data) or cluster(path, dataset[, threads])."""
        if len(args) == 1 and not isinstance(args[0], (str, bytes)):
            try:
                a = np.ascontiguousarray(np.asarray(args[0], dtype=np.float64))
            except (TypeError, ValueError) as e:
                raise TypeError(f"array of type 'double' expected: {e}") from None
            if a.ndim != 2:
                raise TypeError(
                    f"Array must have 2 dimensions.  Given array has {a.ndim} dimensions")
            labels = self._ctx().cluster(a, self._epsilon, self._min_points)
            return labels if as_array else tuple(labels.tolist())
        if len(args) in (2, 3) and all(isinstance(x, str) for x in args[:2]):
            if len(args) == 3 and (isinstance(args[2], bool) or not isinstance(args[2], numbers.Integral)):
                raise TypeError("in method 'HPDBSCAN_cluster', argument 4 of type 'int'")
            # the file is read before the device is touched, as in the
            # reference, so file errors come first
            h5.matrix_info(args[0], args[1])
            labels = h5.cluster_file(args[0], args[1], self._epsilon, self._min_points, self._ctx())
            return labels if as_array else tuple(labels.tolist())
        raise TypeError("Wrong number or type of arguments for overloaded function 'HPDBSCAN_cluster'")

    def cluster_native(self, data) -> np.ndarray:
        """Cluster without converting the dtype (float32 stays float32)."""
        a = np.ascontiguousarray(data)
        return self._ctx().cluster(a, self._epsilon, self._min_points)
