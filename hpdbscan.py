"""``import hpdbscan`` -- the reference's Python module name (swig/hpdbscan.i:1,
README.md:69-72) over the MI355X library.

    import hpdbscan
    labels = hpdbscan.HPDBSCAN(epsilon=0.3, min_points=4).cluster(data)
    labels = hpdbscan.HPDBSCAN(0.3, 4).cluster("file.h5", "DATA")

``HPDBSCAN`` is :class:`hpdbscan_amd.hpdbscan.HPDBSCAN` (float64 conversion of
array input, a tuple of ints as the result, the reference's exceptions).  The
SWIG module's other name, the ``ClusterVector`` proxy of
``std::vector<ptrdiff_t>`` (hpdbscan.i:20), is never constructed by callers:
``cluster`` returns a tuple, as the SWIG typemap does.
"""
from hpdbscan_amd.hpdbscan import HPDBSCAN

__all__ = ["HPDBSCAN"]
