"""Grid reductions (reference `qmtools/cuda/kernel_utils.cu:54-168`, `qmtools.py:2796-2831`).

All act on field 0 of a grid (the kernels index only the spatial block).
Sums use math.fsum, i.e. the correctly rounded FP64 sum.
"""
import math

import numpy


def total(q):
    """sum q (gpu_total, kernel_utils.cu:68-98)."""
    return math.fsum(numpy.asarray(q, dtype=numpy.float64).ravel())


def rescale(q, factor):
    """q * factor (gpu_rescale, kernel_utils.cu:54-65)."""
    return numpy.asarray(q, dtype=numpy.float64) * factor


def absdiff(a, b):
    """sum |a - b| (gpu_qdiff, kernel_utils.cu:101-133)."""
    return math.fsum(numpy.abs(numpy.asarray(a, float) - numpy.asarray(b, float)).ravel())


def reldiff(a, b):
    """100 * sum |a - b| / b (gpu_qdiff_rel kernel_utils.cu:136-168, x100 at qmtools.py:2829)."""
    a = numpy.asarray(a, float)
    b = numpy.asarray(b, float)
    with numpy.errstate(divide="ignore", invalid="ignore"):
        return 100.0 * math.fsum((numpy.abs(a - b) / b).ravel())
