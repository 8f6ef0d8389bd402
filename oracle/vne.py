"""Classical nuclear potential averaged over voxel sub-points (FP64).

Reference `qmtools/cuda/kernel_vne_classic.cu:13-84`, `qmtools.py:2750-2790`:

  V(i) = [sum_{s in NDIV^3} sum_A Z_A / |r_s - R_A|] / (cnt * h^3)
  r_s  = origin + i h + (s + 1/2) h / NDIV                    (:46-60)
  cnt  = number of (s, A) pairs with r_s != R_A               (:66-74)
"""
import numpy

from .density import grid_points


def vne_classic(types, coords, nx, ny, nz, h, origin, ndiv=4):
    types = numpy.asarray(types, dtype=numpy.float64)
    coords = numpy.asarray(coords, dtype=numpy.float64)
    acc = numpy.zeros(nx * ny * nz)
    cnt = numpy.zeros(nx * ny * nz)
    for sx in range(ndiv):
        for sy in range(ndiv):
            for sz in range(ndiv):
                off = ((sx + 0.5) / ndiv, (sy + 0.5) / ndiv, (sz + 0.5) / ndiv)
                pts = grid_points(nx, ny, nz, h, origin, off)
                for z, r in zip(types, coords):
                    d2 = numpy.sum((pts - r) ** 2, axis=1)
                    nz_ = d2 != 0
                    acc[nz_] += z / numpy.sqrt(d2[nz_])
                    cnt += nz_
    return acc / cnt / h ** 3
