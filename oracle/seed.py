"""Nuclear seed charge on grid nodes (reference `qmtools/cuda/kernel_utils.cu:5-50`).

Each nucleus at a = R - origin puts Z * prod_d |a_d - i_d h| / h on the 8
nodes i = floor(a/h) + {0,1}^3 (node positions i*h; the lower node gets the
fractional distance, so the weights still sum to one).  Values are stored,
not added (`qfield[gidx] = p`, :49): a later atom overwrites a shared node.
Nodes outside the grid are skipped (the reference has no bounds check).
"""
import numpy


def seed_grid(types, coords, nx, ny, nz, h, origin):
    out = numpy.zeros(nx * ny * nz)
    o = numpy.asarray(origin, dtype=numpy.float64)
    for z, r in zip(types, numpy.asarray(coords, dtype=numpy.float64)):
        a = r - o
        i0 = numpy.floor(a / h).astype(numpy.int64)
        for tx in (0, 1):
            for ty in (0, 1):
                for tz in (0, 1):
                    idx = i0 + numpy.array([tx, ty, tz])
                    if numpy.any(idx < 0) or idx[0] >= nx or idx[1] >= ny or idx[2] >= nz:
                        continue
                    w = float(z)
                    for d in range(3):
                        w *= abs(a[d] - idx[d] * h) / h
                    out[idx[0] + nx * (idx[1] + ny * idx[2])] = w
    return out
