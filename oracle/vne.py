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


# --------------------------------------------------------------------------
# Diffusion / adsorption pseudo-nuclear field.
#
# Reference `qmtools/cuda/kernel_vne.cu:13-128` (one sweep) and
# `qmtools.py:2677-2745` (the loop).  One sweep, for a voxel r with value v:
#
#   out = sum_{26 neighbours n, loop x, y, z = -1..1} DIF * V(r + n) / sqrt(|nx|+|ny|+|nz|)
#       + (1 - DIF * 16.794682450997072) * v          (:84)
#       - v * ADS                                       (:89)
#       + sum_A [d_A < sqrt(3) h] Z_A exp(-d_A)         (:92-106)
#   d_A = | (r + 1/2) h - (R_A - origin) |  (voxel centre, bohr)
#   out = 0 on the outer voxel layer                   (:110)
#
# Neighbour indices wrap periodically (:63-75); only outer-layer voxels see
# a wrapped neighbour and their result is replaced by 0, so the wrap never
# reaches the returned field.  The loop (qmtools.py:2715-2741) runs pairs of
# sweeps res -> out -> res from res = 0; delta = sum|res - out| of the second
# sweep / npts; it stops when delta <= tolerance or, after the pair count
# exceeds maxiters (default 10 npts^(1/3)), with a "not converged" message.
# --------------------------------------------------------------------------

VNE_SELF = 16.794682450997072          # kernel_vne.cu:84 (not the sum of the 26 weights)


def vne_sweep(v, types, coords, h, origin, adsorb, diff):
    """One sweep on a (nz, ny, nx) array (x fastest); returns the new array."""
    v = numpy.asarray(v, dtype=numpy.float64)
    nz, ny, nx = v.shape
    out = numpy.zeros_like(v)
    for x in (-1, 0, 1):
        for y in (-1, 0, 1):
            for z in (-1, 0, 1):
                d = abs(x) + abs(y) + abs(z)
                if d < 1:
                    continue
                nb = numpy.roll(v, shift=(-z, -y, -x), axis=(0, 1, 2))   # nb[r] = v[r + (x, y, z)]
                out += diff * nb / numpy.sqrt(d)
    out += (1.0 - diff * VNE_SELF) * v
    out -= v * adsorb
    kz, ky, kx = numpy.meshgrid(numpy.arange(nz), numpy.arange(ny), numpy.arange(nx), indexing="ij")
    cx, cy, cz = kx * h + 0.5 * h, ky * h + 0.5 * h, kz * h + 0.5 * h
    rcut = numpy.sqrt(3.0) * h
    for zt, r in zip(numpy.asarray(types, dtype=numpy.float64), numpy.asarray(coords, dtype=numpy.float64)):
        ax, ay, az = r[0] - origin[0], r[1] - origin[1], r[2] - origin[2]
        d = numpy.sqrt((cx - ax) ** 2 + (cy - ay) ** 2 + (cz - az) ** 2)
        out += numpy.where(d < rcut, zt * numpy.exp(-d), 0.0)
    out[0, :, :] = 0.0
    out[-1, :, :] = 0.0
    out[:, 0, :] = 0.0
    out[:, -1, :] = 0.0
    out[:, :, 0] = 0.0
    out[:, :, -1] = 0.0
    return out


def vne_diffusion(types, coords, nx, ny, nz, h, origin, adsorb=0.1, diff=0.01, tolerance=1.0e-9,
                  maxiters=None):
    """Returns (field as (nz, ny, nx), pairs done, last delta, converged)."""
    npts = nx * ny * nz
    if maxiters is None:
        maxiters = 10 * numpy.power(npts, 1.0 / 3)
    res = numpy.zeros((nz, ny, nx))
    delta = 1.0
    niter = 0
    converged = True
    while delta > tolerance:
        out = vne_sweep(res, types, coords, h, origin, adsorb, diff)
        new = vne_sweep(out, types, coords, h, origin, adsorb, diff)
        delta = float(numpy.sum(numpy.abs(out - new))) / npts
        res = new
        niter += 1
        if niter > maxiters:
            converged = False
            break
    return res, niter, delta, converged
