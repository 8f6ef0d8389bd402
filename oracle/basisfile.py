"""Independent reader of the binary basis file (reference `qmtools/qmtools.py:33-108`).

Format: int32 natm; then per element int32 Z, int32 nshells; then per shell
int32 L, int32 nc, f64[nc] exponents, f64[nc] contraction coefficients.
Returns {Z: [(L, alphas, coeffs), ...]} in file order.
"""
import struct

import numpy


def read_basis(path):
    with open(path, "rb") as fh:
        raw = fh.read()
    off = 0
    (natm,) = struct.unpack_from("<i", raw, off)
    off += 4
    out = {}
    for _ in range(natm):
        z, nsh = struct.unpack_from("<ii", raw, off)
        off += 8
        shells = []
        for _ in range(nsh):
            L, nc = struct.unpack_from("<ii", raw, off)
            off += 8
            a = numpy.array(struct.unpack_from("<%dd" % nc, raw, off))
            off += 8 * nc
            c = numpy.array(struct.unpack_from("<%dd" % nc, raw, off))
            off += 8 * nc
            shells.append((L, a, c))
        out[z] = shells
    if off != len(raw):
        raise ValueError("trailing bytes in basis file")
    return out
