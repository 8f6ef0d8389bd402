"""Basis-set tables for the grid hot path.

`BasisSet(filename)` reads the same binary basis file as the reference
(`qmtools/qmtools.py:33-108`, file format `:39-99`):

    int32 natm
    natm x { int32 Z, int32 nshells,
             nshells x { int32 L, int32 nc, f64[nc] alpha, f64[nc] coeff } }

The reference keeps slot-strided FP32 tables (shell s of the global shell list
occupies `[s*MAXAOC, s*MAXAOC + nc)`).  Those tables are kept with the same
names and meaning so that code written against the reference keeps working,
and FP64 copies (`alphas64`, `coeffs64`) are kept beside them: the FP64 grid
path never sees the FP32 rounding of the basis parameters.

The per-element shell lists (`element_shells`) are what the device tables are
built from (see `Molecule.device_tables`).
"""
from __future__ import annotations

import os
import struct

import numpy

MAXAOS = 15   # max shells per element (reference qmtools.py:20, kernel_common.h:9)
MAXAOC = 20   # primitive slots per shell (reference qmtools.py:21, kernel_common.h:10)
MAXZ = 100

# m order of the 2L+1 real solid harmonics of a shell
# (reference qmtools.py:52: rows L=0,1,2 of global_m_values).
M_ORDER = {0: (0,), 1: (0, 1, -1), 2: (0, 1, -1, 2, -2)}


class BasisSet:
    """Contracted real-solid-harmonic GTO basis (reference `qmtools.py:29-108`)."""

    def __init__(self, filename):
        filename = os.fspath(filename)
        with open(filename, "rb") as fh:
            raw = fh.read()
        self.filename = filename
        self._parse(raw)

    # -- parsing --------------------------------------------------------
    def _parse(self, raw: bytes):
        pos = 0

        def take_i32():
            nonlocal pos
            if pos + 4 > len(raw):
                raise ValueError("basis file truncated")
            (v,) = struct.unpack_from("<i", raw, pos)
            pos += 4
            return v

        def take_f64(n):
            nonlocal pos
            if pos + 8 * n > len(raw):
                raise ValueError("basis file truncated")
            v = numpy.frombuffer(raw, dtype="<f8", count=n, offset=pos).copy()
            pos += 8 * n
            return v

        natm = take_i32()
        nslot = MAXAOS * MAXZ
        self.atomOffset = numpy.full(MAXZ, -1, dtype=numpy.int32)
        self.nshells = numpy.zeros(nslot, dtype=numpy.int32)
        self.shellOffset = numpy.zeros(nslot, dtype=numpy.int32)
        self.Ls = numpy.zeros(nslot, dtype=numpy.int32)
        self.alphas64 = numpy.zeros(nslot * MAXAOC, dtype=numpy.float64)
        self.coeffs64 = numpy.zeros(nslot * MAXAOC, dtype=numpy.float64)
        self.global_m_values = numpy.array(
            [[0, 0, 0, 0, 0], [0, 1, -1, 0, 0], [0, 1, -1, 2, -2]], dtype=numpy.int32)
        self.element_shells = {}

        shell = 0
        maxsh = 0
        ncmax = 0
        for _ in range(natm):
            z = take_i32()
            nsh = take_i32()
            if not (0 < z < MAXZ) or nsh < 0 or nsh > MAXAOS:
                raise ValueError(f"bad basis record Z={z} nshells={nsh}")
            self.atomOffset[z] = shell
            self.nshells[z] = nsh
            maxsh = max(maxsh, nsh)
            shells = []
            for _ in range(nsh):
                L = take_i32()
                nc = take_i32()
                if nc < 0 or nc > MAXAOC:
                    raise ValueError(f"shell with {nc} primitives exceeds {MAXAOC}")
                a = take_f64(nc)
                c = take_f64(nc)
                base = shell * MAXAOC
                self.Ls[shell] = L
                self.shellOffset[shell] = base
                self.alphas64[base:base + nc] = a
                self.coeffs64[base:base + nc] = c
                shells.append((L, a, c))
                ncmax = max(ncmax, nc)
                shell += 1
            self.element_shells[z] = shells
        self.nparams = shell * MAXAOC
        self.maxorbs = maxsh
        self.ncmax = ncmax
        # FP32 views with the reference's names (qmtools.py:49-50)
        self.alphas = self.alphas64.astype(numpy.float32)
        self.coeffs = self.coeffs64.astype(numpy.float32)

    # -- queries ----------------------------------------------------------
    def has_element(self, z: int) -> bool:
        return 0 < z < MAXZ and self.atomOffset[z] >= 0

    def nbf_of(self, z: int, lmax: int = 2) -> int:
        """Number of basis functions of element z (shells with L<=lmax)."""
        return sum(2 * L + 1 for (L, _, _) in self.element_shells[z] if L <= lmax)

    def __repr__(self):
        return (f"BasisSet({self.filename!r}: {len(self.element_shells)} elements, "
                f"maxorbs={self.maxorbs}, ncmax={self.ncmax}, nparams={self.nparams})")


def default_basis_path() -> str:
    """Path of the packaged cc-pVDZ table (same bytes as the reference's data/cc-pvdz.bin)."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "cc-pvdz.bin")
