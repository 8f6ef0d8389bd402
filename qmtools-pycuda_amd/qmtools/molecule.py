"""Molecule: nuclei, density matrix and the basis-function table.

Mirrors the reference `Molecule` (`qmtools/qmtools.py:110-191`):

* xyz rows are (Z, x, y, z) in Angstrom and are stored in bohr
  (`coords = xyz * ANG2BOR`, reference `:128-131`);
* `qtot` is the total nuclear charge (`:134`);
* `dm` is the alpha-spin density matrix in the basis-function order of
  `ALMOs` (atom -> shell -> m, `:150-171`); the grid kernels multiply by 2
  for the closed shell, exactly as the reference kernels do;
* `ALMOs[c] = (atom, L, m, shellOffset)` as int16.

Differences that are deliberate: coordinates and the density matrix are kept
in FP64 (the reference casts both to FP32, `:131,145`), shells with L > 2 are
refused with a ValueError (the reference's `global_m_values` only has rows for
L <= 2 and would fail with an IndexError), and a density matrix whose size
does not match the basis is refused.
"""
from __future__ import annotations

import numpy

from .basis import M_ORDER, BasisSet

ANG2BOR = 1.8897259886


class Molecule:
    """A molecule in bohr with its density matrix (reference `qmtools.py:110-191`)."""

    def __init__(self, filexyz, filedm, basisset: BasisSet):
        self.basisset = basisset
        if isinstance(filexyz, (list, tuple, numpy.ndarray)):
            m = numpy.array(filexyz, dtype=numpy.float64)
        else:
            m = numpy.loadtxt(filexyz, dtype=numpy.float64)
        if m.ndim == 1:
            m = m[None, :]
        if m.ndim != 2 or m.shape[1] != 4:
            raise ValueError("xyz data must have rows (Z, x, y, z)")
        self.types = numpy.array(numpy.rint(m[:, 0]), dtype=numpy.int32)
        self.coords = numpy.ascontiguousarray(m[:, 1:] * ANG2BOR, dtype=numpy.float64)
        self.natoms = numpy.int32(self.types.shape[0])
        self.qtot = numpy.int32(int(numpy.sum(self.types)))
        for z in self.types:
            if not basisset.has_element(int(z)):
                raise ValueError(f"element Z={int(z)} is not in the basis set")

        if isinstance(filedm, numpy.ndarray):
            dm = filedm
        else:
            dm = numpy.load(filedm, allow_pickle=False)
        self.dm = numpy.ascontiguousarray(dm, dtype=numpy.float64)
        if self.dm.ndim != 2 or self.dm.shape[0] != self.dm.shape[1]:
            raise ValueError("density matrix must be square")
        self.norbs = numpy.int32(self.dm.shape[0])
        self._build_tables()
        self._dev = None

    # -- basis-function tables --------------------------------------------
    def _build_tables(self):
        bs = self.basisset
        almos = []
        sh_atom, sh_L, sh_nprim, sh_poff, sh_bf = [], [], [], [], []
        alphas, coeffs = [], []
        nbf = 0
        for a in range(int(self.natoms)):
            z = int(self.types[a])
            first = int(bs.atomOffset[z])
            for s, (L, al, co) in enumerate(bs.element_shells[z]):
                if L > 2:
                    raise ValueError(f"Z={z} has an L={L} shell; only L<=2 is supported")
                sh_atom.append(a)
                sh_L.append(L)
                sh_nprim.append(len(al))
                sh_poff.append(len(alphas))
                sh_bf.append(nbf)
                alphas.extend(al.tolist())
                coeffs.extend(co.tolist())
                for m in M_ORDER[L]:
                    almos.append((a, L, m, int(bs.shellOffset[first + s])))
                nbf += 2 * L + 1
        if nbf != int(self.norbs):
            raise ValueError(f"density matrix is {int(self.norbs)}x{int(self.norbs)} "
                             f"but the basis has {nbf} functions")
        self.ALMOs = numpy.array(almos, dtype=numpy.int16).reshape(-1, 4)
        self.shell_atom = numpy.array(sh_atom, dtype=numpy.int32)
        self.shell_L = numpy.array(sh_L, dtype=numpy.int32)
        self.shell_nprim = numpy.array(sh_nprim, dtype=numpy.int32)
        self.shell_poff = numpy.array(sh_poff, dtype=numpy.int32)
        self.shell_bf = numpy.array(sh_bf, dtype=numpy.int32)
        self.prim_alpha = numpy.array(alphas, dtype=numpy.float64)
        self.prim_coeff = numpy.array(coeffs, dtype=numpy.float64)
        self.nshell = len(sh_atom)

    @property
    def nbf(self) -> int:
        return int(self.norbs)

    def Translate(self, dx=0, dy=0, dz=0):
        """Shift all nuclei by (dx, dy, dz) bohr (reference `qmtools.py:182-188`)."""
        self.coords[:, 0] += dx
        self.coords[:, 1] += dy
        self.coords[:, 2] += dz
        if self._dev is not None:
            self._dev.update_coords(self.coords)

    def xyz_angstrom(self):
        return numpy.column_stack([self.types.astype(numpy.float64), self.coords / ANG2BOR])

    def __repr__(self):
        return f"Molecule(natoms={int(self.natoms)}, nbf={self.nbf}, qtot={int(self.qtot)})"
