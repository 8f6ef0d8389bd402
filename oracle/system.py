"""Molecule + basis as plain arrays, in the reference's basis-function order.

Order (reference `qmtools/qmtools.py:150-171`): atom, then the element's shells
in file order, then m in the order of `global_m_values` (`qmtools.py:52`):
L=0: (0); L=1: (0, +1, -1); L=2: (0, +1, -1, +2, -2).
Coordinates are bohr: xyz[Angstrom] * 1.8897259886 (`qmtools.py:18,129`).
"""
from dataclasses import dataclass, field

import numpy

ANG2BOR = 1.8897259886
M_ORDER = {0: (0,), 1: (0, 1, -1), 2: (0, 1, -1, 2, -2)}


@dataclass
class System:
    types: numpy.ndarray            # (natoms,) int
    coords: numpy.ndarray           # (natoms, 3) bohr, float64
    shells: list = field(default_factory=list)   # [(atom, L, alphas, coeffs)]
    dm: numpy.ndarray = None        # alpha-spin density matrix (nbf, nbf)

    @property
    def nbf(self):
        return sum(2 * L + 1 for (_, L, _, _) in self.shells)

    @property
    def qtot(self):
        return int(numpy.sum(self.types))


def make_system(types, coords_bohr, dm, basis):
    """basis: {Z: [(L, alphas, coeffs)]} as returned by `basisfile.read_basis`."""
    types = numpy.asarray(types, dtype=numpy.int64)
    coords = numpy.asarray(coords_bohr, dtype=numpy.float64).reshape(-1, 3)
    shells = []
    for a, z in enumerate(types):
        for (L, al, co) in basis[int(z)]:
            shells.append((a, L, numpy.asarray(al, float), numpy.asarray(co, float)))
    s = System(types=types, coords=coords, shells=shells,
               dm=None if dm is None else numpy.asarray(dm, dtype=numpy.float64))
    if s.dm is not None and s.dm.shape != (s.nbf, s.nbf):
        raise ValueError("density matrix does not match the basis")
    return s


def lower_symmetric(dm):
    """The matrix the reference kernels actually use.

    Both the density and the GTO kernels read only P[p][q] with q <= p and
    double the off-diagonal terms (`kernel_density.cu:74-80,138`,
    `kernel_hartree.cu:352-354,836`), i.e. they use tril(P) mirrored.
    """
    lo = numpy.tril(dm)
    return lo + numpy.tril(dm, -1).T
