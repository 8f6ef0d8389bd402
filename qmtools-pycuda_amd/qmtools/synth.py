"""Seeded synthetic molecules and density matrices for benchmarks and tests.

Molecules: heavy atoms drawn from {C, N, O} grow as a random tree (each new
heavy atom bonds to a random existing heavy atom with free valence, at a
distance U(1.3, 1.5) Angstrom in a uniformly random direction); hydrogens are
then attached at 1.09 Angstrom to heavy atoms with free valence.  Placements
closer than 1.2 Angstrom to any atom other than the bonding partner are
rejected.  The heavy-atom types are adjusted so that the electron count is
even.

Density matrix: P = 2 C C^T with C the Q factor of a QR decomposition of an
N(0,1) [nbf x n_occ] matrix, n_occ = sum(Z)/2.  The drop-in API takes the
alpha-spin matrix (the grid kernels multiply by 2), so `Molecule` receives
P/2 = C C^T.
"""
from __future__ import annotations

import numpy

from .basis import BasisSet, default_basis_path
from .grid import ANG2BOR, Grid, make_float3
from .molecule import Molecule

_VALENCE = {6: 4, 7: 3, 8: 2}


def _unit_vector(rng):
    v = rng.normal(size=3)
    n = numpy.linalg.norm(v)
    while n < 1e-12:
        v = rng.normal(size=3)
        n = numpy.linalg.norm(v)
    return v / n


def random_geometry(n_atoms: int, seed: int, heavy_fraction: float = 0.57):
    """Return an (n_atoms, 4) array of (Z, x, y, z) in Angstrom."""
    if n_atoms < 1:
        raise ValueError("need at least one atom")
    rng = numpy.random.default_rng(seed)
    n_heavy = max(1, int(round(n_atoms * heavy_fraction)))
    n_h = n_atoms - n_heavy
    types = list(rng.choice([6, 7, 8], size=n_heavy, p=[0.6, 0.2, 0.2]))
    # even electron count: flip the first C<->N (changes the parity by one)
    if (sum(types) + n_h) % 2:
        types[0] = 7 if types[0] == 6 else (6 if types[0] == 7 else 7)
    # enough valence for the tree bonds plus all hydrogens?
    def capacity(ts):
        return sum(_VALENCE[t] for t in ts) - 2 * (len(ts) - 1)
    k = 0
    while capacity(types) < n_h and k < len(types):
        if types[k] != 6:
            # swap an O/N for C keeping parity: O(8)->C(6) keeps parity, N needs a partner
            if types[k] == 8:
                types[k] = 6
        k += 1
    if capacity(types) < n_h:
        raise ValueError("cannot place that many hydrogens")

    pos = [numpy.zeros(3)]
    degree = [0]
    for i in range(1, n_heavy):
        for _attempt in range(10000):
            free = [j for j in range(len(pos)) if degree[j] < _VALENCE[types[j]] - 1 or
                    (degree[j] < _VALENCE[types[j]] and i == n_heavy - 1)]
            if not free:
                free = [j for j in range(len(pos)) if degree[j] < _VALENCE[types[j]]]
            parent = int(rng.choice(free))
            p = pos[parent] + rng.uniform(1.3, 1.5) * _unit_vector(rng)
            if all(numpy.linalg.norm(p - q) >= 1.2 for j, q in enumerate(pos) if j != parent):
                pos.append(p)
                degree.append(1)
                degree[parent] += 1
                break
        else:  # pragma: no cover - practically unreachable
            raise RuntimeError("could not grow the heavy-atom tree")

    hpos = []
    for _ in range(n_h):
        for _attempt in range(10000):
            free = [j for j in range(n_heavy) if degree[j] < _VALENCE[types[j]]]
            parent = int(rng.choice(free))
            p = pos[parent] + 1.09 * _unit_vector(rng)
            others = [q for j, q in enumerate(pos) if j != parent] + hpos
            if all(numpy.linalg.norm(p - q) >= 1.2 for q in others):
                hpos.append(p)
                degree[parent] += 1
                break
        else:  # pragma: no cover
            raise RuntimeError("could not place hydrogens")

    zs = numpy.array(list(types) + [1] * n_h, dtype=numpy.float64)
    xyz = numpy.vstack(pos + hpos)
    xyz -= xyz.mean(axis=0)
    return numpy.column_stack([zs, xyz])


def random_density_matrix(nbf: int, n_occ: int, seed: int):
    """Total density matrix P = 2 C C^T with orthonormal columns C (nbf x n_occ)."""
    rng = numpy.random.default_rng(seed)
    a = rng.standard_normal((nbf, n_occ))
    q, _ = numpy.linalg.qr(a)
    return 2.0 * (q @ q.T)


def synthetic_molecule(n_atoms: int, seed: int, basis: BasisSet | None = None):
    """Seeded molecule with P = 2CC^T; the Molecule holds the alpha-spin matrix P/2."""
    if basis is None:
        basis = BasisSet(default_basis_path())
    xyz = random_geometry(n_atoms, seed)
    zsum = int(round(xyz[:, 0].sum()))
    nbf = sum(basis.nbf_of(int(z)) for z in xyz[:, 0])
    p = random_density_matrix(nbf, zsum // 2, seed + 7919)
    return Molecule(xyz, 0.5 * p, basis)


def centred_grid(mol: Molecule, n: int, step_bohr: float = 0.1, pad_ang: float = 3.0,
                 dtype=numpy.float64):
    """Cubic n^3 grid centred on the molecule.

    The step is `step_bohr` unless the molecule plus `pad_ang` padding on each
    side does not fit, in which case the step grows so that it does; the step
    actually used is `grid.step` (bohr).
    """
    xyz = mol.coords
    lo = xyz.min(axis=0) - pad_ang * ANG2BOR
    hi = xyz.max(axis=0) + pad_ang * ANG2BOR
    need = float((hi - lo).max()) / n
    h = max(step_bohr, need)
    centre = 0.5 * (lo + hi)
    g = Grid((1, n, n, n), h / ANG2BOR, dtype=dtype)
    g.step = numpy.float64(h)
    g.origin = make_float3(*(centre - 0.5 * n * h))
    return g


def uniform_points(grid: Grid, npoints: int, seed: int):
    """npoints positions (bohr) drawn uniformly in the grid's box."""
    rng = numpy.random.default_rng(seed)
    nx, ny, nz, h, o = grid.geometry()
    lo = numpy.array(o)
    ext = numpy.array([nx, ny, nz], dtype=numpy.float64) * h
    return lo + rng.uniform(0.0, 1.0, size=(npoints, 3)) * ext


def molecule_12_grid_96(basis: BasisSet | None = None):
    """Seeded 12-atom molecule (about 130 basis functions) on a 96^3 grid, 3 Angstrom padding:
    the CPU-sized density + 2000-sweep Jacobi workload."""
    mol = synthetic_molecule(12, 1, basis)
    return mol, centred_grid(mol, 96, step_bohr=0.1, pad_ang=3.0)


def molecule_60_grid_256(basis: BasisSet | None = None, npoints: int = 10000):
    """Seeded 60-atom molecule (606 basis functions) on a 256^3 grid, 3 Angstrom padding, and
    `npoints` uniformly drawn points for the GTO potential: the full-size density + Hartree workload."""
    mol = synthetic_molecule(60, 2, basis)
    grid = centred_grid(mol, 256, step_bohr=0.1, pad_ang=3.0)
    return mol, grid, uniform_points(grid, npoints, 3)


def dm_probe(dm):
    """A fingerprint of a density matrix: dm times a fixed vector (compares to ~1e-13 across BLAS
    builds, where the QR of the generator may round differently in the last bits)."""
    n = dm.shape[0]
    v = numpy.cos(0.37 * numpy.arange(n)) + 0.5
    return dm @ v


def molecule_batch(count: int, atoms=(5, 80), grid_n=(64, 96, 128, 160, 192), steps=(0.1, 0.15), seed: int = 41,
                   basis: BasisSet | None = None):
    """`count` seeded molecules with atom counts U{atoms}, cubic grids drawn from grid_n and grid steps
    U(steps) bohr (grown where the molecule plus 3 Angstrom of padding does not fit): the batched and
    served workloads (defaults: 5-80 atoms, 64^3-192^3 grids at 0.1-0.15 bohr)."""
    if basis is None:
        basis = BasisSet(default_basis_path())
    rng = numpy.random.default_rng(seed)
    mols, grids = [], []
    for i in range(count):
        na = int(rng.integers(atoms[0], atoms[1] + 1))
        n = int(rng.choice(grid_n))
        st = float(rng.uniform(*steps))
        mol = synthetic_molecule(na, seed * 1000 + i, basis)
        mols.append(mol)
        grids.append(centred_grid(mol, n, step_bohr=st, pad_ang=3.0))
    return mols, grids
