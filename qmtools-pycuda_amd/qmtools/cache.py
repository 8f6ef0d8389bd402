"""Local molecule cache: the offline dataset path of the reference's training scripts.

The reference keeps one file per molecule, `molecule.<CID>.pickle` in its data
folder, holding a dict with keys `CID`, `xyz` (rows Z, x, y, z in Angstrom),
`D-CCSD` (alpha density matrix), `qgrid` (density grid contents) and `gvne`
(diffusion pseudo-V_ne grid contents); `mol` is stored as None
(reference scripts/train-singleshot.py:45-71 writes it, :78-108 reads it
back and rebuilds the Molecule and the two grids on DensityGrid(mol, 0.1, 3.0)).

Here the same fields go to `molecule.<CID>.npz` (NumPy archive, read with
allow_pickle=False, so loading executes nothing from the file).  Files in the
reference's pickle layout can be read with `legacy_pickle=True`; that reader
accepts only the pickle globals NumPy arrays need and refuses every other one.
"""
from __future__ import annotations

import glob
import os

import numpy

from . import _safeload
from .grid import Grid
from .molecule import Molecule

KEYS = ("CID", "xyz", "D-CCSD", "qgrid", "gvne")


def entry_path(directory, cid, ext="npz"):
    return os.path.join(directory, f"molecule.{cid}.{ext}")


def save_entry(directory, cid, xyz, dm, qgrid=None, gvne=None):
    """Write one cache entry; qgrid / gvne are Grids or arrays (their `qube` contents)."""
    os.makedirs(directory, exist_ok=True)
    arrays = {"CID": numpy.array(str(cid)), "xyz": numpy.asarray(xyz, dtype=numpy.float64),
              "D-CCSD": numpy.asarray(dm, dtype=numpy.float64)}
    for key, g in (("qgrid", qgrid), ("gvne", gvne)):
        if g is not None:
            arrays[key] = numpy.asarray(g.qube if isinstance(g, Grid) else g)
    path = entry_path(directory, cid)
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        numpy.savez(fh, **arrays)
    os.replace(tmp, path)
    return path


def read_entry(path):
    """The stored fields of one `.npz` entry as a dict (missing grids are None)."""
    with numpy.load(path, allow_pickle=False) as z:
        out = {k: (z[k] if k in z.files else None) for k in KEYS}
    out["CID"] = str(out["CID"]) if out["CID"] is not None else None
    return out


def read_legacy_pickle(path):
    """Fields of a reference-layout `molecule.<CID>.pickle` (dicts, lists, numbers, strings, arrays only)."""
    d = _safeload.load(path)
    if not isinstance(d, dict):
        raise ValueError(f"{path}: not a molecule cache entry")
    return {k: d.get(k) for k in KEYS}


def _rebuild(fields, basisset, step, fat):
    mol = Molecule([list(map(float, r)) for r in numpy.asarray(fields["xyz"], dtype=numpy.float64)],
                   numpy.asarray(fields["D-CCSD"]), basisset)
    template = Grid.DensityGrid(mol, step, fat)
    out = dict(fields)
    out["mol"] = mol
    for key in ("qgrid", "gvne"):
        if fields[key] is None:
            out[key] = None
            continue
        g = Grid.emptyAs(template, dtype=numpy.asarray(fields[key]).dtype)
        g.LoadFromNPY(numpy.asarray(fields[key]))
        out[key] = g
    return out


def load_batch(directory, basisset, step=0.1, fat=3.0, legacy_pickle=False):
    """Every cached molecule in `directory` with its Molecule and grids rebuilt (reference LoadBatch)."""
    out = []
    for path in sorted(glob.glob(os.path.join(directory, "molecule.*.npz"))):
        out.append(_rebuild(read_entry(path), basisset, step, fat))
    if legacy_pickle:
        for path in sorted(glob.glob(os.path.join(directory, "molecule.*.pickle"))):
            out.append(_rebuild(read_legacy_pickle(path), basisset, step, fat))
    return out


def compute_entry(cid, xyz, dm, basisset, step=0.1, fat=3.0, subgrid=2, adsorb=0.05, diff=0.02,
                  tolerance=1.0e-9, maxiters_factor=20.0):
    """Molecule, density grid and diffusion V_ne as the reference's ingest computes them
    (train-singleshot.py:44-53), on the GPU path; returns the dict `load_batch` gives."""
    from .qmtools import QMTools
    mol = Molecule(xyz, dm, basisset)
    template = Grid.DensityGrid(mol, step, fat)
    qgrid = QMTools.Compute_density(template, mol, subgrid=subgrid, copyBack=True)
    maxiters = maxiters_factor * numpy.power(qgrid.npts, 1.0 / 3)
    gvne = QMTools.Compute_VNe(template, mol, adsorb=adsorb, diff=diff, tolerance=tolerance, maxiters=maxiters,
                               copyBack=True)
    return {"CID": str(cid), "xyz": numpy.asarray(xyz, dtype=numpy.float64), "D-CCSD": numpy.asarray(dm),
            "mol": mol, "qgrid": qgrid, "gvne": gvne}
