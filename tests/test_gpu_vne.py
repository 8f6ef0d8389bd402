"""GPU parity of the diffusion pseudo-V_ne relaxation (Compute_VNe) against the FP64 oracle.

FP64: field within 1e-12 (normwise relative), same pair count and stop reason.
FP32 storage (an optional variant; the mixed density + Hartree path keeps V_ne in FP64): 21 fixed
pairs within 2e-6, the accumulated rounding of 42 sweeps stored in FP32.  Inputs are printed before
each GPU step.
"""
import os
import warnings

import numpy
import pytest
from conftest import BASIS, MOL29766

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = numpy.asarray(a, dtype=numpy.float64).ravel()
    b = numpy.asarray(b, dtype=numpy.float64).ravel()
    return float(numpy.abs(a - b).max() / numpy.abs(b).max())


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def qm():
    import qmtools
    return qmtools


def _oracle(mol, grid, **kw):
    from oracle import vne as ov
    nx, ny, nz, h, o = grid.geometry()
    return ov.vne_diffusion(mol.types, mol.coords, nx, ny, nz, h, o, **kw)


def test_vne_diffusion_29766_converged(qm):
    basis = qm.BasisSet(BASIS)
    mol = qm.Molecule(os.path.join(MOL29766, "GEOM-B3LYP.xyz"), os.path.join(MOL29766, "D-CCSD.npy"), basis)
    grid = qm.Grid.DensityGrid(mol, 0.3, 2.0)
    _say("vne diffusion 29766", grid.shape)
    v = qm.QMTools.Compute_VNe(grid, mol, tolerance=1e-7)
    want, pairs, delta, conv = _oracle(mol, grid, tolerance=1e-7)
    assert conv and v.pairs == pairs
    assert abs(v.delta - delta) <= 1e-9 * delta
    assert _rel(v.qube, want) <= 1e-12


def test_vne_diffusion_ragged_many_atoms(qm):
    """C-ABI call on a 70 x 13 x 9 grid (no tile multiple) with more nuclei near one workgroup
    than its list holds (80 atoms, 2 bohr step), stopped by the pair cap."""
    import ctypes
    import torch
    from qmtools import _native as nat
    from qmtools import synth
    from qmtools.qmtools import device_molecule
    from oracle import vne as ov
    basis = qm.BasisSet(BASIS)
    mol = synth.synthetic_molecule(80, 5, basis)
    c = numpy.asarray(mol.coords)
    nx, ny, nz, h = 70, 13, 9, 2.0
    o = tuple(float(x) for x in (c.min(axis=0) - 1.5))
    _say("vne diffusion ragged", (nx, ny, nz), "atoms", mol.natoms)
    v = torch.empty(nx * ny * nz, dtype=torch.float64, device="cuda")
    t = torch.empty_like(v)
    pairs, delta, conv = ctypes.c_int64(0), ctypes.c_double(0), ctypes.c_int(1)
    nat.call("qm_vne_diffusion", device_molecule(mol).handle, nat.stream(), ctypes.byref(nat.geom((nx, ny, nz, h, o))),
             nat.QM_F64, 0.05, 0.02, -1.0, 30.0, nat.dptr(v), nat.dptr(t), ctypes.byref(pairs), ctypes.byref(delta),
             ctypes.byref(conv))
    want, wp, wd, wc = ov.vne_diffusion(mol.types, mol.coords, nx, ny, nz, h, o, adsorb=0.05, diff=0.02,
                                        tolerance=-1.0, maxiters=30)
    assert not wc and conv.value == 0 and wp == pairs.value == 31
    assert abs(delta.value - wd) <= 1e-10 * wd
    assert _rel(v.cpu().numpy(), want) <= 1e-12


def test_vne_diffusion_fp32_fixed_pairs(qm):
    from qmtools import synth
    basis = qm.BasisSet(BASIS)
    mol = synth.synthetic_molecule(10, 8, basis)
    grid = synth.centred_grid(mol, 40, step_bohr=0.3, pad_ang=2.0)
    _say("vne diffusion fp32", grid.shape)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        v = qm.QMTools.Compute_VNe(grid, mol, tolerance=-1.0, maxiters=20, precision="mixed")
    want, pairs, _, _ = _oracle(mol, grid, tolerance=-1.0, maxiters=20)
    assert v.qube.dtype == numpy.float32 and v.pairs == pairs == 21
    assert _rel(v.qube, want) <= 2e-6


def test_vne_diffusion_no_sweep_when_tolerance_large(qm):
    from qmtools import synth
    basis = qm.BasisSet(BASIS)
    mol = synth.synthetic_molecule(4, 2, basis)
    grid = synth.centred_grid(mol, 16, step_bohr=0.4, pad_ang=2.0)
    _say("vne diffusion tol>=1", grid.shape)
    v = qm.QMTools.Compute_VNe(grid, mol, tolerance=1.0)
    assert v.pairs == 0 and not numpy.asarray(v.qube).any()


def test_cache_ingest_round_trip(qm, tmp_path):
    """Density (subgrid 2) and diffusion V_ne of the ingest step, stored and read back unchanged."""
    from qmtools import cache
    from oracle import vne as ov
    basis = qm.BasisSet(BASIS)
    xyz = numpy.loadtxt(os.path.join(MOL29766, "GEOM-B3LYP.xyz"))
    dm = numpy.load(os.path.join(MOL29766, "D-CCSD.npy"), allow_pickle=False)
    _say("cache ingest 29766 step 0.4")
    e = cache.compute_entry(29766, xyz.tolist(), dm, basis, step=0.4, fat=2.0, tolerance=1e-6)
    assert abs(float(e["qgrid"].qube.sum()) - 108.0) < 1e-9
    nx, ny, nz, h, o = e["gvne"].geometry()
    want, pairs, _, _ = ov.vne_diffusion(e["mol"].types, e["mol"].coords, nx, ny, nz, h, o, adsorb=0.05, diff=0.02,
                                         tolerance=1e-6, maxiters=20 * numpy.power(nx * ny * nz, 1.0 / 3))
    assert e["gvne"].pairs == pairs and _rel(e["gvne"].qube, want) <= 1e-12
    cache.save_entry(str(tmp_path), 29766, xyz, dm, e["qgrid"], e["gvne"])
    b = cache.load_batch(str(tmp_path), basis, step=0.4, fat=2.0)[0]
    assert numpy.array_equal(b["qgrid"].qube, e["qgrid"].qube)
    assert numpy.array_equal(b["gvne"].qube, e["gvne"].qube)


def test_vne_diffusion_negative_cap_runs_one_pair(qm):
    from qmtools import synth
    basis = qm.BasisSet(BASIS)
    mol = synth.synthetic_molecule(4, 2, basis)
    grid = synth.centred_grid(mol, 16, step_bohr=0.4, pad_ang=2.0)
    _say("vne diffusion maxiters=-0.5", grid.shape)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        v = qm.QMTools.Compute_VNe(grid, mol, tolerance=-1.0, maxiters=-0.5)
    want, pairs, _, conv = _oracle(mol, grid, tolerance=-1.0, maxiters=-0.5)
    assert v.pairs == pairs == 1 and not conv
    assert _rel(v.qube, want) <= 1e-12
