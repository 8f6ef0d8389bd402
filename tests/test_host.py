"""Host-side logic: data layer, grid formats, synthetic inputs, and the C-ABI library's exports (no GPU)."""
import ctypes
import math
import os
import re
import struct

import numpy
import pytest
from conftest import BASIS, MOL29766, ROOT


def test_molecule_29766_tables(basis):
    from qmtools import Grid, Molecule
    mol = Molecule(os.path.join(MOL29766, "GEOM-B3LYP.xyz"), os.path.join(MOL29766, "D-CCSD.npy"), basis)
    assert int(mol.natoms) == 12 and int(mol.qtot) == 108 and mol.nbf == 162
    # ALMO order: atom -> shell -> m in (0, +1, -1, +2, -2) (reference qmtools.py:52,150-171)
    assert [tuple(r) for r in mol.ALMOs[:4]] == [(0, 0, 0, basis.shellOffset[basis.atomOffset[35]])] + \
        [(0, 0, 0, basis.shellOffset[basis.atomOffset[35] + s]) for s in (1, 2, 3)]
    p_rows = [tuple(r[1:3]) for r in mol.ALMOs if r[0] == 0 and r[1] == 1][:3]
    assert p_rows == [(1, 0), (1, 1), (1, -1)]
    d_rows = [int(r[2]) for r in mol.ALMOs if r[0] == 0 and r[1] == 2][:5]
    assert d_rows == [0, 1, -1, 2, -2]
    numpy.testing.assert_allclose(mol.coords[0], numpy.array([2.5775248495833942, 0.007162049772548163,
                                                              -0.009535952472364858]) * 1.8897259886)
    g = Grid.DensityGrid(mol, 0.1, 3.0)
    assert g.shape == (1, 120, 120, 64)
    assert g.GPUblocks == (15, 15, 8)
    assert float(g.step) == pytest.approx(0.1 * 1.8897259886)


def test_molecule_rejects_bad_inputs(basis):
    from qmtools import Molecule
    with pytest.raises(ValueError):
        Molecule([[6, 0, 0, 0]], numpy.eye(13), basis)          # C has 14 functions
    with pytest.raises(ValueError):
        Molecule([[26, 0, 0, 0]], numpy.eye(43), basis)        # Fe has f shells
    with pytest.raises(ValueError):
        Molecule([[19, 0, 0, 0]], numpy.eye(1), basis)         # K is not in the file


def test_grid_shape_rules():
    from qmtools import Grid
    with pytest.raises(ValueError):
        Grid((1, 8, 12, 8), 0.1)
    g = Grid((8, 16, 24), 0.2)
    assert g.shape == (1, 8, 16, 24) and int(g.npts) == 8 * 16 * 24


def test_savebin_loadbin_roundtrip(tmp_path, basis):
    from qmtools import Grid, Molecule
    mol = Molecule([[8, 0.1, 0.2, 0.3], [1, 0.9, 0.0, 0.1], [1, -0.3, 0.8, 0.0]], numpy.eye(24) * 0.1, basis)
    g = Grid((1, 8, 16, 8), 0.3, origin=(-1.5, -2.0, 0.25), dtype=numpy.float32)
    rng = numpy.random.default_rng(0)
    g.qube = rng.standard_normal(g.shape).astype(numpy.float32)
    f = tmp_path / "g.bin"
    g.SaveBIN(str(f), mol)
    raw = f.read_bytes()
    # header: natoms, 3 x (Z, xyz f32), nx ny nz, step, origin
    assert struct.unpack_from("<i", raw, 0)[0] == 3
    assert struct.unpack_from("<i", raw, 4)[0] == 8
    off = 4 + 3 * 16
    assert struct.unpack_from("<iii", raw, off) == (8, 16, 8)
    assert len(raw) == off + 12 + 16 + 4 * 8 * 16 * 8
    h = Grid.LoadBIN(str(f), dtype=numpy.float32)
    assert h.shape == g.shape
    numpy.testing.assert_array_equal(h.qube, g.qube)
    assert float(h.step) == float(numpy.float32(g.step))
    # second round trip is byte-identical
    f2 = tmp_path / "g2.bin"
    h.SaveBIN(str(f2), mol)
    assert f2.read_bytes() == raw
    g2 = Grid((2, 8, 8, 8), 0.3)
    g2.qube = rng.standard_normal(g2.shape)
    g2.SaveBINmulti(str(tmp_path / "m"))
    for fld in range(2):
        back = Grid.LoadBIN(str(tmp_path / f"m-{fld}.bin"))
        numpy.testing.assert_array_equal(back.qube[0], g2.qube[fld].astype(numpy.float32))


def test_save_load_pickle(tmp_path):
    from qmtools import Grid
    g = Grid((1, 8, 8, 8), 0.3, origin=(1.0, 2.0, 3.0))
    g.qube[:] = numpy.arange(512).reshape(g.shape)
    g.Save(str(tmp_path / "g.pkl"))
    h = Grid.Load(None, str(tmp_path / "g.pkl"))
    numpy.testing.assert_array_equal(h.qube, g.qube)
    assert h.origin["y"] == 2.0 and h.step == g.step


class _Planted:
    """Unpickles to a call of os.getcwd: the kind of global Grid.Load must refuse."""
    def __reduce__(self):
        import os
        return (os.getcwd, ())


def test_load_pickle_refuses_foreign_globals(tmp_path):
    import pickle
    from qmtools import Grid
    path = str(tmp_path / "bad.pkl")
    with open(path, "wb") as fh:
        pickle.dump({"shape": (1, 8, 8, 8), "origin": None, "step": 0.1, "lattice": None,
                     "qube": _Planted()}, fh)
    with pytest.raises(pickle.UnpicklingError, match="refused pickle global"):
        Grid.Load(None, path)
    with open(path, "wb") as fh:
        pickle.dump([1, 2, 3], fh)
    with pytest.raises(ValueError, match="not a grid"):
        Grid.Load(None, path)


def test_field_xyz_is_x_fastest():
    from qmtools import Grid
    g = Grid((1, 8, 16, 24), 0.1)
    g.qube[0].reshape(-1)[3 + 8 * (5 + 16 * 7)] = 1.0
    assert g.field_xyz()[3, 5, 7] == 1.0


@pytest.mark.parametrize("natoms,seed", [(12, 1), (60, 2), (5, 3)])
def test_synthetic_molecule_rules(basis, natoms, seed):
    from qmtools import synth
    xyz = synth.random_geometry(natoms, seed)
    z = xyz[:, 0]
    assert len(z) == natoms and int(z.sum()) % 2 == 0
    pos = xyz[:, 1:]
    d = numpy.linalg.norm(pos[:, None] - pos[None], axis=-1) + numpy.eye(natoms) * 99
    heavy = z > 1
    for i in numpy.nonzero(~heavy)[0]:
        assert abs(d[i][heavy].min() - 1.09) < 1e-9          # each H bonded at 1.09 A
    hh = numpy.nonzero(heavy)[0]
    if len(hh) > 1:
        # every heavy atom has a bonded heavy neighbour at 1.3-1.5 A; no contact is below 1.2 A
        dh = d[numpy.ix_(hh, hh)]
        assert numpy.all(numpy.any((dh >= 1.3 - 1e-9) & (dh <= 1.5 + 1e-9), axis=1))
    assert numpy.all(d[numpy.ix_(hh, hh)] >= 1.2 - 1e-9)
    mol = synth.synthetic_molecule(natoms, seed, basis)
    P = 2 * mol.dm
    # P = 2 C C^T with orthonormal C: (P/2)^2 = P/2, trace = n_occ
    numpy.testing.assert_allclose(mol.dm @ mol.dm, mol.dm, atol=1e-12)
    assert numpy.trace(P) == pytest.approx(float(mol.qtot), rel=1e-12)


def _declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "qmgrid.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(qm_\w+)\s*\(", hdr, re.M)))


@pytest.fixture(scope="module")
def built_lib():
    from qmtools import _build, _native
    if not os.path.exists(_native.LIB_PATH):
        _build.build(verbose=False)
    return _native.LIB_PATH


def test_library_exports_every_declared_symbol(built_lib):
    from qmtools import _native
    names = _declared_symbols()
    assert len(names) >= 20
    lib = ctypes.CDLL(built_lib)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_native.SIGNATURES) == names
    assert lib.qm_version() == 10000


def test_compute_without_gpu_fails_loudly(basis):
    """No CPU fallback: a compute call on a machine without a HIP device raises."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from qmtools import Molecule, QMTools, Grid
    mol = Molecule([[1, 0, 0, 0], [1, 0.74, 0, 0]], numpy.eye(10) * 0.1, basis)
    with pytest.raises(RuntimeError):
        QMTools.Compute_density(Grid((1, 8, 8, 8), 0.3), mol)
