"""Molecule cache (offline dataset path): file round trips and the restricted legacy-pickle reader (CPU)."""
import os
import pickle

import numpy
import pytest
from conftest import BASIS, MOL29766


@pytest.fixture(scope="module")
def basis():
    import qmtools
    return qmtools.BasisSet(BASIS)


def _xyz_dm():
    xyz = numpy.loadtxt(os.path.join(MOL29766, "GEOM-B3LYP.xyz"))
    dm = numpy.load(os.path.join(MOL29766, "D-CCSD.npy"), allow_pickle=False)
    return xyz, dm


def test_npz_round_trip_and_batch(tmp_path, basis):
    import qmtools
    from qmtools import cache
    xyz, dm = _xyz_dm()
    mol = qmtools.Molecule(xyz.tolist(), dm, basis)
    tmpl = qmtools.Grid.DensityGrid(mol, 0.4, 2.0)
    rng = numpy.random.default_rng(0)
    q = rng.standard_normal(tmpl.qube.shape)
    g = rng.standard_normal(tmpl.qube.shape).astype(numpy.float32)
    cache.save_entry(str(tmp_path), 29766, xyz, dm, qgrid=q, gvne=g)
    e = cache.read_entry(cache.entry_path(str(tmp_path), 29766))
    assert e["CID"] == "29766"
    assert numpy.array_equal(e["xyz"], xyz) and numpy.array_equal(e["D-CCSD"], dm)
    assert numpy.array_equal(e["qgrid"], q) and numpy.array_equal(e["gvne"], g)
    batch = cache.load_batch(str(tmp_path), basis, step=0.4, fat=2.0)
    assert len(batch) == 1
    b = batch[0]
    assert b["mol"].natoms == 12 and b["mol"].nbf == 162
    assert b["qgrid"].shape == tmpl.shape and numpy.array_equal(b["qgrid"].qube, q)
    assert b["gvne"].qube.dtype == numpy.float32 and numpy.array_equal(b["gvne"].qube, g)


def test_legacy_pickle_layout_is_read(tmp_path, basis):
    from qmtools import cache
    xyz, dm = _xyz_dm()
    d = {"CID": 29766, "xyz": xyz.tolist(), "D-CCSD": dm, "mol": None, "qgrid": None, "gvne": None}
    with open(cache.entry_path(str(tmp_path), 29766, "pickle"), "wb") as fh:
        pickle.dump(d, fh, protocol=pickle.HIGHEST_PROTOCOL)
    e = cache.read_legacy_pickle(cache.entry_path(str(tmp_path), 29766, "pickle"))
    assert e["CID"] == 29766 and numpy.array_equal(e["D-CCSD"], dm)
    assert cache.load_batch(str(tmp_path), basis, step=0.4, fat=2.0) == []        # npz only by default
    b = cache.load_batch(str(tmp_path), basis, step=0.4, fat=2.0, legacy_pickle=True)
    assert len(b) == 1 and b[0]["qgrid"] is None and b[0]["mol"].qtot == 108


class _Evil:
    def __reduce__(self):
        return (os.getcwd, ())


def test_legacy_pickle_refuses_other_globals(tmp_path):
    from qmtools import cache
    p = tmp_path / "molecule.1.pickle"
    p.write_bytes(pickle.dumps({"CID": 1, "xyz": _Evil()}))
    with pytest.raises(pickle.UnpicklingError, match="refused pickle global"):
        cache.read_legacy_pickle(str(p))
