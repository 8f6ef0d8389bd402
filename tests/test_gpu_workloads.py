"""The hot path at full workload sizes against the FP64 oracle.

* 12-atom molecule, 96^3 grid: FP64 density and 2000 Jacobi sweeps with multipole
  boundaries, compared on every voxel with the oracle's closed form (and with the
  committed fixture); the mixed path within 1e-6 of it.
* 60-atom molecule (606 basis functions), 256^3 grid: FP64 density at 4096 sampled
  voxels, its total over all 16.8M voxels, the free-space FFT potential (512^3
  transforms) at 64 voxels and the GTO potential at 64 of the 10^4 points, against
  fixtures written by tests/golden/make_workload_fixtures.py from the oracle; the
  mixed path (FP16 hi/lo density, complex64 FFT, FP32 Jacobi) within 1e-6 of FP64.
* 256 molecules of 5-80 atoms on 64^3-192^3 grids, batched: charge conservation per
  molecule, batch results bitwise equal to single-molecule runs, mixed within 1e-6 of FP64.
* 200 requests arriving as a Poisson stream at 50 req/s into the serving loop: no
  failure, each sampled result bitwise the same as the molecule computed alone.

Tolerances: FP64 1e-10 relative, mixed path 1e-6 relative (normwise: max|diff| / max|ref|
over the compared values; integrated charge |sum - ref| / ref).  BF16 has its own bound.
"""
import math
import os

import numpy
import pytest
from conftest import BASIS, GOLDEN

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _rel(a, b):
    a = numpy.asarray(a, dtype=numpy.float64).ravel()
    b = numpy.asarray(b, dtype=numpy.float64).ravel()
    return float(numpy.abs(a - b).max() / numpy.abs(b).max())


def _fixture(name):
    with numpy.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _same_inputs(fx, mol, grid):
    from qmtools import synth
    nx, ny, nz, h, o = grid.geometry()
    numpy.testing.assert_array_equal(mol.types, fx["types"])
    numpy.testing.assert_allclose(mol.coords, fx["coords"], rtol=0, atol=1e-12)
    assert _rel(synth.dm_probe(mol.dm), fx["dm_probe"]) <= 1e-12
    assert (nx, ny, nz) == tuple(fx["shape"]) and h == pytest.approx(float(fx["step"]), rel=1e-15)
    numpy.testing.assert_allclose(o, fx["origin"], rtol=0, atol=1e-12)


@pytest.fixture(scope="module")
def basis():
    import qmtools
    return qmtools.BasisSet(BASIS)


# ------------------------------------------------------------------ 12 atoms, 96^3, 2000 sweeps
def test_mol12_grid96_density_and_2000_jacobi_sweeps(basis):
    import qmtools
    from oracle import jacobi as oj
    from oracle import seed as osd
    from qmtools import synth
    fx = _fixture("mol12_grid96.npz")
    mol, grid = synth.molecule_12_grid_96(basis)
    _same_inputs(fx, mol, grid)
    nx, ny, nz, h, o = grid.geometry()
    idx = fx["idx"]
    _say("12 atoms 96^3 fp64 density")
    q = qmtools.QMTools.Compute_density(grid, mol)
    qh = q.qube.ravel()
    assert _rel(qh[idx], fx["q"]) <= 1e-10
    assert q.raw_total == pytest.approx(float(fx["raw_total"]), rel=1e-10)
    assert math.fsum(qh) == pytest.approx(float(mol.qtot), rel=1e-12)
    _say("12 atoms 96^3 fp64 jacobi 2000 multipole")
    v = qmtools.QMTools.Compute_hartree_iter(q, mol, tolerance=None, max_sweeps=2000, boundary="multipole")
    assert v.sweeps == 2000
    vh = v.qube.ravel()
    rho = osd.seed_grid(mol.types, mol.coords, nx, ny, nz, h, o) - qh
    want = oj.jacobi_closed_form(rho, nx, ny, nz, h, 2000, bvals=oj.multipole_boundary(rho, nx, ny, nz, h, o))
    assert _rel(vh, want) <= 1e-10                      # every voxel
    assert _rel(vh[idx], fx["v_jacobi"]) <= 1e-10
    assert math.fsum(vh) == pytest.approx(float(fx["v_sum"]), rel=1e-10)
    assert math.fsum(numpy.abs(vh)) == pytest.approx(float(fx["v_abs_sum"]), rel=1e-10)
    _say("12 atoms 96^3 mixed density + jacobi 2000")
    qm = qmtools.QMTools.Compute_density(grid, mol, precision="mixed")
    qmh = qm.qube.ravel()
    assert _rel(qmh, qh) <= 1e-6
    assert _rel(qmh[idx], fx["q"]) <= 1e-6
    assert abs(math.fsum(qmh) - float(mol.qtot)) / float(mol.qtot) <= 1e-6
    vm = qmtools.QMTools.Compute_hartree_iter(qm, mol, tolerance=None, max_sweeps=2000, boundary="multipole",
                                              precision="mixed")
    err = _rel(vm.qube, want)
    print(f"12 atoms 96^3: mixed rho {_rel(qmh, qh):.2e}, mixed V after 2000 sweeps {err:.2e}")
    assert err <= 1e-6


# ------------------------------------------------------------------ 60 atoms, 256^3
@pytest.fixture(scope="module")
def mol60(basis):
    from qmtools import synth
    fx = _fixture("mol60_grid256.npz")
    mol, grid, pts = synth.molecule_60_grid_256(basis)
    _same_inputs(fx, mol, grid)
    numpy.testing.assert_allclose(pts[:64], fx["gto_points"], rtol=0, atol=1e-12)
    return mol, grid, pts, fx


@pytest.fixture(scope="module")
def mol60_fp64(mol60):
    """FP64 density (host copy) and free-space FFT potential of the 60-atom workload."""
    import qmtools
    mol, grid, pts, fx = mol60
    _say("60 atoms 256^3 fp64 density")
    q = qmtools.QMTools.Compute_density(grid, mol)
    _say("60 atoms 256^3 fp64 free-space FFT (512^3)")
    v = qmtools.QMTools.Compute_hartree_fft(q, mol, sigma=float(fx["sigma"]), cutoff=float(fx["cutoff"]),
                                            mode="freespace")
    return q, v


def _fft_samples(v, grid, fx):
    nx, ny, _, _, _ = grid.geometry()
    t = fx["fft_idx"]
    return v[t % nx, (t // nx) % ny, t // (nx * ny)]


def test_mol60_grid256_fp64_density(mol60, mol60_fp64):
    mol, grid, _, fx = mol60
    q, _ = mol60_fp64
    qh = q.qube.ravel()
    want = fx["raw"] * (float(mol.qtot) / float(fx["raw_total"]))
    assert _rel(qh[fx["idx"]], want) <= 1e-10
    assert q.raw_total == pytest.approx(float(fx["raw_total"]), rel=1e-10)
    assert math.fsum(qh) == pytest.approx(float(mol.qtot), rel=1e-12)


def test_mol60_grid256_fp64_fft_freespace(mol60, mol60_fp64):
    _, grid, _, fx = mol60
    _, v = mol60_fp64
    assert v.shape == tuple(fx["shape"])
    assert _rel(_fft_samples(v, grid, fx), fx["v_fft"]) <= 1e-10


def test_mol60_gto_points(mol60):
    """64 points (the distribution list split over workgroups, partial sums reduced) and the
    first 64 of all 10^4 points (one pass) against the oracle."""
    import qmtools
    mol, _, pts, fx = mol60
    _say("60 atoms GTO potential at 64 points")
    v64 = qmtools.QMTools.Compute_hartree_points(mol, pts[:64])
    assert _rel(v64, fx["v_gto"]) <= 1e-10
    _say("60 atoms GTO potential at 10^4 points")
    vall = qmtools.QMTools.Compute_hartree_points(mol, pts)
    assert vall.shape == (pts.shape[0],)
    assert _rel(vall[:64], fx["v_gto"]) <= 1e-10
    assert numpy.all(numpy.isfinite(vall))


def test_mol60_grid256_mixed(mol60, mol60_fp64):
    """FP16 hi/lo density and complex64 FFT within 1e-6 of FP64 (whole grid) and of the oracle."""
    import qmtools
    mol, grid, _, fx = mol60
    q64, v64 = mol60_fp64
    q64h = q64.qube.ravel()
    _say("60 atoms 256^3 mixed density")
    q = qmtools.QMTools.Compute_density(grid, mol, precision="mixed")
    qh = q.qube.ravel()
    e_rho = _rel(qh, q64h)
    e_fix = _rel(qh[fx["idx"]], fx["raw"] * (float(mol.qtot) / float(fx["raw_total"])))
    e_chg = abs(math.fsum(qh) - float(mol.qtot)) / float(mol.qtot)
    e_raw = abs(q.raw_total - float(fx["raw_total"])) / float(fx["raw_total"])
    _say("60 atoms 256^3 FP32 free-space FFT")
    v = qmtools.QMTools.Compute_hartree_fft(q, mol, sigma=float(fx["sigma"]), cutoff=float(fx["cutoff"]),
                                            mode="freespace", precision="mixed")
    e_v = _rel(v, v64)
    e_vfix = _rel(_fft_samples(v, grid, fx), fx["v_fft"])
    print(f"60 atoms 256^3 mixed: rho {e_rho:.2e} (fixture {e_fix:.2e}), charge {e_chg:.2e}, raw total {e_raw:.2e}, "
          f"V_H {e_v:.2e} (fixture {e_vfix:.2e})")
    assert max(e_rho, e_fix, e_chg, e_raw, e_v, e_vfix) <= 1e-6


def test_mol60_grid256_bf16_density(mol60, mol60_fp64):
    """BF16 hi/lo keeps 16 significant bits: its own bound, 2e-5 relative."""
    import qmtools
    mol, grid, _, _ = mol60
    q64, _ = mol60_fp64
    _say("60 atoms 256^3 bf16 density")
    q = qmtools.QMTools.Compute_density(grid, mol, precision="bf16")
    err = _rel(q.qube, q64.qube)
    print(f"60 atoms 256^3 bf16 rho error {err:.2e}")
    assert err <= 2e-5


def test_mol60_grid256_jacobi_fp32_vs_fp64(mol60, mol60_fp64):
    """2000 sweeps with multipole boundaries on 256^3: FP32 defect correction within 1e-6 of FP64."""
    import qmtools
    mol, _, _, _ = mol60
    q64, _ = mol60_fp64
    _say("60 atoms 256^3 jacobi 2000 fp64 / fp32")
    a = qmtools.QMTools.Compute_hartree_iter(q64, mol, tolerance=None, max_sweeps=2000, boundary="multipole",
                                             copyBack=False)
    b = qmtools.QMTools.Compute_hartree_iter(q64, mol, tolerance=None, max_sweeps=2000, boundary="multipole",
                                             precision="mixed", copyBack=False)
    err = float((a.d_qube - b.d_qube).abs().max() / a.d_qube.abs().max())
    print(f"256^3 Jacobi 2000 sweeps: FP32 vs FP64 {err:.2e}")
    assert a.sweeps == b.sweeps == 2000
    assert err <= 1e-6


# ------------------------------------------------------------------ 256 molecules, batched
@pytest.fixture(scope="module")
def batch256(basis):
    from qmtools import synth
    return synth.molecule_batch(256, basis=basis)


@pytest.mark.parametrize("poisson", ["jacobi", "fft"])
def test_batch_256_molecules(batch256, poisson):
    import torch
    from qmtools import batch
    mols, grids = batch256
    _say("batch 256 molecules", poisson, batch.grid_points(grids), "points")
    lane = batch.Lane(default=True)
    qs, vs = batch.run(mols, grids, "mixed", 2000, lane=lane, poisson=poisson)
    q64, v64 = batch.run(mols, grids, "fp64", 2000, lane=lane, poisson=poisson)
    torch.cuda.synchronize()
    worst = {"rho": 0.0, "charge": 0.0, "v": 0.0}
    for m, q, v, qr, vr in zip(mols, qs, vs, q64, v64):
        z = float(m.qtot)
        assert abs(float(qr.sum()) - z) / z <= 1e-12
        worst["charge"] = max(worst["charge"], abs(float(q.double().sum()) - z) / z)
        worst["rho"] = max(worst["rho"], float((q.double() - qr).abs().max() / qr.abs().max()))
        worst["v"] = max(worst["v"], float((v.double() - vr).abs().max() / vr.abs().max()))
    print(f"256 molecules ({poisson}): mixed vs FP64 worst {worst}")
    assert max(worst.values()) <= 1e-6
    for i in range(0, 256, 32):   # a molecule's results are the same alone as in the batch
        _say("batch 256 alone", poisson, i)
        qa, va = batch.run([mols[i]], [grids[i]], "mixed", 2000, lane=lane, poisson=poisson)
        assert torch.equal(qa[0], qs[i]) and torch.equal(va[0], vs[i]), i
    del qs, vs, q64, v64


# ------------------------------------------------------------------ served, Poisson arrivals
def test_served_poisson_stream_50_per_s(basis):
    import torch
    from qmtools import batch, serve, synth
    mols, grids = synth.molecule_batch(200, (5, 120), (64, 96, 128, 160, 192, 224, 256), (0.1, 0.15), 77, basis)
    _say("served 200 requests at 50 req/s", batch.grid_points(grids), "points")
    srv = serve.Server(nlanes=2, precision="mixed", max_batch=8, sweeps=2000, to_host=False).start()
    reqs = serve.poisson_arrivals(srv, list(zip(mols, grids)), 50.0, seed=5)
    for r in reqs:
        r.wait(600)
    srv.stop()
    st = srv.stats()
    print(f"served: {st}")
    assert st["requests"] == 200 and st["failed"] == 0
    assert st["max_concurrent_batches"] >= 2
    lane = batch.Lane(default=True)
    for i in range(0, 200, 40):
        _say("served alone", i)
        qa, va = batch.run([mols[i]], [grids[i]], "mixed", 2000, lane=lane)
        assert torch.equal(qa[0], reqs[i].density) and torch.equal(va[0], reqs[i].potential), i
        q64, v64 = batch.run([mols[i]], [grids[i]], "fp64", 2000, lane=lane)
        assert float((va[0] - v64[0]).abs().max() / v64[0].abs().max()) <= 1e-6
        assert float((qa[0].double() - q64[0]).abs().max() / q64[0].abs().max()) <= 1e-6
