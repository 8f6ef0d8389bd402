"""GPU parity of every hot-path entry point against the FP64 oracle (small seeded inputs).

Tolerances: FP64 paths <= 1e-10 relative (normwise: max|diff| / max|ref|);
mixed paths as stated per test.  Inputs are printed before each GPU step so
that a failure names its case.
"""
import math
import os

import numpy
import pytest
from conftest import BASIS, MOL29766

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = numpy.asarray(a, dtype=numpy.float64).ravel()
    b = numpy.asarray(b, dtype=numpy.float64).ravel()
    return float(numpy.abs(a - b).max() / numpy.abs(b).max())


@pytest.fixture(scope="module")
def qm():
    import qmtools
    return qmtools


@pytest.fixture(scope="module")
def obasis():
    from oracle.basisfile import read_basis
    return read_basis(BASIS)


def _system(mol, obasis):
    from oracle.system import make_system
    return make_system(mol.types, mol.coords, mol.dm, obasis)


@pytest.fixture(scope="module")
def small(qm, obasis):
    from qmtools import synth
    basis = qm.BasisSet(BASIS)
    mol = synth.synthetic_molecule(6, 21, basis)
    grid = synth.centred_grid(mol, 32, step_bohr=0.35, pad_ang=2.0)
    return mol, grid, _system(mol, obasis)


@pytest.fixture(scope="module")
def mol29766(qm, obasis):
    basis = qm.BasisSet(BASIS)
    mol = qm.Molecule(os.path.join(MOL29766, "GEOM-B3LYP.xyz"), os.path.join(MOL29766, "D-CCSD.npy"), basis)
    grid = qm.Grid.DensityGrid(mol, 0.4, 3.0)
    return mol, grid, _system(mol, obasis)


def _say(*a):
    print("[case]", *a, flush=True)


# ---------------------------------------------------------------- density
@pytest.mark.parametrize("subgrid,variant", [(1, 0), (2, 0), (1, 1), (2, 1), (1, 2), (2, 2), (1, 3), (2, 3), (1, 4), (2, 4)])
def test_density_fp64_synthetic(qm, small, subgrid, variant):
    from oracle import density as od
    from qmtools import autotune
    mol, grid, sysm = small
    nx, ny, nz, h, o = grid.geometry()
    _say("density fp64 synthetic", grid.shape, "subgrid", subgrid, "nbf", mol.nbf, "variant", variant)
    old = autotune.get("f64_variant")
    autotune.set("f64_variant", variant)
    try:
        q = qm.QMTools.Compute_density(grid, mol, subgrid=subgrid)
    finally:
        autotune.set("f64_variant", old)
    want, raw = od.density_grid(sysm, nx, ny, nz, h, o, subgrid=subgrid)
    assert _rel(q.qube, want) <= 1e-10
    assert q.raw_total == pytest.approx(raw, rel=1e-10)
    assert math.fsum(q.qube.ravel()) == pytest.approx(float(mol.qtot), rel=1e-12)


def test_density_fp64_mol29766(qm, mol29766):
    from oracle import density as od
    mol, grid, sysm = mol29766
    nx, ny, nz, h, o = grid.geometry()
    _say("density fp64 29766", grid.shape)
    q = qm.QMTools.Compute_density(grid, mol)
    want, raw = od.density_grid(sysm, nx, ny, nz, h, o)
    assert _rel(q.qube, want) <= 1e-10
    assert q.raw_total == pytest.approx(raw, rel=1e-10)


@pytest.mark.parametrize("prec,tol", [("f16", 1e-6), ("bf16", 2e-5)])
def test_density_mixed(qm, small, prec, tol):
    from oracle import density as od
    mol, grid, sysm = small
    nx, ny, nz, h, o = grid.geometry()
    _say("density", prec, grid.shape)
    q = qm.QMTools.Compute_density(grid, mol, precision=prec)
    want, raw = od.density_grid(sysm, nx, ny, nz, h, o)
    err = _rel(q.qube, want)
    print(f"{prec}: normwise rho error {err:.3e}, raw total rel err {abs(q.raw_total - raw) / raw:.3e}")
    assert err <= tol
    assert abs(q.raw_total - raw) / raw <= tol


@pytest.mark.parametrize("natoms,seed", [(1, 3), (2, 5), (3, 7), (9, 2), (12, 4), (13, 8), (15, 9)])
def test_density_block_alignment(qm, obasis, natoms, seed):
    """nbf 14 .. 156: the default contractions right-align their 128-column blocks on nbf, so
    block 0 holds 0 .. 112 padding columns below column 0 (and nbf = 128 none at all)."""
    from oracle import density as od
    from qmtools import synth
    mol = synth.synthetic_molecule(natoms, seed, qm.BasisSet(BASIS))
    grid = synth.centred_grid(mol, 16, step_bohr=0.5, pad_ang=2.0)
    nx, ny, nz, h, o = grid.geometry()
    _say("density block alignment", natoms, seed, "nbf", mol.nbf)
    want, raw = od.density_grid(_system(mol, obasis), nx, ny, nz, h, o)
    q = qm.QMTools.Compute_density(grid, mol)
    assert _rel(q.qube, want) <= 1e-10
    q = qm.QMTools.Compute_density(grid, mol, precision="mixed")
    assert _rel(q.qube, want) <= 1e-6


@pytest.mark.parametrize("prec,tol", [("f16", 1e-6), ("bf16", 2e-5)])
@pytest.mark.parametrize("subgrid", [1, 2])
@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
def test_density_mixed_tiled(qm, small, prec, tol, subgrid, variant):
    """The tiled split contractions (partial sums per column block; 4 evaluates the basis values inside
    the contraction kernel) against the oracle."""
    from oracle import density as od
    from qmtools import autotune
    mol, grid, sysm = small
    nx, ny, nz, h, o = grid.geometry()
    _say("density tiled", prec, grid.shape, "subgrid", subgrid, "variant", variant)
    old = autotune.get("split_variant")
    autotune.set("split_variant", variant)
    try:
        q = qm.QMTools.Compute_density(grid, mol, subgrid=subgrid, precision=prec)
    finally:
        autotune.set("split_variant", old)
    want, raw = od.density_grid(sysm, nx, ny, nz, h, o, subgrid=subgrid)
    err = _rel(q.qube, want)
    print(f"{prec} tiled subgrid {subgrid}: normwise rho error {err:.3e}")
    assert err <= tol
    assert abs(q.raw_total - raw) / raw <= tol


# ---------------------------------------------------------------- reductions / seed / vne
def test_reductions(qm, small):
    from oracle import reduce as orr
    mol, grid, _ = small
    rng = numpy.random.default_rng(0)
    a = qm.Grid.emptyAs(grid)
    b = qm.Grid.emptyAs(grid)
    a.LoadFromNPY(rng.standard_normal(a.shape))
    b.LoadFromNPY(rng.uniform(1.0, 2.0, b.shape))
    _say("reductions", grid.shape)
    assert qm.QMTools.Compute_qtot(a) == pytest.approx(orr.total(a.qube), rel=1e-13, abs=1e-12)
    assert qm.QMTools.Compute_qdiff(a, b) == pytest.approx(orr.absdiff(a.qube, b.qube), rel=1e-13)
    assert qm.QMTools.Compute_qdiff(a, b, rel=True) == pytest.approx(orr.reldiff(a.qube, b.qube), rel=1e-13)
    before = a.qube.copy()
    qm.QMTools.Compute_qscale(a, 2.5, copyBack=True)
    numpy.testing.assert_array_equal(a.qube, 2.5 * before)


def test_seed(qm, small):
    from oracle import seed as osd
    mol, grid, _ = small
    nx, ny, nz, h, o = grid.geometry()
    _say("seed", grid.shape)
    s = qm.QMTools.Compute_qseed(grid, mol)
    want = osd.seed_grid(mol.types, mol.coords, nx, ny, nz, h, o)
    numpy.testing.assert_allclose(s.qube.ravel(), want, rtol=1e-13, atol=1e-14)
    assert math.fsum(s.qube.ravel()) == pytest.approx(float(mol.qtot), rel=1e-13)


@pytest.mark.parametrize("divisions,n", [(2, 32), (4, 32), (16, 16)])
def test_vne_classic(qm, small, divisions, n):
    """divisions 4 is the default of Compute_VNe_classic (reference qmtools.py:2750), 16 what the
    reference's scripts/test-vneclassic.py:24 uses."""
    from oracle import vne as ov
    from qmtools import synth
    mol = small[0]
    grid = synth.centred_grid(mol, n, step_bohr=0.35, pad_ang=2.0)
    nx, ny, nz, h, o = grid.geometry()
    _say("vne classic", grid.shape, "divisions", divisions)
    v = qm.QMTools.Compute_VNe_classic(grid, mol, divisions=divisions)
    want = ov.vne_classic(mol.types, mol.coords, nx, ny, nz, h, o, divisions)
    assert _rel(v.qube, want) <= 1e-12


# ---------------------------------------------------------------- Jacobi
@pytest.mark.parametrize("precision,sweeps", [("fp64", 250), ("fp64", 7), ("fp32", 250), ("fp32", 53),
                                              ("fp32", 1), ("fp32", 2)])
def test_jacobi_ragged_grid(qm, precision, sweeps):
    """qm_jacobi on a 70 x 13 x 9 grid (no strip, tile or slab multiple) with a held nonzero outer layer:
    two-sweep passes, single sweeps and graph replays against the oracle's plain sweeps."""
    import ctypes
    import torch
    from qmtools import _native as nat
    from oracle import jacobi as oj
    nx, ny, nz, h = 70, 13, 9, 0.3
    rng = numpy.random.default_rng(5)
    rho = rng.standard_normal(nx * ny * nz)
    bvals = rng.standard_normal(nx * ny * nz)
    v0 = numpy.where(oj.boundary_mask(nx, ny, nz), bvals, 0.0)
    _say("jacobi ragged", precision, sweeps)
    dr = torch.tensor(rho, device="cuda")
    dv = torch.tensor(v0, device="cuda")
    dt = dv.clone()
    done, delta = ctypes.c_int64(0), ctypes.c_double(0)
    g = nat.geom((nx, ny, nz, h, (0.0, 0.0, 0.0)))
    if precision == "fp64":
        args = (nat.QM_PREC_FP64, nat.dptr(dr), nat.dptr(dv), nat.dptr(dt), sweeps, -1.0, 0)
    else:
        args = (nat.QM_PREC_FP32, nat.dptr(dr), nat.dptr(dv), None, sweeps, -1.0, 50)
    nat.call("qm_jacobi", nat.stream(), ctypes.byref(g), *args, ctypes.byref(done), ctypes.byref(delta))
    want = oj.jacobi(rho, nx, ny, nz, h, sweeps, bvals=bvals, v0=v0)
    assert done.value == sweeps
    assert _rel(dv.cpu().numpy(), want) <= (1e-12 if precision == "fp64" else 1e-6)



@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_jacobi_slab_size_invariant(qm, precision):
    """Sweeps on 64 x 40 x 61 are bitwise the same for every z slab per workgroup: odd and even
    plane counts take different unrolled copies of the two-sweep loop, which must round alike."""
    import ctypes
    import torch
    from qmtools import _native as nat, autotune
    nx, ny, nz, h, sweeps = 64, 40, 61, 0.25, 46
    rng = numpy.random.default_rng(11)
    rho = torch.tensor(rng.standard_normal(nx * ny * nz), device="cuda")
    g = nat.geom((nx, ny, nz, h, (0.0, 0.0, 0.0)))
    old = autotune.get("jacobi_zs")
    outs = {}
    try:
        for zs in (0, 4, 5, 7, 13, 61):
            _say("jacobi slab size", precision, zs)
            autotune.set("jacobi_zs", zs)
            dv = torch.zeros_like(rho)
            dt = torch.zeros_like(rho)
            done, delta = ctypes.c_int64(0), ctypes.c_double(0)
            if precision == "fp64":
                args = (nat.QM_PREC_FP64, nat.dptr(rho), nat.dptr(dv), nat.dptr(dt), sweeps, -1.0, 0)
            else:
                args = (nat.QM_PREC_FP32, nat.dptr(rho), nat.dptr(dv), None, sweeps, -1.0, 50)
            nat.call("qm_jacobi", nat.stream(), ctypes.byref(g), *args, ctypes.byref(done), ctypes.byref(delta))
            outs[zs] = dv.cpu().numpy()
    finally:
        autotune.set("jacobi_zs", old)
    for zs, v in outs.items():
        numpy.testing.assert_array_equal(v, outs[0], err_msg=f"z slab {zs}")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_jacobi_checked_equals_unchecked(qm, precision):
    """Blocks with convergence checks (each ends in two single sweeps; FP32 clears its correction
    buffers) and blocks without (two-sweep passes; FP32 forms the first correction sweep in the
    residual kernel) give bitwise the same V after the same 157 sweeps (3 blocks + 7)."""
    import ctypes
    import torch
    from qmtools import _native as nat
    nx, ny, nz, h, sweeps = 40, 48, 33, 0.2, 157
    rng = numpy.random.default_rng(23)
    rho = torch.tensor(rng.standard_normal(nx * ny * nz), device="cuda")
    g = nat.geom((nx, ny, nz, h, (0.0, 0.0, 0.0)))
    lib, ctx = nat.library(), nat.context()
    outs = []
    for tol in (-1.0, 1e-300):
        _say("jacobi checked vs unchecked", precision, tol)
        dv = torch.zeros_like(rho)
        dt = torch.zeros_like(rho)
        done, delta = ctypes.c_int64(0), ctypes.c_double(0)
        if precision == "fp64":
            args = (nat.QM_PREC_FP64, nat.dptr(rho), nat.dptr(dv), nat.dptr(dt), sweeps, tol, 50)
        else:
            args = (nat.QM_PREC_FP32, nat.dptr(rho), nat.dptr(dv), None, sweeps, tol, 50)
        rc = lib.qm_jacobi(ctx, nat.stream(), ctypes.byref(g), *args, ctypes.byref(done), ctypes.byref(delta))
        assert rc == (nat.QM_OK if tol < 0 else nat.QM_ENOCONV), rc   # 1e-300 is never reached
        assert done.value == sweeps
        outs.append(dv.cpu().numpy())
    numpy.testing.assert_array_equal(outs[1], outs[0])


@pytest.mark.parametrize("boundary", ["zero", "multipole"])
def test_jacobi_fp64_fixed_sweeps(qm, small, boundary):
    from oracle import density as od
    from oracle import jacobi as oj
    from oracle import seed as osd
    mol, grid, sysm = small
    nx, ny, nz, h, o = grid.geometry()
    _say("jacobi fp64", boundary, grid.shape)
    q = qm.QMTools.Compute_density(grid, mol)
    v = qm.QMTools.Compute_hartree_iter(q, mol, tolerance=None, max_sweeps=300, boundary=boundary)
    rho = osd.seed_grid(mol.types, mol.coords, nx, ny, nz, h, o) - q.qube.ravel()
    b = oj.multipole_boundary(rho, nx, ny, nz, h, o) if boundary == "multipole" else None
    want = oj.jacobi(rho, nx, ny, nz, h, 300, bvals=b)
    assert v.sweeps == 300
    assert _rel(v.qube, want) <= 1e-10


def test_jacobi_convergence_loop(qm, small):
    from oracle import jacobi as oj
    from oracle import seed as osd
    mol, grid, _ = small
    nx, ny, nz, h, o = grid.geometry()
    _say("jacobi convergence loop", grid.shape)
    q = qm.QMTools.Compute_density(grid, mol)
    v = qm.QMTools.Compute_hartree_iter(q, mol, tolerance=1e-6)
    rho = osd.seed_grid(mol.types, mol.coords, nx, ny, nz, h, o) - q.qube.ravel()
    want, itr, delta = oj.hartree_iter(rho, nx, ny, nz, h, tolerance=1e-6)
    assert v.sweeps == 2 * (itr - 1) or v.sweeps == 2 * itr
    assert _rel(v.qube, want) <= 1e-9


def test_jacobi_mixed_2000_sweeps(qm, small):
    from oracle import jacobi as oj
    from oracle import seed as osd
    mol, grid, _ = small
    nx, ny, nz, h, o = grid.geometry()
    _say("jacobi mixed 2000", grid.shape)
    q = qm.QMTools.Compute_density(grid, mol)
    v = qm.QMTools.Compute_hartree_iter(q, mol, tolerance=None, max_sweeps=2000, boundary="multipole",
                                        precision="mixed")
    rho = osd.seed_grid(mol.types, mol.coords, nx, ny, nz, h, o) - q.qube.ravel()
    b = oj.multipole_boundary(rho, nx, ny, nz, h, o)
    want = oj.jacobi(rho, nx, ny, nz, h, 2000, bvals=b)
    err = _rel(v.qube, want)
    print(f"mixed Jacobi normwise error after 2000 sweeps: {err:.3e}")
    assert err <= 1e-6


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("mode", ["periodic", "freespace"])
def test_fft_fp64(qm, small, mode):
    from oracle import fft as off
    mol, grid, _ = small
    nx, ny, nz, h, o = grid.geometry()
    q = qm.QMTools.Compute_density(grid, mol)
    _say("fft fp64", mode, grid.shape)
    v = qm.QMTools.Compute_hartree_fft(q, mol, sigma=0.2, cutoff=6, mode=mode)
    want = off.hartree_fft(q.qube.ravel(), mol.types, mol.coords, nx, ny, nz, h, o, 0.2, 6, mode=mode)
    assert v.shape == (nx, ny, nz)
    assert _rel(v, want) <= 1e-10


@pytest.mark.parametrize("shape", [(24, 40, 16), (88, 16, 24), (152, 16, 24)])
def test_fft_mixed_radix_lengths(qm, small, shape):
    """Lengths with factors 3, 5 (in-place radices) and 11, 19 (generic radix path, periodic); in free
    space 2n = 176 and 304 are padded to 180 and 315 (no prime factor above 7), against the oracle's
    2n-point convolution."""
    from oracle import fft as off
    mol, _, _ = small
    nx, ny, nz = shape
    grid = qm.Grid((1, nx, ny, nz), 0.25)
    from qmtools.grid import make_float3
    c = mol.coords.mean(axis=0)
    grid.origin = make_float3(*(c - 0.5 * numpy.array(shape) * float(grid.step)))
    rng = numpy.random.default_rng(7)
    grid.LoadFromNPY(rng.uniform(0.0, 0.01, grid.shape))
    nx, ny, nz, h, o = grid.geometry()
    for mode in ("periodic", "freespace"):
        want = off.hartree_fft(grid.qube.ravel(), mol.types, mol.coords, nx, ny, nz, h, o, mode=mode)
        # FP64 to the oracle's precision; FP32 (complex64 passes) to the mixed path's 1e-6
        for precision, tol in (("fp64", 1e-10), ("mixed", 1e-6)):
            _say("fft radix", shape, mode, precision)
            v = qm.QMTools.Compute_hartree_fft(grid, mol, mode=mode, precision=precision)
            assert _rel(v, want) <= tol, (mode, precision)


def test_fft_freespace_green_cache_across_steps(qm, small):
    """Grids of one shape and different steps share the cached Green's-function transform (kept at unit
    step, scaled by h^2 at use): each step's free-space potential equals the oracle's (FP64 1e-10, FP32
    1e-6), and the context's resident bytes do not grow after the first step."""
    import ctypes
    from oracle import fft as off
    from qmtools import _native as nat
    from qmtools.grid import make_float3
    mol, _, _ = small
    shape = (40, 32, 24)
    c = mol.coords.mean(axis=0)
    rng = numpy.random.default_rng(11)
    dens = rng.uniform(0.0, 0.01, (1,) + shape)
    resident = []
    for step in (0.25, 0.31, 0.4):
        grid = qm.Grid((1,) + shape, step)
        grid.origin = make_float3(*(c - 0.5 * numpy.array(shape) * float(grid.step)))
        grid.LoadFromNPY(dens)
        nx, ny, nz, h, o = grid.geometry()
        want = off.hartree_fft(grid.qube.ravel(), mol.types, mol.coords, nx, ny, nz, h, o, mode="freespace")
        for precision, tol in (("fp64", 1e-10), ("mixed", 1e-6)):
            _say("fft green cache", shape, step, precision)
            v = qm.QMTools.Compute_hartree_fft(grid, mol, mode="freespace", precision=precision)
            assert _rel(v, want) <= tol, (step, precision)
        b = ctypes.c_int64(0)
        nat.check(nat.library().qm_ctx_memory(nat.context(), ctypes.byref(b)), nat.context())
        resident.append(int(b.value))
    assert resident[1] == resident[0] and resident[2] == resident[0], resident


def test_fft_fp32(qm, small):
    from oracle import fft as off
    mol, grid, _ = small
    nx, ny, nz, h, o = grid.geometry()
    q = qm.QMTools.Compute_density(grid, mol)
    _say("fft fp32 freespace", grid.shape)
    v = qm.QMTools.Compute_hartree_fft(q, mol, mode="freespace", precision="mixed")
    want = off.hartree_fft(q.qube.ravel(), mol.types, mol.coords, nx, ny, nz, h, o, mode="freespace")
    err = _rel(v, want)
    print(f"FP32 free-space FFT normwise error {err:.3e}")
    assert err <= 1e-6


# ---------------------------------------------------------------- GTO
def test_gto_points(qm, small):
    from oracle import gto as og
    from qmtools import synth
    mol, grid, sysm = small
    pts = synth.uniform_points(grid, 48, 5)
    # include a nucleus position and a point very close to one (T -> 0 and r < 1e-8 paths)
    pts = numpy.vstack([pts, mol.coords[:1], mol.coords[1:2] + 1e-9])
    _say("gto points", pts.shape)
    v = qm.QMTools.Compute_hartree_points(mol, pts)
    want = og.potential(sysm, pts)
    assert _rel(v, want) <= 1e-10


def test_gto_grid(qm, obasis):
    from oracle import gto as og
    from qmtools import synth
    basis = qm.BasisSet(BASIS)
    mol = synth.synthetic_molecule(3, 5, basis)
    grid = synth.centred_grid(mol, 8, step_bohr=0.9, pad_ang=1.0)
    nx, ny, nz, h, o = grid.geometry()
    sysm = _system(mol, obasis)
    _say("gto grid", grid.shape)
    vg = qm.Grid.emptyAs(grid)
    qm.QMTools.Compute_hartree(vg, mol)
    want = og.potential_grid(sysm, nx, ny, nz, h, o)
    assert _rel(vg.qube, want) <= 1e-10


# ---------------------------------------------------------------- contraction tile paths
@pytest.fixture(scope="module", params=[56, 70])
def large(request, qm, obasis):
    """nbf_pad 576 (one column pass of 5 tiles per wave) and 768 (two passes)."""
    from qmtools import synth
    basis = qm.BasisSet(BASIS)
    mol = synth.synthetic_molecule(request.param, 40 + request.param, basis)
    grid = synth.centred_grid(mol, 16, step_bohr=1.6, pad_ang=2.0)
    return mol, grid, _system(mol, obasis)


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_density_fp64_large_basis(qm, large, variant):
    from oracle import density as od
    from qmtools import autotune
    mol, grid, sysm = large
    nx, ny, nz, h, o = grid.geometry()
    _say("density fp64 large basis", grid.shape, "nbf", mol.nbf, "variant", variant)
    old = autotune.get("f64_variant")
    autotune.set("f64_variant", variant)
    try:
        q = qm.QMTools.Compute_density(grid, mol)
    finally:
        autotune.set("f64_variant", old)
    want, raw = od.density_grid(sysm, nx, ny, nz, h, o)
    assert _rel(q.qube, want) <= 1e-10
    assert q.raw_total == pytest.approx(raw, rel=1e-10)


@pytest.mark.parametrize("rows", [0, 1, 2, 3, 4, 5, 64, 128])
def test_density_f16_large_basis_tiles(qm, large, rows):
    from oracle import density as od
    from qmtools import autotune
    mol, grid, sysm = large
    nx, ny, nz, h, o = grid.geometry()
    _say("density f16 large basis", grid.shape, "nbf", mol.nbf, "rows", rows)
    old = autotune.get("split_variant")
    autotune.set("split_variant", rows)
    try:
        q = qm.QMTools.Compute_density(grid, mol, precision="f16")
    finally:
        autotune.set("split_variant", old)
    want, raw = od.density_grid(sysm, nx, ny, nz, h, o)
    err = _rel(q.qube, want)
    print(f"f16 rows {rows} nbf {mol.nbf}: normwise error {err:.3e}")
    assert err <= 1e-6
    assert abs(q.raw_total - raw) / raw <= 1e-6


def test_autotune_cache(qm, small, tmp_path):
    import json
    from qmtools import autotune
    mol, grid, _ = small
    path = str(tmp_path / "tune.json")
    _say("autotune", grid.shape)
    old = autotune.get("split_variant")
    e = autotune.autotune_density(mol, grid, "mixed", path=path)
    assert e["split_variant"] in (0, 1, 2, 3, 4, 5, 64, 128) and e["ms"] > 0
    table = json.load(open(path))
    assert autotune.key_for(mol, "mixed") in table
    autotune.set("split_variant", 64)
    again = autotune.autotune_density(mol, grid, "mixed", path=path)   # served from the cache
    assert again == e and autotune.get("split_variant") == e["split_variant"]
    autotune.set("split_variant", old)


def test_seed_shared_nodes_deterministic(qm):
    """Nuclei closer than a grid step share seed nodes: the later atom wins, as in the oracle."""
    from oracle import seed as osd
    basis = qm.BasisSet(BASIS)
    xyz = [[6, 0.0, 0.0, 0.0], [1, 0.4, 0.3, 0.2], [1, -0.3, 0.35, -0.25], [8, 1.2, 0.1, 0.0]]
    nbf = 14 + 5 + 5 + 14
    mol = qm.Molecule(xyz, numpy.eye(nbf) * 0.1, basis)
    grid = qm.Grid((1, 16, 16, 16), 1.0)
    from qmtools.grid import make_float3
    grid.origin = make_float3(-8.0, -8.0, -8.0)
    nx, ny, nz, h, o = grid.geometry()
    want = osd.seed_grid(mol.types, mol.coords, nx, ny, nz, h, o)
    _say("seed shared nodes")
    for _ in range(3):
        s = qm.QMTools.Compute_qseed(grid, mol)
        numpy.testing.assert_allclose(s.qube.ravel(), want, rtol=1e-14, atol=1e-15)


# ---------------------------------------------------------------- Translate with resident tables
def test_translate_reuses_resident_tables(qm, obasis):
    """Molecule.Translate (reference qmtools.py:182-188) moves the nuclei on the device: density, GTO
    potential at points, free-space FFT potential and the batched density all follow the moved
    nuclei (oracle at the new coordinates, 1e-10), the GTO table is rebuilt once, and neither the
    density matrix nor the basis tables are uploaded again."""
    from oracle import density as od
    from oracle import fft as off
    from oracle import gto as og
    from qmtools import batch, synth
    from qmtools.qmtools import device_molecule
    basis = qm.BasisSet(BASIS)
    mol = synth.synthetic_molecule(7, 33, basis)
    grid = synth.centred_grid(mol, 32, step_bohr=0.35, pad_ang=2.5)
    pts = synth.uniform_points(grid, 40, 8)
    nx, ny, nz, h, o = grid.geometry()
    _say("translate: before")
    qm.QMTools.Compute_density(grid, mol)
    qm.QMTools.Compute_hartree_points(mol, pts)
    dev = device_molecule(mol)
    before = dev.stats()
    assert before == {"p_uploads": 1, "coord_updates": 0, "gto_builds": 1}
    shift = (0.31, -0.47, 0.22)    # bohr, as the reference's Translate
    moved = mol.coords + numpy.array(shift)
    mol.Translate(*shift)
    numpy.testing.assert_array_equal(mol.coords, moved)
    assert device_molecule(mol) is dev
    sysm = _system(mol, obasis)
    _say("translate: density, GTO, FFT after the move")
    q = qm.QMTools.Compute_density(grid, mol)
    want, raw = od.density_grid(sysm, nx, ny, nz, h, o)
    assert _rel(q.qube, want) <= 1e-10
    v = qm.QMTools.Compute_hartree_points(mol, pts)
    assert _rel(v, og.potential(sysm, pts)) <= 1e-10
    vf = qm.QMTools.Compute_hartree_fft(q, mol, mode="freespace")
    vfw = off.hartree_fft(q.qube.ravel(), mol.types, moved, nx, ny, nz, h, o, mode="freespace")
    assert _rel(vf, vfw) <= 1e-10
    qb = batch.density_batch([mol], [grid], "fp64")[0]
    numpy.testing.assert_array_equal(qb.cpu().numpy(), q.qube.ravel())
    after = dev.stats()
    assert after == {"p_uploads": 1, "coord_updates": 1, "gto_builds": 2}
    # a second GTO call reuses the rebuilt table
    qm.QMTools.Compute_hartree_points(mol, pts)
    assert dev.stats()["gto_builds"] == 2


@pytest.mark.parametrize("subgrid", [1, 2])
def test_f64_tile_skip_bitwise(qm, subgrid):
    """The FP64 LDS-DMA contraction (single and batched) skips the products of 16-function groups whose
    basis values are zero on a whole row tile (beyond every shell's cutoff, by the tile's bounding box):
    they add exact zeros, so the density equals the one with every product formed, bit for bit.  The workload (60
    atoms in a 29 bohr box) has such groups: checked on the host for the tiles' bounding boxes."""
    from qmtools import autotune, synth
    basis = qm.BasisSet(BASIS)
    mol = synth.synthetic_molecule(60, 2, basis)
    grid = synth.centred_grid(mol, 64, step_bohr=0.45, pad_ang=3.0)
    # host estimate: fraction of (tile, shell) pairs beyond the cutoff for 128-point tiles
    nx, ny, nz, h, o = grid.geometry()
    amin = numpy.array([mol.prim_alpha[mol.shell_poff[s]:mol.shell_poff[s] + mol.shell_nprim[s]].min()
                        for s in range(mol.nshell)])
    centres = mol.coords[numpy.asarray(mol.shell_atom)]
    idx = numpy.arange(nx * ny * nz)
    pts = numpy.stack([o[0] + (idx % nx + 0.5) * h, o[1] + ((idx // nx) % ny + 0.5) * h,
                       o[2] + (idx // (nx * ny) + 0.5) * h], 1).reshape(-1, 128, 3)
    lo, hi = pts.min(1), pts.max(1)
    gap = numpy.maximum(0.0, numpy.maximum(lo[:, None, :] - centres[None], centres[None] - hi[:, None, :]))
    zero = ((gap ** 2).sum(-1) > 50.0 / amin[None, :]).mean()
    assert zero > 0.2, zero
    old = autotune.get("tile_skip"), autotune.get("f64_variant")
    autotune.set("f64_variant", 4)
    out = {}
    try:
        for skip in (1, 0):
            autotune.set("tile_skip", skip)
            _say("f64 tile skip", skip, "subgrid", subgrid, f"zero tile-shell pairs {zero:.2f}")
            out[skip] = qm.QMTools.Compute_density(grid, mol, subgrid=subgrid).qube.copy()
            if subgrid == 1:   # the batched kernel, with a second molecule in the same round
                from qmtools import batch
                other = synth.synthetic_molecule(9, 4, basis)
                qs = batch.density_batch([other, mol], [synth.centred_grid(other, 40, step_bohr=0.5), grid], "fp64")
                out[(skip, "batch")] = qs[1].cpu().numpy()
    finally:
        autotune.set("tile_skip", old[0])
        autotune.set("f64_variant", old[1])
    assert numpy.abs(out[0]).max() > 0
    numpy.testing.assert_array_equal(out[1], out[0])
    if subgrid == 1:
        numpy.testing.assert_array_equal(out[(1, "batch")], out[(0, "batch")])
        numpy.testing.assert_array_equal(out[(1, "batch")], out[1].ravel())


def test_gto_table_rebuild_is_deterministic(qm, obasis):
    """The GTO distribution table is built by several host threads, one shell at a time, and
    concatenated in shell order: rebuilding it for the same coordinates (a coordinate update to the
    same positions) gives bitwise the same potentials, and they match the oracle (1e-10)."""
    from oracle import gto as og
    from qmtools import synth
    from qmtools.qmtools import device_molecule
    basis = qm.BasisSet(BASIS)
    mol = synth.synthetic_molecule(30, 17, basis)
    grid = synth.centred_grid(mol, 32, step_bohr=0.5, pad_ang=2.5)
    pts = synth.uniform_points(grid, 200, 5)
    _say("gto rebuild", mol.natoms, mol.nbf)
    v1 = qm.QMTools.Compute_hartree_points(mol, pts)
    dev = device_molecule(mol)
    builds = dev.stats()["gto_builds"]
    vs = []
    for _ in range(3):
        dev.update_coords(mol.coords)
        vs.append(qm.QMTools.Compute_hartree_points(mol, pts))
    assert dev.stats()["gto_builds"] == builds + 3
    for v in vs:
        numpy.testing.assert_array_equal(v, v1)
    assert _rel(v1, og.potential(_system(mol, obasis), pts)) <= 1e-10


@pytest.mark.parametrize("natoms,seed,n", [(3, 7, 24), (12, 4, 16), (56, 96, 16), (120, 5, 8)])
@pytest.mark.parametrize("prec", ["f16", "bf16"])
def test_density_fused_equals_two_kernel_path(qm, natoms, seed, n, prec):
    """The fused kernel (basis values evaluated into LDS) evaluates the basis values with the roundings
    of the basis-value kernel and forms the same MFMA products in the same k order as the two-kernel
    path (variant 3); the row-dot terms are grouped and rounded differently (measured differences
    4e-10 and below, against 1e-7 for the split itself), so the gate is 1e-9: a wrong or missing
    basis value or product would exceed it.  nbf 33 .. 1200: one to four column groups, padding columns
    in group 0, and more than 1024 basis functions."""
    import ctypes
    import torch
    from qmtools import _native as nat, autotune, synth
    from qmtools.qmtools import device_molecule
    mol = synth.synthetic_molecule(natoms, seed, qm.BasisSet(BASIS))
    grid = synth.centred_grid(mol, n, step_bohr=0.6, pad_ang=2.0)
    _say("fused vs two-kernel", prec, natoms, "nbf", mol.nbf, grid.shape)
    g = nat.geom(grid)
    code = nat.QM_PREC_F16X3 if prec == "f16" else nat.QM_PREC_BF16X3
    outs = []
    old = autotune.get("split_variant")
    try:
        for v in (3, 4):   # FP64 output: the FP32 rounding of the drop-in grid would hide the comparison
            autotune.set("split_variant", v)
            o = torch.empty(int(grid.npts), dtype=torch.float64, device="cuda")
            nat.call("qm_density", device_molecule(mol).handle, nat.stream(), ctypes.byref(g), 1, code, nat.QM_F64,
                     nat.dptr(o), 0, None)
            outs.append(o.cpu().numpy())
    finally:
        autotune.set("split_variant", old)
    err = _rel(outs[1], outs[0])
    print(f"fused vs two-kernel {prec} nbf {mol.nbf}: {err:.2e}")
    assert err <= 1e-9
