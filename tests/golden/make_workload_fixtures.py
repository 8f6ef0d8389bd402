"""Writes the full-size workload fixtures from the FP64 NumPy oracle (run once, on the CPU).

    python tests/golden/make_workload_fixtures.py [--workers 8] [--only mol12|mol60]

mol12_grid96.npz  -- seeded 12-atom molecule, 96^3 grid (`synth.molecule_12_grid_96`):
    normalised density at 4096 sampled voxels, the un-normalised total, and V after
    2000 Jacobi sweeps with multipole boundaries (oracle closed form) at the same voxels
    plus sum(V) and sum|V| over the grid.
mol60_grid256.npz -- seeded 60-atom molecule, 256^3 grid (`synth.molecule_60_grid_256`):
    un-normalised density at 4096 sampled voxels and its total over all 16.8M voxels,
    the GTO potential at the first 64 of the 10^4 sampled points, and the free-space FFT
    potential (sigma 0.2 A, cutoff 6) at 64 sampled voxels by direct summation over the
    whole grid (the convolution the 2x-padded FFT computes).

Both record the molecule (types, coordinates, a density-matrix fingerprint) and the grid
geometry, so the tests can check that `synth` rebuilds the same inputs.  The oracle's full
256^3 density takes a few minutes with 8 worker processes.
"""
import argparse
import math
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for _p in (ROOT, os.path.join(ROOT, "qmtools-pycuda_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)
os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")

import numpy  # noqa: E402

BASIS = os.path.join(ROOT, "qmtools-pycuda_amd", "qmtools", "data", "cc-pvdz.bin")
NSAMPLE = 4096
_SYS = None


def _system(mol):
    from oracle.basisfile import read_basis
    from oracle.system import make_system
    return make_system(mol.types, mol.coords, mol.dm, read_basis(BASIS))


def _init(mol_args):
    global _SYS
    from qmtools import synth
    mol = synth.synthetic_molecule(*mol_args)
    _SYS = _system(mol)


def _rho_slab(args):
    """Un-normalised density 2 h^3 rho of z planes [k0, k1) (flat x-fastest)."""
    from oracle import density as od
    nx, ny, k0, k1, h, o = args
    oz = (o[0], o[1], o[2] + k0 * h)
    pts = od.grid_points(nx, ny, k1 - k0, h, oz)
    return k0, 2.0 * od.rho_on_points(_SYS, pts, chunk=16384) * h ** 3


def _gto_part(args):
    from oracle import gto as og
    pts, shells = args
    return og.potential(_SYS, pts, nuclei=False, shells_a=shells)


def _full_density(pool, grid, planes=2):
    nx, ny, nz, h, o = grid.geometry()
    out = numpy.empty(nx * ny * nz)
    jobs = [(nx, ny, k, min(k + planes, nz), h, o) for k in range(0, nz, planes)]
    t0 = time.time()
    for n, (k0, q) in enumerate(pool.imap_unordered(_rho_slab, jobs)):
        out[k0 * nx * ny:k0 * nx * ny + q.size] = q
        if n % 16 == 0:
            print(f"  density planes {n * planes}/{nz} ({time.time() - t0:.0f} s)", flush=True)
    return out


def _sample(n, seed):
    return numpy.sort(numpy.random.default_rng(seed).choice(n, size=NSAMPLE, replace=False))


def _header(mol, grid):
    from qmtools import synth
    nx, ny, nz, h, o = grid.geometry()
    return {"types": mol.types, "coords": mol.coords, "dm_probe": synth.dm_probe(mol.dm),
            "shape": numpy.array([nx, ny, nz]), "step": numpy.float64(h), "origin": numpy.array(o)}


def make_mol12(workers):
    from oracle import jacobi as oj
    from oracle import seed as osd
    from qmtools import synth
    import qmtools
    mol, grid = synth.molecule_12_grid_96(qmtools.BasisSet(BASIS))
    nx, ny, nz, h, o = grid.geometry()
    with mp.Pool(workers, initializer=_init, initargs=((12, 1),)) as pool:
        raw = _full_density(pool, grid, planes=4)
    total = math.fsum(raw)
    q = raw * (float(mol.qtot) / total)
    rho = osd.seed_grid(mol.types, mol.coords, nx, ny, nz, h, o) - q
    b = oj.multipole_boundary(rho, nx, ny, nz, h, o)
    v = oj.jacobi_closed_form(rho, nx, ny, nz, h, 2000, bvals=b)
    idx = _sample(nx * ny * nz, 12)
    out = _header(mol, grid)
    out.update({"idx": idx, "q": q[idx], "raw_total": numpy.float64(total), "sweeps": numpy.int64(2000),
                "v_jacobi": v[idx], "v_sum": numpy.float64(math.fsum(v)),
                "v_abs_sum": numpy.float64(math.fsum(numpy.abs(v)))})
    numpy.savez_compressed(os.path.join(HERE, "mol12_grid96.npz"), **out)
    print(f"mol12_grid96: nbf {mol.nbf}, raw total {total:.15g}, V sum {out['v_sum']:.15g}")


def make_mol60(workers):
    from oracle import fft as off
    from qmtools import synth
    import qmtools
    mol, grid, pts = synth.molecule_60_grid_256(qmtools.BasisSet(BASIS))
    nx, ny, nz, h, o = grid.geometry()
    out = _header(mol, grid)
    with mp.Pool(workers, initializer=_init, initargs=((60, 2),)) as pool:
        # GTO potential at 64 points: the outer shells dealt round-robin to the workers
        gp = pts[:64]
        ns = int(mol.nshell)
        t0 = time.time()
        parts = pool.map(_gto_part, [(gp, list(range(w, ns, workers))) for w in range(workers)])
        from oracle import gto as og
        vg = numpy.sum(parts, axis=0) + og.potential(_system(mol), gp, nuclei=True, shells_a=[])
        print(f"  GTO at 64 points: {time.time() - t0:.0f} s", flush=True)
        raw = _full_density(pool, grid)
    total = math.fsum(raw)
    q = raw * (float(mol.qtot) / total)
    idx = _sample(nx * ny * nz, 60)
    # free-space FFT potential by direct summation over the whole grid at 64 voxels
    rhs = off.total_charge_rhs(q, mol.types, mol.coords, nx, ny, nz, h, o, 0.2, 6.0)   # C order [i, j, k]
    h_ang = h / off.ANG2BOR
    qc = ((-rhs * off.EPS0) * h_ang ** 3 / (4.0 * math.pi * off.EPS0)).ravel()
    del rhs
    tgt = _sample(nx * ny * nz, 61)[::64]
    ti, tj, tk = tgt % nx, (tgt // nx) % ny, tgt // (nx * ny)
    ax = [numpy.arange(n, dtype=numpy.float64) for n in (nx, ny, nz)]
    c0 = off.cube_self_potential() / h_ang
    vf = numpy.empty(tgt.size)
    for n in range(tgt.size):
        dx2 = ((ax[0] - ti[n]) * h_ang) ** 2
        dy2 = ((ax[1] - tj[n]) * h_ang) ** 2
        dz2 = ((ax[2] - tk[n]) * h_ang) ** 2
        d = numpy.sqrt(dx2[:, None, None] + dy2[None, :, None] + dz2[None, None, :]).ravel()
        d[(ti[n] * ny + tj[n]) * nz + tk[n]] = 1.0
        g = 1.0 / d
        g[(ti[n] * ny + tj[n]) * nz + tk[n]] = c0
        vf[n] = math.fsum(g * qc)
    out.update({"idx": idx, "raw": raw[idx], "raw_total": numpy.float64(total),
                "gto_points": gp, "v_gto": vg, "fft_idx": tgt, "v_fft": vf,
                "sigma": numpy.float64(0.2), "cutoff": numpy.float64(6.0)})
    numpy.savez_compressed(os.path.join(HERE, "mol60_grid256.npz"), **out)
    print(f"mol60_grid256: nbf {mol.nbf}, raw total {total:.15g}")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--only", choices=["mol12", "mol60"])
    a = ap.parse_args()
    if a.only in (None, "mol12"):
        make_mol12(a.workers)
    if a.only in (None, "mol60"):
        make_mol60(a.workers)
