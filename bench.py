"""Grid hot-path benchmark: density + Hartree on one MI355X.

Workload (one "step"): seeded 60-atom C/N/O/H molecule in cc-pVDZ (606 basis
functions), 256^3 cell-centred grid with the molecule centred and 3 Angstrom of
padding (step 0.112 bohr: 0.1 bohr does not fit the molecule in 256 points);
  1. density  rho = diag(Phi P Phi^T) on all 256^3 points, normalised to sum Z;
  2. Hartree potential by FFT Poisson with free-space boundaries (2x zero-padded,
     512^3 complex transforms) of electrons + Gaussian nuclei;
  3. analytic GTO-pair + nuclear potential at 10^4 uniformly sampled points.
All inputs (basis tables, P, points, grids) are resident in HBM before timing.

Default path: plain FP64 (FP64 MFMA contraction, complex128 FFT, FP64 GTO).
`--precision mixed` measures the mixed path (FP32 basis values, FP16 hi/lo MFMA
contraction, complex64 FFT; GTO stays FP64).  Each run also measures the other
path once and reports its throughput and its error against the FP64 GPU result,
which itself is checked against the FP64 NumPy oracle on a sample of voxels.

Prints one JSON line (see the keys at the end of main()).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "qmtools-pycuda_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per contraction launch measured with rocprofv3 FETCH_SIZE (x2 calibration) and
# WRITE_SIZE (x1) for this workload; source: profiles/mol60_grid256_pmc_traffic.md
# (precision, nbf_pad, variant): (points per launch, HBM bytes per launch)
MEASURED_TRAFFIC = {("fp64", 640, "tiles_dma"): (838784, 3499848 * 1024 * 2 + 32765 * 1024),
                    ("fp64", 640, "tiles"): (838784, 6138183 * 1024 * 2 + 32765 * 1024),
                    ("fp64", 640, "rows"): (838784, 3870778 * 1024 * 2 + 65536 * 1024),
                    ("fp64", 640, "lds"): (838784, 4444220 * 1024 * 2 + 190037 * 1024),
                    ("fp64", 640, "direct"): (838784, 7883104 * 1024 * 2 + 6553 * 1024),
                    ("mixed", 640, 3): (1677696, 1995799 * 1024 * 2 + 65535 * 1024),
                    ("mixed", 640, 1): (1677696, 2246457 * 1024 * 2 + 65535 * 1024),
                    ("mixed", 640, 0): (1677696, 2397406 * 1024 * 2 + 65535 * 1024),
                    ("mixed", 640, 2): (1677696, 2526254 * 1024 * 2 + 65535 * 1024),
                    ("mixed", 640, 64): (1677696, 3798200 * 1024 * 2 + 530834 * 1024),
                    ("mixed", 640, 128): (1677824, 6713930 * 1024 * 2 + 8984396 * 1024)}
FP64_MFMA_PEAK_TFLOPS = 78.6
F16_MFMA_PEAK_TFLOPS = 2500.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


class Workload:
    def __init__(self, natoms=60, seed=2, n=256, npoints=10000, precision="fp64"):
        import torch

        import qmtools
        from qmtools import _device, _native as nat, synth
        from qmtools.qmtools import device_molecule
        self.torch = torch
        self.nat = nat
        self.basis = qmtools.BasisSet(os.path.join(PKG, "qmtools", "data", "cc-pvdz.bin"))
        self.mol = synth.synthetic_molecule(natoms, seed, self.basis)
        self.grid = synth.centred_grid(self.mol, n, step_bohr=0.1, pad_ang=3.0)
        self.n = n
        self.N = n ** 3
        self.pts_host = synth.uniform_points(self.grid, npoints, seed + 1)
        self.npoints = npoints
        self.precision = precision
        dev = _device.require()
        self.dm = device_molecule(self.mol)
        self.geom = nat.geom(self.grid)
        self.q64 = torch.empty(self.N, dtype=torch.float64, device=dev)
        self.v64 = torch.empty(self.N, dtype=torch.float64, device=dev)
        self.q32 = torch.empty(self.N, dtype=torch.float32, device=dev)
        self.v32 = torch.empty(self.N, dtype=torch.float32, device=dev)
        self.pts = torch.from_numpy(self.pts_host.ravel().copy()).to(dev)
        self.vpts = torch.empty(npoints, dtype=torch.float64, device=dev)
        self.stream = torch.cuda.current_stream(dev)

    def step(self, precision=None):
        nat = self.nat
        p = self.precision if precision is None else precision
        s = ctypes.c_void_p(self.stream.cuda_stream)
        g = ctypes.byref(self.geom)
        if p == "fp64":
            nat.call("qm_density", self.dm.handle, s, g, 1, nat.QM_PREC_FP64, nat.QM_F64, nat.dptr(self.q64), 1,
                     None)
            nat.call("qm_hartree_fft", self.dm.handle, s, g, nat.QM_PREC_FP64, nat.QM_F64, nat.dptr(self.q64),
                     0.2, 6.0, nat.QM_FFT_FREESPACE, nat.dptr(self.v64))
        else:
            nat.call("qm_density", self.dm.handle, s, g, 1, nat.QM_PREC_F16X3, nat.QM_F32, nat.dptr(self.q32), 1,
                     None)
            nat.call("qm_hartree_fft", self.dm.handle, s, g, nat.QM_PREC_FP32, nat.QM_F32, nat.dptr(self.q32),
                     0.2, 6.0, nat.QM_FFT_FREESPACE, nat.dptr(self.v32))
        nat.call("qm_hartree_gto_points", self.dm.handle, s, nat.dptr(self.pts), self.npoints, nat.dptr(self.vpts))

    def timed(self, steps, warmup, precision=None, profile=False):
        torch = self.torch
        nat = self.nat
        for _ in range(warmup):
            self.step(precision)
        torch.cuda.synchronize()
        ctx = nat.context()
        if profile:
            nat.library().qm_profile(ctx, 1)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a, b in ev:
            a.record(self.stream)
            self.step(precision)
            b.record(self.stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if profile:
            nat.library().qm_profile(ctx, 0)
        lat = sorted(a.elapsed_time(b) for a, b in ev)
        return wall, lat

    def kernel_ms(self, kid):
        tot = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        self.nat.check(self.nat.library().qm_profile_read(self.nat.context(), kid, ctypes.byref(tot),
                                                          ctypes.byref(n)), self.nat.context())
        return tot.value, n.value


def pct(vals, q):
    if not vals:
        return None
    k = min(len(vals) - 1, max(0, int(math.ceil(q / 100.0 * len(vals))) - 1))
    return vals[k]


def oracle_sample_check(w, nsample=4096, seed=9):
    """Density at a random sample of voxels vs the FP64 NumPy oracle (normalised with the GPU's sum)."""
    from oracle import density as od
    from oracle.basisfile import read_basis
    from oracle.system import make_system
    sysm = make_system(w.mol.types, w.mol.coords, w.mol.dm, read_basis(os.path.join(PKG, "qmtools", "data",
                                                                                     "cc-pvdz.bin")))
    rng = numpy.random.default_rng(seed)
    nx, ny, nz, h, o = w.grid.geometry()
    idx = rng.choice(w.N, size=nsample, replace=False)
    i, j, k = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    pts = numpy.asarray(o) + (numpy.stack([i, j, k], 1) + 0.5) * h
    raw = 2.0 * od.rho_on_points(sysm, pts) * h ** 3
    q = w.q64.cpu().numpy()
    gpu_total = float(numpy.sum(q))   # = sum Z after normalisation
    # reproduce the normalisation with the GPU's un-normalised total
    nat = w.nat
    tot = ctypes.c_double(0.0)
    nat.call("qm_density", w.dm.handle, ctypes.c_void_p(w.stream.cuda_stream), ctypes.byref(w.geom), 1,
             nat.QM_PREC_FP64, nat.QM_F64, nat.dptr(w.v64), 0, ctypes.byref(tot))
    unnorm = w.v64.cpu().numpy()[idx]
    err = float(numpy.abs(unnorm - raw).max() / numpy.abs(raw).max())
    return err, gpu_total, tot.value


def cpu_baseline(w):
    """FP64 NumPy oracle on the host cores, on bounded samples of the same workload, per component."""
    from oracle import density as od
    from oracle import fft as off
    from oracle import gto as og
    from oracle.basisfile import read_basis
    from oracle.system import make_system
    cores = len(os.sched_getaffinity(0))
    sysm = make_system(w.mol.types, w.mol.coords, w.mol.dm, read_basis(os.path.join(PKG, "qmtools", "data",
                                                                                     "cc-pvdz.bin")))
    nx, ny, nz, h, o = w.grid.geometry()
    # density: random voxels of the grid, scaled by the voxel count
    ns = 32768
    rng = numpy.random.default_rng(5)
    idx = rng.choice(w.N, size=ns, replace=False)
    i, j, k = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    pts = numpy.asarray(o) + (numpy.stack([i, j, k], 1) + 0.5) * h
    t0 = time.perf_counter()
    od.rho_on_points(sysm, pts, chunk=8192)
    t_dens = (time.perf_counter() - t0) * w.N / ns
    # free-space FFT: the same solve on a 64^3 grid (128^3 padded), scaled by M log M to 512^3
    m = 64
    qs = numpy.random.default_rng(6).uniform(0, 1e-3, m ** 3)
    t0 = time.perf_counter()
    off.hartree_fft(qs, w.mol.types, w.mol.coords, m, m, m, h * nx / m, o, mode="freespace")
    t_small = time.perf_counter() - t0
    M_big, M_small = (2 * nx) ** 3, (2 * m) ** 3
    t_fft = t_small * (M_big * math.log2(M_big)) / (M_small * math.log2(M_small))
    # GTO: the oracle's shell-pair loop over all 10^4 points (it is vectorised over points), for every
    # 16th outer shell, scaled by the primitive pairs of the whole pair sum
    npr = [len(a) for (_, _, a, _) in sysm.shells]
    pair_cost = [npr[sa] * sum(npr[:sa + 1]) for sa in range(len(npr))]
    sub = list(range(0, len(npr), 16))
    t0 = time.perf_counter()
    og.potential(sysm, w.pts_host, nuclei=False, shells_a=sub)
    t_gto = (time.perf_counter() - t0) * sum(pair_cost) / sum(pair_cost[sa] for sa in sub)
    total = t_dens + t_fft + t_gto
    return {
        "value": w.N / total,
        "unit": "grid pts/s",
        "cores": cores,
        "kind": "port",
        "seconds_per_step": {"density": t_dens, "fft_poisson": t_fft, "gto_points": t_gto},
        "sample": (f"NumPy FP64 oracle (OpenBLAS, {cores} threads): density on {ns} random voxels of 256^3 "
                   f"scaled x{w.N // ns}; free-space FFT Poisson timed on 64^3 (128^3 padded) and scaled by "
                   f"M log M to 512^3; GTO potential at all {w.npoints} points for every 16th outer shell "
                   f"({len(sub)} of {len(npr)}) scaled by primitive pairs; seconds per step (extrapolated): "
                   f"density {t_dens:.1f}, FFT {t_fft:.1f}, GTO {t_gto:.1f}"),
    }


def cpu_small_workload():
    """The 12-atom, 96^3 workload run in full by the NumPy oracle on the host: FP64 density, multipole
    boundary, 2000 Jacobi sweeps (the oracle's sweep loop, as the GPU does them)."""
    import qmtools
    from oracle import density as od
    from oracle import jacobi as oj
    from oracle import seed as osd
    from oracle.basisfile import read_basis
    from oracle.system import make_system
    from qmtools import synth
    mol, grid = synth.molecule_12_grid_96(qmtools.BasisSet(os.path.join(PKG, "qmtools", "data", "cc-pvdz.bin")))
    nx, ny, nz, h, o = grid.geometry()
    sysm = make_system(mol.types, mol.coords, mol.dm, read_basis(os.path.join(PKG, "qmtools", "data",
                                                                                     "cc-pvdz.bin")))
    t0 = time.perf_counter()
    q, _ = od.density_grid(sysm, nx, ny, nz, h, o)
    t1 = time.perf_counter()
    rho = osd.seed_grid(mol.types, mol.coords, nx, ny, nz, h, o) - q
    v = oj.jacobi(rho, nx, ny, nz, h, 2000, bvals=oj.multipole_boundary(rho, nx, ny, nz, h, o))
    t2 = time.perf_counter()
    npts = nx * ny * nz
    return {"workload": "12-atom cc-pVDZ (123 bf), 96^3: density + multipole boundary + 2000 Jacobi sweeps",
            "cores": len(os.sched_getaffinity(0)), "kind": "port", "seconds": {"density": t1 - t0, "jacobi": t2 - t1},
            "grid_pts_per_s": npts / (t2 - t0), "v_checksum": float(numpy.sum(v))}


def gpu_small_workload():
    """The same 12-atom, 96^3 workload on the GPU through the drop-in API (FP64 and mixed)."""
    import torch
    import qmtools
    from qmtools import synth
    mol, grid = synth.molecule_12_grid_96(qmtools.BasisSet(os.path.join(PKG, "qmtools", "data", "cc-pvdz.bin")))
    out = {}
    for prec in ("fp64", "mixed"):
        for rep in range(2):   # the first run uploads the molecule and builds the caches
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            q = qmtools.QMTools.Compute_density(grid, mol, precision=prec, copyBack=False)
            v = qmtools.QMTools.Compute_hartree_iter(q, mol, tolerance=None, max_sweeps=2000, boundary="multipole",
                                                     precision=prec, copyBack=False)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        out[prec] = {"seconds": dt, "grid_pts_per_s": int(grid.npts) / dt, "sweeps": int(v.sweeps)}
    out["workload"] = "12-atom cc-pVDZ (123 bf), 96^3: density + multipole boundary + 2000 Jacobi sweeps"
    return out


def cold_request(basis):
    """First request of a molecule never seen before, on a grid shape never seen before: molecule upload,
    GTO distribution table and Green's-function transform included, against the same call repeated."""
    import torch
    import qmtools
    from qmtools import synth
    mol = synth.synthetic_molecule(40, 913, basis)
    grid = synth.centred_grid(mol, 152, step_bohr=0.13, pad_ang=3.0)
    pts = synth.uniform_points(grid, 10000, 4)
    out = {"workload": f"40-atom molecule ({mol.nbf} bf), 152^3, mixed density + free-space FFT + GTO at 1e4 points"}
    for rep in ("cold", "warm"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        q = qmtools.QMTools.Compute_density(grid, mol, precision="mixed", copyBack=False)
        qmtools.QMTools.Compute_hartree_fft(q, mol, mode="freespace", precision="mixed")
        qmtools.QMTools.Compute_hartree_points(mol, pts)
        torch.cuda.synchronize()
        out[rep + "_s"] = time.perf_counter() - t0
    return out


def jacobi_section(w, sweeps=200):
    """Jacobi sweeps on the 256^3 grid of the 60-atom workload: ms per sweep and HBM roofline fraction."""
    import torch
    import qmtools
    from qmtools import _native as nat
    out = {}
    qg = qmtools.Grid.emptyAs(w.grid)
    qg.d_qube = w.q64.clone()
    for prec, bps in (("fp64", 24.0), ("mixed", 12.0)):
        qmtools.QMTools.Compute_hartree_iter(qg, w.mol, tolerance=None, max_sweeps=50, boundary="multipole",
                                             precision=prec, copyBack=False)
        torch.cuda.synchronize()
        nat.library().qm_profile(nat.context(), 1)
        t0 = time.perf_counter()
        qmtools.QMTools.Compute_hartree_iter(qg, w.mol, tolerance=None, max_sweeps=sweeps, boundary="multipole",
                                             precision=prec, copyBack=False)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        nat.library().qm_profile(nat.context(), 0)
        ms, n = w.kernel_ms(nat.QM_K_JACOBI)
        per = ms / max(n, 1)
        gbs = bps * w.N / (per * 1e-3) / 1e9
        out[prec] = {"ms_per_sweep": per, "sweeps": n, "wall_ms_per_sweep": wall * 1000 / sweeps,
                     "algorithmic_bytes_per_point": bps, "achieved_GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
    return out


def vne_section(w, pairs=25):
    """Diffusion pseudo-V_ne on the 256^3 grid: fixed pair count, per-sweep time and HBM fraction
    (16 B/point/sweep FP64: read V, write V')."""
    import warnings
    import qmtools
    from qmtools import _native as nat
    nat.library().qm_profile(nat.context(), 1)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        t0 = time.perf_counter()
        v = qmtools.QMTools.Compute_VNe(w.grid, w.mol, tolerance=-1.0, maxiters=pairs - 1, copyBack=False)
        wall = time.perf_counter() - t0
    nat.library().qm_profile(nat.context(), 0)
    ms, n = w.kernel_ms(nat.QM_K_VNE)
    per_sweep = ms / max(n, 1) / 2
    gbs = 16.0 * w.N / (per_sweep * 1e-3) / 1e9
    return {"pairs": int(v.pairs), "ms_per_sweep": per_sweep, "wall_ms_per_pair": wall * 1e3 / max(v.pairs, 1),
            "algorithmic_bytes_per_point": 16.0, "achieved_GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}


def batched_section(basis, count, sweeps, seed=41):
    """A batch of molecules (5-80 atoms, 64^3-192^3 grids): mixed batched fused launches vs the plain
    FP64 path (one molecule at a time through the drop-in API), for each Poisson method: Jacobi
    (hipGraph, `sweeps` sweeps, multipole boundary) and free-space FFT.  Same method and sweep count
    on both sides of each ratio."""
    import torch
    import qmtools
    from qmtools import batch, synth
    mols, grids = synth.molecule_batch(count, (5, 80), (64, 96, 128, 160, 192), (0.1, 0.15), seed, basis)
    npts = batch.grid_points(grids)
    log(f"batched: {count} molecules, {npts} grid points, {sweeps} Jacobi sweeps")
    lane = batch.Lane(default=True)
    out = {"molecules": count, "grid_points": npts, "sweeps": sweeps}
    # untimed densities of the whole batch on both paths: workspaces and the allocator's cache reach
    # their size for the batch before either side is timed
    batch.density_batch(mols, grids, "mixed", lane)
    for m, g in zip(mols, grids):
        qmtools.QMTools.Compute_density(g, m, precision="fp64", copyBack=False)
    torch.cuda.synchronize()
    for poisson in ("jacobi", "fft"):
        batch.run(mols[:2], grids[:2], "mixed", 100, lane=lane, poisson=poisson)     # warm-up
        torch.cuda.synchronize()
        # density, then Poisson, for the whole batch on each side (the split costs one synchronization)
        t0 = time.perf_counter()
        qs = batch.density_batch(mols, grids, "mixed", lane)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if poisson == "jacobi":
            vs = batch.hartree_jacobi_batch(mols, grids, qs, sweeps, "mixed", "multipole", lane)
        else:
            vs = batch.hartree_fft_batch(mols, grids, qs, "mixed", lane=lane)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        dens = [qmtools.QMTools.Compute_density(g, m, precision="fp64", copyBack=False) for m, g in zip(mols, grids)]
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        plain = []
        for m, g, q in zip(mols, grids, dens):
            if poisson == "jacobi":
                v = qmtools.QMTools.Compute_hartree_iter(q, m, tolerance=None, max_sweeps=sweeps,
                                                         boundary="multipole", copyBack=False).d_qube
            else:
                v = batch.hartree_fft_batch([m], [g], [q.d_qube], "fp64", lane=lane)[0]
            plain.append((q.d_qube, v))
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        del dens
        t_mixed, t_plain = t2 - t0, t4 - t2
        parts = {"density": {"mixed_batched": t1 - t0, "plain_fp64": t3 - t2},
                 "poisson": {"mixed_batched": t2 - t1, "plain_fp64": t4 - t3}}
        rho_err = max(float((q.double() - p[0]).abs().max() / p[0].abs().max()) for q, p in zip(qs, plain))
        chg_err = max(float(abs(q.double().sum() - p[0].sum()) / p[0].sum()) for q, p in zip(qs, plain))
        v_err = max(float((v.double() - p[1]).abs().max() / p[1].abs().max()) for v, p in zip(vs, plain))
        out[poisson] = {"mixed_batched_pts_per_s": npts / t_mixed, "plain_fp64_pts_per_s": npts / t_plain,
                        "speedup": t_plain / t_mixed, "seconds": {"mixed_batched": t_mixed, "plain_fp64": t_plain},
                        "seconds_by_part": parts,
                        "max_rel_err_vs_fp64": {"rho": rho_err, "charge": chg_err, "v_hartree": v_err}}
        del qs, vs, plain
    return out


def serve_pool(basis, count=250, seed=77, grid_n=(64, 96, 128, 160, 192, 224, 256)):
    """Seeded molecules of 5-120 atoms on 64^3-256^3 grids (0.1-0.15 bohr) that requests are drawn from."""
    from qmtools import synth
    return list(zip(*synth.molecule_batch(count, (5, 120), grid_n, (0.1, 0.15), seed, basis)))


def served_section(pool, nreq, rate, sweeps, precision="mixed", poisson="fft", seed=5, round_sweeps=None):
    """Poisson arrivals at `rate` req/s into the continuous-batching server; each request draws a
    molecule and its grid from `pool`.  Results are dropped on completion (throughput and latency only).
    round_sweeps: Jacobi lanes admit between rounds of that many sweeps (None: whole batches)."""
    from qmtools import serve
    rng = numpy.random.default_rng(seed)
    items = [pool[int(i)] for i in rng.integers(0, len(pool), size=nreq)]
    srv = serve.Server(nlanes=2, precision=precision, max_batch=8, sweeps=sweeps, hbm_budget=160 << 30,
                       poisson=poisson, results="drop", round_sweeps=round_sweeps).start()
    for m, g in pool[:2]:                                            # warm-up
        srv.submit(m, g).wait(600)
    srv.completed.clear()
    t0 = time.perf_counter()
    reqs = serve.poisson_arrivals(srv, items, rate, seed)
    for r in reqs:
        r.wait(1200)
    srv.stop()
    st = srv.stats()
    st.update({"rate_req_per_s": rate, "precision": precision, "poisson": poisson,
               "sweeps": sweeps if poisson == "jacobi" else None, "wall_s": time.perf_counter() - t0})
    if poisson == "jacobi":
        st.update({"round_sweeps": srv.round_sweeps, "joined_running": srv.joined_running})
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-other", action="store_true", help="skip the comparison run of the other precision")
    ap.add_argument("--batch", type=int, default=256, help="molecules in the batched section (0: skip)")
    ap.add_argument("--serve", type=int, default=1000, help="requests per rate in the served section (0: skip)")
    ap.add_argument("--rate", type=str, default="50,100,200,500",
                    help="Poisson arrival rates of the served section (FFT Poisson), comma-separated (req/s)")
    ap.add_argument("--serve-jacobi", type=int, default=100,
                    help="requests of the served Jacobi runs (2000 sweeps, 10 req/s, mixed and FP64; 0: skip)")
    ap.add_argument("--no-cpu-full", action="store_true",
                    help="skip the full host run of the 12-atom, 96^3 workload")
    ap.add_argument("--sweeps", type=int, default=2000, help="Jacobi sweeps in the batched/served sections")
    ap.add_argument("--no-extra", action="store_true", help="skip the 256^3 Jacobi and diffusion V_ne sections")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device-to-device (a kernel of known traffic for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build runs on one MI355X; "
              "molecules and grid slabs shard across GPUs only by design)", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OPENBLAS_NUM_THREADS", str(len(os.sched_getaffinity(0))))
    import torch
    from qmtools import _native as nat
    if args.calibrate_copy:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    log("building workload")
    w = Workload(n=args.grid, precision=args.precision)
    log(f"molecule: {int(w.mol.natoms)} atoms, nbf {w.mol.nbf}; grid {w.n}^3, step {float(w.grid.step):.5f} bohr")
    other = "mixed" if args.precision == "fp64" else "fp64"
    from qmtools import autotune
    tune = {"fp64": autotune.autotune_density(w.mol, w.grid, "fp64", use_cache=False),
            "mixed": autotune.autotune_density(w.mol, w.grid, "mixed", use_cache=False)}
    log(f"autotune: {tune}")

    # main measurement
    wall, lat = w.timed(args.steps, args.warmup, profile=True)
    kid = nat.QM_K_DENSITY_CONTRACT if args.precision == "fp64" else nat.QM_K_DENSITY_SPLIT
    k_ms, k_n = w.kernel_ms(kid)
    phi_ms, phi_n = w.kernel_ms(nat.QM_K_PHI if args.precision == "fp64" else nat.QM_K_PHI_SPLIT)
    fft_ms, fft_n = w.kernel_ms(nat.QM_K_FFT_AXIS)
    gto_ms, gto_n = w.kernel_ms(nat.QM_K_GTO)
    ms_step = wall * 1000.0 / args.steps
    log(f"{args.precision}: {ms_step:.2f} ms/step; contraction {k_ms / max(k_n, 1):.3f} ms x {k_n}, "
        f"phi {phi_ms / max(phi_n, 1):.3f} ms x {phi_n}, fft pass {fft_ms / max(fft_n, 1):.3f} ms x {fft_n}, "
        f"gto {gto_ms / max(gto_n, 1):.3f} ms x {gto_n}")

    # keep the results of this path, run the other path once for the comparison
    other_out, err_out = None, None
    if not args.no_other:
        q_main = (w.q64 if args.precision == "fp64" else w.q32).double().clone()
        v_main = (w.v64 if args.precision == "fp64" else w.v32).double().clone()
        wall_o, lat_o = w.timed(max(2, args.steps // 2), 1, precision=other)
        q_other = (w.q64 if other == "fp64" else w.q32).double()
        v_other = (w.v64 if other == "fp64" else w.v32).double()
        q64, q_mx = (q_main, q_other) if args.precision == "fp64" else (q_other, q_main)
        v64, v_mx = (v_main, v_other) if args.precision == "fp64" else (v_other, v_main)
        err_out = {"rho_maxabs_rel": float((q_mx - q64).abs().max() / q64.abs().max()),
                   "charge_rel": float(abs(q_mx.sum() - q64.sum()) / q64.sum()),
                   "vh_fft_maxabs_rel": float((v_mx - v64).abs().max() / v64.abs().max())}
        other_out = {"value": w.N / (wall_o / max(2, args.steps // 2)), "p50_ms": pct(lat_o, 50),
                     "p99_ms": pct(lat_o, 99)}
        if other == "mixed":
            wall_m, _ = w.timed(2, 0, precision="mixed", profile=True)
            km, kn = w.kernel_ms(nat.QM_K_DENSITY_SPLIT)
            pm, _ = w.kernel_ms(nat.QM_K_PHI_SPLIT)
            fm, _ = w.kernel_ms(nat.QM_K_FFT_AXIS)
            gm, _ = w.kernel_ms(nat.QM_K_GTO)
            other_out["breakdown_ms_per_step"] = {"contraction": km / 2, "phi": pm / 2, "fft_passes": fm / 2,
                                                  "gto": gm / 2}
            other_out["contraction_avg_ms"] = km / max(kn, 1)
            other_out["contraction_launches_per_step"] = kn / 2
            # the split contraction on the FP16 MFMAs: 3 products of the triangular form per point
            nbf_m = w.mol.nbf
            pts_m = w.N / max(kn / 2, 1)
            ach_m = 3.0 * (nbf_m * (nbf_m + 1.0) + 2.0 * nbf_m) * pts_m / (km / max(kn, 1) * 1e-3) / 1e12
            other_out["contraction_roofline"] = {"bound": "mfma", "achieved": ach_m, "peak": F16_MFMA_PEAK_TFLOPS,
                                                 "unit": "TFLOP/s", "frac": ach_m / F16_MFMA_PEAK_TFLOPS,
                                                 "algorithmic": "3 x (nbf (nbf + 1) + 2 nbf) FP16 MFMA flops/point"}
            log(f"mixed breakdown: {other_out['breakdown_ms_per_step']}")
        w.q64.copy_(q64)

    # roofline of the dominant kernel (the density contraction).  The LDS-staged
    # kernels contract with the upper-triangular M = D + 2U of the symmetric P,
    # nbf (nbf + 1) + 2 nbf flops per point (half the dense 2 nbf^2 + 2 nbf);
    # the barrier-free FP64 variant is dense.
    nbf = w.mol.nbf
    launch_pts = w.N / max(k_n / args.steps, 1)
    dense = args.precision == "fp64" and tune["fp64"].get("f64_variant", 0) == 1
    per_point = (2.0 * nbf * nbf if dense else nbf * (nbf + 1.0)) + 2.0 * nbf
    flops_per_launch = per_point * launch_pts
    mult = 1.0 if args.precision == "fp64" else 3.0     # hi/lo split executes 3 products
    avg_ms = k_ms / max(k_n, 1)
    achieved = flops_per_launch * mult / (avg_ms * 1e-3) / 1e12
    dense_equiv = (2.0 * nbf * nbf + 2.0 * nbf) * launch_pts * mult / (avg_ms * 1e-3) / 1e12
    peak = FP64_MFMA_PEAK_TFLOPS if args.precision == "fp64" else F16_MFMA_PEAK_TFLOPS
    variant = ({0: "lds", 1: "direct", 2: "tiles", 3: "rows", 4: "tiles_dma"}[tune["fp64"].get("f64_variant", 0)]
               if args.precision == "fp64"
               else tune["mixed"].get("split_variant"))
    meas = MEASURED_TRAFFIC.get((args.precision, (nbf + 63) // 64 * 64, variant))
    traffic = None if meas is None else meas[1] / meas[0] * launch_pts

    check = None
    if not args.no_check and args.precision == "fp64":
        log("oracle sample check")
        err, gsum, raw = oracle_sample_check(w)
        check = {"density_sample_rel_err_vs_numpy_fp64": err, "voxels": 4096, "normalised_sum": gsum,
                 "raw_sum": raw}
    cpu = cpu_small = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(w)
        log(f"cpu baseline: {cpu['seconds_per_step']}")
        if not args.no_cpu_full:
            cpu_small = cpu_small_workload()
            log(f"cpu 12 atoms 96^3: {cpu_small}")
    gpu_small = gpu_small_workload()
    log(f"gpu 12 atoms 96^3: {gpu_small}")
    cold = cold_request(w.basis)
    log(f"cold request: {cold}")

    jac = vne = None
    if not args.no_extra:
        jac = jacobi_section(w)
        log(f"jacobi: {jac}")
        vne = vne_section(w)
        log(f"vne diffusion: {vne}")
    batched = served = None
    if args.batch > 0:
        batched = batched_section(w.basis, args.batch, args.sweeps)
        for k in ("jacobi", "fft"):
            log(f"batched ({k}): {batched[k]['speedup']:.2f}x plain FP64, errors {batched[k]['max_rel_err_vs_fp64']}")
    if args.serve > 0 or args.serve_jacobi > 0:
        pool = serve_pool(w.basis)
        served = {"fft_mixed": [], "fft_plain_fp64": None, "jacobi": None, "jacobi_rounds": None,
                  "requests": "Poisson arrivals; molecules of 5-120 atoms on 64^3-256^3 grids drawn from a "
                              "seeded pool of 250"}
        if args.serve > 0:
            for rate in (float(x) for x in args.rate.split(",")):
                st = served_section(pool, args.serve, rate, args.sweeps)
                log(f"served fft mixed: {st}")
                served["fft_mixed"].append(st)
            st = served_section(pool, args.serve, 100.0, args.sweeps, precision="fp64")
            log(f"served fft fp64: {st}")
            served["fft_plain_fp64"] = st
        if args.serve_jacobi > 0:
            served["jacobi"] = [served_section(pool, args.serve_jacobi, 10.0, args.sweeps, precision=p,
                                               poisson="jacobi") for p in ("mixed", "fp64")]
            log(f"served jacobi: {served['jacobi']}")
            # the same arrivals with admission between rounds of 200 sweeps
            served["jacobi_rounds"] = served_section(pool, args.serve_jacobi, 10.0, args.sweeps, precision="mixed",
                                                     poisson="jacobi", round_sweeps=200)
            log(f"served jacobi, rounds: {served['jacobi_rounds']}")
        del pool

    value = w.N / (ms_step / 1000.0)
    out = {
        "metric": "grid pts/s density+Hartree at <=1e-6 rel err vs FP64; p50/p99 latency",
        "value": value,
        "unit": "grid pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "p50_ms": pct(lat, 50),
        "p99_ms": pct(lat, 99),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f16x3/f32",
        "data": "synthetic (seeded 60-atom molecule, P = 2 C C^T with QR-orthonormal C)",
        "config": {"workload": "60-atom cc-pVDZ (606 bf), 256^3 grid: density + free-space FFT Poisson "
                               "(512^3 padded) + GTO V_H at 1e4 points",
                   "precision": args.precision, "grid": [w.n] * 3, "step_bohr": float(w.grid.step),
                   "nbf": nbf, "points": w.npoints},
        "roofline": {"bound": "mfma", "kernel": "density contraction (Phi P Phi^T diagonal)",
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic,
                     "traffic_source": "profiles/mol60_grid256_pmc_traffic.md",
                     "algorithmic": (f"({'2 nbf^2' if dense else 'nbf (nbf + 1)'} + 2 nbf) flops/point x "
                                     f"{launch_pts:.0f} points/launch"
                                     + (" x 3 split products" if mult > 1 else "")
                                     + ("" if dense else " (triangular M = D + 2U of symmetric P)")
                                     + ("" if mult > 1 else "; counts the whole triangle, of which the kernel "
                                        "skips the groups of 16 functions that are zero on a row tile")),
                     "dense_equivalent_tflops": dense_equiv,
                     "avg_ms": avg_ms, "launches": k_n},
        "breakdown_ms_per_step": {"contraction": k_ms / args.steps, "phi": phi_ms / args.steps,
                                  "fft_passes": fft_ms / args.steps, "gto": gto_ms / args.steps},
        other + "_path": other_out,
        "mixed_vs_fp64": err_out,
        "jacobi_256": jac,
        "vne_diffusion_256": vne,
        "batched": batched,
        "served": served,
        "oracle_check": check,
        "autotune": tune,
        "small_workload_gpu": gpu_small,
        "small_workload_cpu": cpu_small,
        "cold_request": cold,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
