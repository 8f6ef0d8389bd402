"""Drop-in module `qmtools.qmtools` (reference `qmtools/qmtools.py`).

Same names and call surface as the reference for the grid hot path:
`BasisSet`, `Molecule`, `Grid`, and the static `QMTools.Compute_*` methods
(reference `qmtools.py:2327-2831`).  Compute calls run hand-written gfx950
HIP kernels from libqmgrid.so through ctypes; there is no CPU fallback.

Precision: `QMTools.precision` (default "fp64") selects the plain FP64 path;
"mixed" runs the density contraction on FP16 MFMA with a hi/lo split (FP32
basis values), FP32 Jacobi sweeps with FP64 residuals and an FP32 FFT, with
charge normalisation, boundary multipoles and V_ne in FP64.  Every call also
accepts `precision=` to override the default.
"""
from __future__ import annotations

import ctypes
import warnings

import numpy

from . import _device
from . import _native as nat
from .basis import MAXAOC, MAXAOS, BasisSet  # noqa: F401
from .grid import Grid, make_float3  # noqa: F401
from .molecule import ANG2BOR, Molecule  # noqa: F401

MOLFAT = 2.0
CUDA_SRC_DIR = None   # the reference's kernel-source directory; kernels here are precompiled

_PREC = {"fp64": nat.QM_PREC_FP64, "mixed": nat.QM_PREC_F16X3, "f16": nat.QM_PREC_F16X3,
         "bf16": nat.QM_PREC_BF16X3}


def _np_dtype_code(dt):
    return nat.QM_F64 if numpy.dtype(dt) == numpy.float64 else nat.QM_F32


class _DeviceMol:
    """Device-resident nuclei, basis tables and density matrix of a Molecule."""

    def __init__(self, mol: Molecule):
        lib = nat.library()
        ctx = nat.context()
        out = ctypes.c_void_p()
        z = numpy.ascontiguousarray(mol.types, dtype=numpy.int32)
        xyz = numpy.ascontiguousarray(mol.coords, dtype=numpy.float64)
        rc = lib.qm_mol_create(ctx, int(mol.natoms), nat.i32p(z), nat.f64p(xyz), int(mol.nshell),
                               nat.i32p(mol.shell_atom), nat.i32p(mol.shell_L), nat.i32p(mol.shell_nprim),
                               nat.i32p(mol.shell_poff), int(mol.prim_alpha.size), nat.f64p(mol.prim_alpha),
                               nat.f64p(mol.prim_coeff), ctypes.byref(out))
        nat.check(rc, ctx)
        self.handle = out.value
        self._ctx = ctx
        self.set_dm(mol.dm)

    def set_dm(self, dm):
        d = numpy.ascontiguousarray(dm, dtype=numpy.float64)
        nat.check(nat.library().qm_mol_set_dm(self._ctx, self.handle, nat.stream(), nat.f64p(d)), self._ctx)

    def update_coords(self, coords):
        """New nuclear positions, ordered on the current stream; P and the basis tables stay resident."""
        xyz = numpy.ascontiguousarray(coords, dtype=numpy.float64)
        nat.check(nat.library().qm_mol_set_coords(self._ctx, self.handle, nat.stream(), nat.f64p(xyz)), self._ctx)

    def stats(self):
        """{'p_uploads', 'coord_updates', 'gto_builds'} counters of the resident tables."""
        c = (ctypes.c_int64 * 3)()
        nat.check(nat.library().qm_mol_stats(self.handle, c), self._ctx)
        return {"p_uploads": c[0], "coord_updates": c[1], "gto_builds": c[2]}

    def __del__(self):
        try:
            if self.handle:
                nat.library().qm_mol_destroy(self._ctx, self.handle)
        except Exception:  # pragma: no cover - interpreter shutdown
            pass
        self.handle = None


def device_molecule(mol: Molecule) -> _DeviceMol:
    """The molecule's device tables (uploaded on first use; `mol.upload()` refreshes)."""
    if mol._dev is None:
        mol._dev = _DeviceMol(mol)
    return mol._dev


def _upload_molecule(self):
    if self._dev is None:
        self._dev = _DeviceMol(self)
    else:
        self._dev.update_coords(self.coords)
        self._dev.set_dm(self.dm)


Molecule.upload = _upload_molecule


def _prec(precision):
    p = QMTools.precision if precision is None else precision
    if p not in _PREC:
        raise ValueError(f"unknown precision {p!r}; use one of {sorted(_PREC)}")
    return p


def _field0(grid):
    """Device tensor view of field 0."""
    return grid.d_qube[: int(grid.npts)]


def _as_device(grid, dtype):
    """Field 0 of `grid` on the device in `dtype` (a view when the dtype matches)."""
    t = _field0(grid)
    want = _device._dtype(dtype)
    return t if t.dtype == want else t.to(want)


def _result_grid(template, dtype):
    """A new single-field grid with the geometry of `template` for a computed result: its device
    buffer is allocated and cleared on the device, and the host `qube` keeps the zeros of a fresh
    grid (nothing is copied between host and device)."""
    g = Grid.emptyAs(template, nfields=1, dtype=dtype)
    g.d_qube = _device.zeros(g.qube.size, dtype)
    return g


class QMTools:
    """Grid hot-path compute API (reference `qmtools.py:2320-2831`)."""

    precision = "fp64"

    # ---------------------------------------------------------------- density
    @staticmethod
    def Compute_density(grid, molecule, subgrid=1, copyBack=True, debug=False, precision=None):
        """Electrons per voxel, normalised to the nuclear charge (reference `qmtools.py:2381-2430`).

        Returns a new Grid with the geometry of `grid`.  fp64 grids for
        precision "fp64", float32 grids for the mixed paths.
        """
        p = _prec(precision)
        dt = numpy.float64 if p == "fp64" else numpy.float32
        cgrid = _result_grid(grid, dt)
        dm = device_molecule(molecule)
        raw = ctypes.c_double(0.0)
        nat.call("qm_density", dm.handle, nat.stream(), ctypes.byref(nat.geom(cgrid)), int(subgrid), _PREC[p],
                 _np_dtype_code(dt), nat.dptr(cgrid.d_qube), 1, ctypes.byref(raw))
        cgrid.raw_total = raw.value
        if debug:
            print(f"density: raw total {raw.value:.12g}, scaled to {int(molecule.qtot)}")
        if copyBack:
            cgrid.download()
        return cgrid

    # ---------------------------------------------------------------- Hartree (analytic)
    @staticmethod
    def Compute_hartree(vgrid, molecule, copyBack=True):
        """Potential of nuclei + GTO density at the grid nodes, written into `vgrid` (reference `qmtools.py:2436-2489`)."""
        dm = device_molecule(molecule)
        nat.call("qm_hartree_gto_grid", dm.handle, nat.stream(), ctypes.byref(nat.geom(vgrid)),
                 _np_dtype_code(vgrid.dtype), nat.dptr(vgrid.d_qube))
        if copyBack:
            vgrid.download()
        return None

    @staticmethod
    def Compute_hartree_points(molecule, points):
        """Potential of nuclei + GTO density (a.u.) at arbitrary points (bohr, shape (n, 3)).

        No reference counterpart: the reference evaluates only on full grids.
        """
        pts = numpy.ascontiguousarray(points, dtype=numpy.float64).reshape(-1, 3)
        dm = device_molecule(molecule)
        dpts = _device.from_host(pts.ravel())
        out = _device.empty(pts.shape[0], numpy.float64)
        nat.call("qm_hartree_gto_points", dm.handle, nat.stream(), nat.dptr(dpts), pts.shape[0], nat.dptr(out))
        return _device.to_host(out)

    # ---------------------------------------------------------------- Hartree (Jacobi)
    @staticmethod
    def Compute_hartree_iter(qgrid, molecule, tolerance=1.0e-24, copyBack=True, boundary="zero",
                             max_sweeps=100000, check_every=200, precision=None):
        """Jacobi relaxation of lap V = rho/h^3 with rho = seed - Q (reference `qmtools.py:2497-2573`).

        The result is V = -phi/(4 pi) (phi: potential of the seed nuclei and
        electrons, a.u.).  Like the reference the loop stops when
        sum |V - T| <= tolerance, checked every `check_every` sweeps (200 =
        every 100 iterations of two sweeps).  Unlike the reference it stops
        after `max_sweeps` (with a RuntimeWarning); `tolerance=None` runs
        exactly `max_sweeps` sweeps.  boundary: "zero" (reference) or
        "multipole" (outer layer held at the multipole potential of rho).
        Mixed precision: FP32 sweeps on the correction with an FP64 residual
        every 50 sweeps.
        """
        p = _prec(precision)
        seed = QMTools.Compute_qseed(qgrid, molecule, copyBack=False)
        rho = seed.d_qube[: int(qgrid.npts)] - _as_device(qgrid, numpy.float64)
        vgrid = _result_grid(qgrid, numpy.float64)
        v = vgrid.d_qube
        g = nat.geom(vgrid)
        if boundary == "multipole":
            mom = (ctypes.c_double * 10)()
            nat.call("qm_multipole_boundary", nat.stream(), ctypes.byref(g), nat.QM_F64, nat.dptr(rho),
                     nat.dptr(v), mom)
            vgrid.multipoles = numpy.array(mom[:])
        elif boundary != "zero":
            raise ValueError("boundary must be 'zero' or 'multipole'")
        tol = -1.0 if tolerance is None else float(tolerance)
        done = ctypes.c_int64(0)
        delta = ctypes.c_double(-1.0)
        if p == "fp64":
            tmp = v.clone()
            prec, every = nat.QM_PREC_FP64, int(check_every)
            tptr = nat.dptr(tmp)
        else:
            prec, every, tptr = nat.QM_PREC_FP32, 50, None
        ctx = nat.context()
        rc = nat.library().qm_jacobi(ctx, nat.stream(), ctypes.byref(g), prec, nat.dptr(rho), nat.dptr(v), tptr,
                                     int(max_sweeps), tol, every, ctypes.byref(done), ctypes.byref(delta))
        if rc == nat.QM_ENOCONV:
            warnings.warn(f"Compute_hartree_iter: not converged after {done.value} sweeps "
                          f"(delta {delta.value:.3e} > {tol:.3e})", RuntimeWarning)
        else:
            nat.check(rc, ctx)
        vgrid.sweeps = done.value
        vgrid.delta = delta.value
        if copyBack:
            vgrid.download()
        return vgrid

    # ---------------------------------------------------------------- Hartree (FFT)
    @staticmethod
    def Compute_hartree_fft(qgrid, molecule, sigma=0.2, cutoff=6, mode="periodic", precision=None):
        """FFT Poisson solve of electrons + Gaussian nuclei (reference `qmtools.py:2578-2639`).

        Returns a host array V[i, j, k] in volts (C order, shape (nx, ny, nz)).
        mode "periodic" follows the reference; "freespace" uses a 2x zero-padded
        grid and the 1/r Green's function (open boundaries).
        """
        p = _prec(precision)
        m = {"periodic": nat.QM_FFT_PERIODIC, "freespace": nat.QM_FFT_FREESPACE}[mode]
        dm = device_molecule(molecule)
        q = _as_device(qgrid, numpy.float64)
        out = _device.empty(int(qgrid.npts), numpy.float64)
        prec = nat.QM_PREC_FP64 if p == "fp64" else nat.QM_PREC_FP32
        nat.call("qm_hartree_fft", dm.handle, nat.stream(), ctypes.byref(nat.geom(qgrid)), prec, nat.QM_F64,
                 nat.dptr(q), float(sigma), float(cutoff), m, nat.dptr(out))
        return _device.to_host(out).reshape(qgrid.space_shape())

    # ---------------------------------------------------------------- seed, V_ne
    @staticmethod
    def Compute_qseed(grid, molecule, copyBack=True):
        """Nuclear charges spread on the 8 surrounding nodes (reference `qmtools.py:2645-2672`)."""
        qgrid = _result_grid(grid, numpy.float64)
        dm = device_molecule(molecule)
        nat.call("qm_seed", dm.handle, nat.stream(), ctypes.byref(nat.geom(qgrid)), nat.QM_F64,
                 nat.dptr(qgrid.d_qube))
        if copyBack:
            qgrid.download()
        return qgrid

    @staticmethod
    def Compute_VNe(grid, molecule, adsorb=0.1, diff=0.01, tolerance=1.0e-9, maxiters=None, copyBack=True,
                    precision=None):
        """Diffusion / adsorption pseudo-nuclear field (reference `qmtools.py:2677-2745`, `kernel_vne.cu`).

        Pairs of sweeps from 0 until sum|res - out| / npts <= tolerance; after
        more than `maxiters` pairs (default 10 npts^(1/3)) it stops with a
        RuntimeWarning, where the reference prints "VNe not converged!".  The
        grid argument is a template; the result is a new single-field Grid
        (FP64, or FP32 for the reduced-precision modes).  The pair count and
        last delta are kept as `.pairs` and `.delta` on the result.
        """
        p = _prec(precision)
        dt = numpy.float64 if p == "fp64" else numpy.float32
        gres = _result_grid(grid, dt)
        tmp = _result_grid(grid, dt)
        if maxiters is None:
            maxiters = 10 * numpy.power(grid.npts, 1.0 / 3)
        dm = device_molecule(molecule)
        pairs = ctypes.c_int64(0)
        delta = ctypes.c_double(0.0)
        conv = ctypes.c_int(1)
        nat.call("qm_vne_diffusion", dm.handle, nat.stream(), ctypes.byref(nat.geom(gres)),
                 nat.QM_F64 if dt == numpy.float64 else nat.QM_F32, float(adsorb), float(diff), float(tolerance),
                 float(maxiters), nat.dptr(gres.d_qube), nat.dptr(tmp.d_qube), ctypes.byref(pairs),
                 ctypes.byref(delta), ctypes.byref(conv))
        gres.pairs = pairs.value
        gres.delta = delta.value
        if not conv.value:
            warnings.warn(f"Compute_VNe: not converged after {pairs.value} pairs (delta {delta.value:.3e}, "
                          f"maxiters {maxiters})", RuntimeWarning)
        if copyBack:
            gres.download()
        return gres

    @staticmethod
    def Compute_VNe_classic(grid, molecule, divisions=4, copyBack=True):
        """Classical nuclear potential averaged over divisions^3 sub-points (reference `qmtools.py:2750-2790`)."""
        gout = _result_grid(grid, numpy.float64)
        dm = device_molecule(molecule)
        nat.call("qm_vne_classic", dm.handle, nat.stream(), ctypes.byref(nat.geom(gout)), int(divisions),
                 nat.QM_F64, nat.dptr(gout.d_qube))
        if copyBack:
            gout.download()
        return gout

    # ---------------------------------------------------------------- reductions
    @staticmethod
    def Compute_qtot(grid):
        """Sum of field 0 (reference `qmtools.py:2796-2803`), FP64, reproducible."""
        out = ctypes.c_double(0.0)
        nat.call("qm_reduce", nat.stream(), nat.QM_RED_SUM, _np_dtype_code(grid.dtype), nat.dptr(_field0(grid)),
                 None, int(grid.npts), ctypes.byref(out))
        return float(out.value)

    @staticmethod
    def Compute_qscale(grid, factor, copyBack=False):
        """Field 0 *= factor (reference `qmtools.py:2807-2812`)."""
        nat.call("qm_scale", nat.stream(), _np_dtype_code(grid.dtype), nat.dptr(_field0(grid)), int(grid.npts),
                 float(factor))
        if copyBack:
            grid.download()

    @staticmethod
    def Compute_qdiff(grid1, grid2, rel=False):
        """sum |a - b| of field 0, or 100 * sum |a - b| / b with rel=True (reference `qmtools.py:2817-2831`)."""
        dt = numpy.float64 if numpy.float64 in (grid1.dtype, grid2.dtype) else numpy.float32
        a = _as_device(grid1, dt)
        b = _as_device(grid2, dt)
        out = ctypes.c_double(0.0)
        op = nat.QM_RED_RELDIFF if rel else nat.QM_RED_ABSDIFF
        nat.call("qm_reduce", nat.stream(), op, _np_dtype_code(dt), nat.dptr(a), nat.dptr(b), int(grid1.npts),
                 ctypes.byref(out))
        r = float(out.value)
        return r * 100.0 if rel else r
