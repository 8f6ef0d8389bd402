"""Multi-field regular grid (reference `qmtools/qmtools.py:193-432`).

Geometry and layout follow the reference exactly:

* `shape = (nfields, nx, ny, nz)`, each spatial size a multiple of 8
  (`:203-213`); a 3-tuple is augmented with nfields=1 (`:199-201`);
* `step` is given in Angstrom and stored in bohr (`:221`);
* the device buffer of field f holds point (i, j, k) at flat index
  `f*npts + i + nx*(j + ny*k)` (x fastest, reference `kernel_density.cu:151-154`),
  and `qube` is a host array with the NumPy shape `(nfields, nx, ny, nz)` whose
  flat contents are that same x-fastest order (`:215`).  So `qube[0].ravel()[i + nx*(j+ny*k)]`
  is point (i, j, k); `Grid.field_xyz(f)` returns the field as an array indexed [i, j, k].

The device buffer is a torch tensor allocated on first use (`d_qube`), so
that grids can be built, saved and loaded on a machine without a GPU.
Precision: the default dtype is float64 (the plain FP64 path); float32 grids
are used by the mixed-precision path.  On-disk formats (`SaveBIN`, `LoadBIN`,
`SaveBINmulti`) keep the reference's float32 records byte for byte.
"""
from __future__ import annotations

import pickle
import struct

import numpy

from . import _device, _safeload

ANG2BOR = 1.8897259886
_F3 = numpy.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8")])


def make_float3(x, y, z):
    """A record with ['x'], ['y'], ['z'] access, like pycuda's `make_float3` (FP64 here)."""
    return numpy.array((float(x), float(y), float(z)), dtype=_F3)[()]


def origin_tuple(o):
    return (float(o["x"]), float(o["y"]), float(o["z"]))


class Grid:
    """Regular grid with `nfields` fields (reference `qmtools.py:193-278`)."""

    def __init__(self, shape, step, origin=(0.0, 0.0, 0.0), dtype=numpy.float64):
        shape = tuple(int(s) for s in shape)
        if len(shape) == 3:
            shape = (1,) + shape
        if len(shape) != 4:
            raise ValueError("shape has to be 4-dimensional: nfields, nx, ny, nz")
        if shape[0] < 1 or any(s <= 0 or s % 8 for s in shape[1:]):
            raise ValueError("shape for xyz have to be multiple of 8")
        self.shape = shape
        self.GPUblocks = tuple(s // 8 for s in shape[1:])
        self.dtype = numpy.dtype(dtype)
        if self.dtype not in (numpy.dtype(numpy.float32), numpy.dtype(numpy.float64)):
            raise ValueError("grid dtype must be float32 or float64")
        self.qube = numpy.zeros(shape, dtype=self.dtype)
        self.npts = numpy.uint32(shape[1] * shape[2] * shape[3])
        self.nfields = numpy.uint32(shape[0])
        self.step = numpy.float64(step * ANG2BOR)
        self.origin = make_float3(*origin) if not isinstance(origin, numpy.void) else origin
        self.lattice = numpy.eye(3, dtype=numpy.float64)
        self._d_qube = None

    def __str__(self):
        return "origin: {}\nshape: {} -- points: {} gpublocks: {}".format(
            self.origin, self.shape, self.npts, self.GPUblocks)

    # -- geometry -------------------------------------------------------------
    def space_shape(self):
        return (self.shape[1], self.shape[2], self.shape[3])

    def space_shape_uint3(self):
        return numpy.array(self.space_shape(), dtype=numpy.uint32)

    @property
    def nbytes(self):
        return self.qube.nbytes

    def geometry(self):
        """(nx, ny, nz, step_bohr, (x0, y0, z0)) as Python numbers."""
        nx, ny, nz = self.space_shape()
        return nx, ny, nz, float(self.step), origin_tuple(self.origin)

    def field_xyz(self, f=0):
        """Field f as a host array indexed [i, j, k] (a view of `qube`)."""
        nx, ny, nz = self.space_shape()
        return self.qube[f].reshape(-1).reshape(nz, ny, nx).transpose(2, 1, 0)

    # -- device mirror ----------------------------------------------------------
    @property
    def d_qube(self):
        """Device buffer (flat torch tensor, all fields), uploaded from `qube` on first use."""
        if self._d_qube is None:
            self._d_qube = _device.from_host(self.qube.reshape(-1))
        return self._d_qube

    @d_qube.setter
    def d_qube(self, t):
        if t is not None and t.numel() != self.qube.size:
            raise ValueError("device buffer size does not match the grid")
        self._d_qube = t

    def upload(self):
        """Copy `qube` to the device buffer."""
        if self._d_qube is None:
            self._d_qube = _device.from_host(self.qube.reshape(-1))
        else:
            _device.copy_from_host(self._d_qube, self.qube.reshape(-1))

    def download(self):
        """Copy the device buffer back into `qube`."""
        if self._d_qube is not None:
            self.qube = _device.to_host(self._d_qube).reshape(self.shape).astype(self.dtype, copy=False)
        return self.qube

    # -- constructors -----------------------------------------------------------
    @staticmethod
    def DensityGrid(molecule, step, fat, dtype=numpy.float64):
        """Grid around a molecule with `fat` Angstrom padding (reference `qmtools.py:243-260`)."""
        xyz = numpy.array(molecule.coords, dtype=numpy.float64, copy=True)
        crdmax = numpy.amax(xyz, axis=0) + fat * ANG2BOR
        crdmin = numpy.amin(xyz, axis=0) - fat * ANG2BOR
        grd = (crdmax - crdmin) / (step * ANG2BOR) / 8.0
        grd = numpy.ceil(grd).astype(numpy.int64) * 8
        grid = Grid((1, int(grd[0]), int(grd[1]), int(grd[2])), step, dtype=dtype)
        grid.origin = make_float3(*crdmin)
        return grid

    @staticmethod
    def emptyAs(g, nfields=None, dtype=None):
        """Zero grid with the geometry of `g` (reference `qmtools.py:266-278`)."""
        shape = g.shape if nfields is None else (int(nfields),) + tuple(g.shape[1:])
        grid = Grid(shape, 1.0, dtype=g.dtype if dtype is None else dtype)
        grid.origin = g.origin
        grid.lattice = g.lattice
        grid.step = g.step
        return grid

    # -- data in/out ------------------------------------------------------------
    def LoadData(self, filename):
        """Load `qube` from an .npy file of the same shape (reference `qmtools.py:283-287`)."""
        self.LoadFromNPY(numpy.load(filename, allow_pickle=False))

    def LoadFromNPY(self, npyarray):
        """Set `qube` from an array of the grid's shape (reference `qmtools.py:317-327`)."""
        if npyarray.ndim != self.qube.ndim:
            raise ValueError("dimensionality mismatch")
        if tuple(npyarray.shape) != tuple(self.qube.shape):
            raise ValueError("array shape mismatch {} (expected {})".format(npyarray.shape, self.qube.shape))
        self.qube = numpy.ascontiguousarray(npyarray, dtype=self.dtype)
        if self._d_qube is not None or _device.available():
            self.upload()

    def Save(self, filename):
        """Pickle the grid (reference `qmtools.py:290-300`)."""
        d = {"shape": self.shape, "origin": self.origin, "step": self.step,
             "lattice": self.lattice, "qube": self.qube}
        with open(filename, "wb") as fh:
            pickle.dump(d, fh)

    @staticmethod
    def Load(molecule, filename):
        """Restore a grid written by `Save` (reference `qmtools.py:303-314`).

        The file is read with an unpickler that rebuilds plain data and NumPy arrays
        only and refuses any other pickle global, so loading runs nothing from the file.
        """
        d = _safeload.load(filename)
        if not isinstance(d, dict) or not {"shape", "origin", "step", "lattice", "qube"} <= set(d):
            raise ValueError(f"{filename}: not a grid written by Grid.Save")
        q = numpy.asarray(d["qube"])
        grid = Grid(d["shape"], 0.1, dtype=q.dtype if q.dtype in (numpy.float32, numpy.float64) else numpy.float64)
        o = d["origin"]
        grid.origin = make_float3(o["x"], o["y"], o["z"])
        grid.step = numpy.float64(d["step"])
        grid.lattice = numpy.asarray(d["lattice"])
        grid.qube = numpy.ascontiguousarray(q, dtype=grid.dtype)
        return grid

    def _bin_header(self, mol):
        out = []
        if mol is None:
            out.append(struct.pack("<i", 0))
        else:
            out.append(struct.pack("<i", int(mol.natoms)))
            for i in range(int(mol.natoms)):
                out.append(struct.pack("<i", int(mol.types[i])))
                out.append(struct.pack("<fff", *[float(numpy.float32(c)) for c in mol.coords[i]]))
        out.append(struct.pack("<iii", self.shape[1], self.shape[2], self.shape[3]))
        out.append(struct.pack("<f", float(numpy.float32(self.step))))
        ox, oy, oz = origin_tuple(self.origin)
        out.append(struct.pack("<fff", ox, oy, oz))
        return b"".join(out)

    def SaveBIN(self, filename, mol=None):
        """Binary grid file, first field only (reference `qmtools.py:331-367`).

        Layout: int32 natoms; natoms x (int32 Z, f32 x, f32 y, f32 z [bohr]);
        int32 nx, ny, nz; f32 step [bohr]; f32 origin x, y, z [bohr];
        f32[nx*ny*nz] values in x-fastest order.
        """
        npts = self.shape[1] * self.shape[2] * self.shape[3]
        linear = numpy.reshape(self.qube, [npts])  # like the reference: needs nfields == 1
        with open(filename, "wb") as fh:
            fh.write(self._bin_header(mol))
            fh.write(linear.astype("<f4").tobytes())

    def SaveBINmulti(self, filename, mol=None):
        """One `SaveBIN`-format file per field, named '<filename>-<f>.bin' (reference `qmtools.py:370-401`)."""
        npts = self.shape[1] * self.shape[2] * self.shape[3]
        for f in range(self.shape[0]):
            with open("{}-{}.bin".format(filename, f), "wb") as fh:
                fh.write(self._bin_header(mol))
                fh.write(numpy.reshape(self.qube[f], [npts]).astype("<f4").tobytes())

    @staticmethod
    def LoadBIN(filename, dtype=numpy.float64):
        """Read a `SaveBIN` file (reference `qmtools.py:406-432`)."""
        with open(filename, "rb") as fh:
            raw = fh.read()
        (natm,) = struct.unpack_from("<i", raw, 0)
        pos = 4 + 16 * natm
        shape = struct.unpack_from("<iii", raw, pos); pos += 12
        (step,) = struct.unpack_from("<f", raw, pos); pos += 4
        origin = struct.unpack_from("<fff", raw, pos); pos += 12
        n = shape[0] * shape[1] * shape[2]
        linear = numpy.frombuffer(raw, dtype="<f4", count=n, offset=pos)
        grid = Grid(shape, step, dtype=dtype)
        grid.step = numpy.float64(numpy.float32(step))
        grid.origin = make_float3(*origin)
        grid.qube = linear.astype(grid.dtype).reshape(grid.shape)
        return grid
