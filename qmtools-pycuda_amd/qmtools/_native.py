"""ctypes binding of libqmgrid.so (the C-ABI declared in include/qmgrid.h).

The library is loaded from the package directory.  There is no fallback: if
the library is missing or a GPU is absent, every compute call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

from . import _device

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libqmgrid.so")

QM_OK, QM_EINVAL, QM_ENOMEM, QM_EHIP, QM_ENOCONV, QM_ELIMIT = range(6)
QM_F64, QM_F32 = 0, 1
QM_PREC_FP64, QM_PREC_F16X3, QM_PREC_BF16X3, QM_PREC_FP32 = 0, 1, 2, 3
QM_BC_ZERO, QM_BC_MULTIPOLE = 0, 1
QM_FFT_PERIODIC, QM_FFT_FREESPACE = 0, 1
QM_RED_SUM, QM_RED_ABSDIFF, QM_RED_RELDIFF = 0, 1, 2

_ERRNAMES = {1: "invalid argument", 2: "out of device memory", 3: "HIP error", 4: "not converged",
             5: "size limit"}


class QMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qmgrid error {code} ({_ERRNAMES.get(code, '?')}): {msg}")
        self.code = code


class NotConverged(QMError):
    pass


class Geom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("step", ctypes.c_double), ("origin", ctypes.c_double * 3)]


_c = ctypes
_vp = _c.c_void_p
_i = _c.c_int
_i64 = _c.c_int64
_d = _c.c_double
_pd = _c.POINTER(_c.c_double)
_pi = _c.POINTER(_c.c_int)
_pg = _c.POINTER(Geom)

# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "qm_version": [],
    "qm_ctx_create": [_i, _c.POINTER(_vp)],
    "qm_ctx_destroy": [_vp],
    "qm_last_error": [_vp],
    "qm_ctx_memory": [_vp, _c.POINTER(_i64)],
    "qm_mol_create": [_vp, _i, _pi, _pd, _i, _pi, _pi, _pi, _pi, _i, _pd, _pd, _c.POINTER(_vp)],
    "qm_mol_set_dm": [_vp, _vp, _vp, _pd],
    "qm_mol_set_coords": [_vp, _vp, _vp, _pd],
    "qm_mol_nbf": [_vp],
    "qm_mol_stats": [_vp, _c.POINTER(_i64)],
    "qm_mol_destroy": [_vp, _vp],
    "qm_density": [_vp, _vp, _vp, _pg, _i, _i, _i, _vp, _i, _pd],
    "qm_basis_points": [_vp, _vp, _vp, _vp, _i64, _vp],
    "qm_reduce": [_vp, _vp, _i, _i, _vp, _vp, _i64, _pd],
    "qm_reduce_noalloc": [_vp, _vp, _i, _i, _vp, _vp, _i64, _pd],
    "qm_scale": [_vp, _vp, _i, _vp, _i64, _d],
    "qm_seed": [_vp, _vp, _vp, _pg, _i, _vp],
    "qm_multipole_boundary": [_vp, _vp, _pg, _i, _vp, _vp, _pd],
    "qm_jacobi": [_vp, _vp, _pg, _i, _vp, _vp, _vp, _i64, _d, _i, _c.POINTER(_i64), _pd],
    "qm_hartree_fft": [_vp, _vp, _vp, _pg, _i, _i, _vp, _d, _d, _i, _vp],
    "qm_hartree_gto_points": [_vp, _vp, _vp, _vp, _i64, _vp],
    "qm_hartree_gto_grid": [_vp, _vp, _vp, _pg, _i, _vp],
    "qm_vne_classic": [_vp, _vp, _vp, _pg, _i, _i, _vp],
    "qm_vne_diffusion": [_vp, _vp, _vp, _pg, _i, _d, _d, _d, _d, _vp, _vp, _c.POINTER(_i64), _pd, _c.POINTER(_i)],
    "qm_density_batch": [_vp, _vp, _i, _c.POINTER(_vp), _pg, _c.POINTER(_vp), _i, _i, _i],
    "qm_jacobi_batch": [_vp, _vp, _i, _pg, _c.POINTER(_vp), _c.POINTER(_vp), _c.POINTER(_vp), _i, _i64, _i, _i],
    "qm_tune_set": [_vp, _i, _i],
    "qm_tune_get": [_vp, _i, _pi],
    "qm_autotune_density": [_vp, _vp, _vp, _pg, _i, _pd],
    "qm_profile": [_vp, _i],
    "qm_profile_read": [_vp, _i, _pd, _c.POINTER(_i64)],
}
QM_TUNE_SPLIT_VARIANT, QM_TUNE_F64_VARIANT, QM_TUNE_JACOBI_ZS, QM_TUNE_JACOBI_PAIRS, QM_TUNE_TILE_SKIP = 0, 1, 2, 3, 4
QM_K_DENSITY_CONTRACT, QM_K_DENSITY_SPLIT, QM_K_PHI, QM_K_PHI_SPLIT, QM_K_FFT_AXIS, QM_K_GTO, QM_K_JACOBI, QM_K_VNE = range(8)
_RESTYPE = {"qm_last_error": _c.c_char_p}

_lock = threading.Lock()
_lib = None
_ctx = {}


def library():
    """Load libqmgrid.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(f"qmtools: native library {LIB_PATH} is not built "
                                       "(run `python -m qmtools._build` or __graft_entry__.build())")
                # torch first: the library binds to the HIP runtime torch loaded
                _device.torch()
                lib = ctypes.CDLL(LIB_PATH)
                for name, args in SIGNATURES.items():
                    fn = getattr(lib, name)
                    fn.argtypes = args
                    fn.restype = _RESTYPE.get(name, _c.c_int)
                _lib = lib
    return _lib


def exported_symbols():
    return list(SIGNATURES)


def context(device=None):
    """Per-device qm_ctx handle (created on first use)."""
    dev = _device.device_index() if device is None else int(device)
    h = _ctx.get(dev)
    if h is None:
        _device.require()
        lib = library()
        out = _vp()
        rc = lib.qm_ctx_create(dev, _c.byref(out))
        if rc != QM_OK:
            raise QMError(rc, f"cannot create a context on device {dev}")
        h = out.value
        _ctx[dev] = h
    return h


def new_context(device=None):
    """An additional qm_ctx (own workspaces and graph stream), e.g. one per serving lane."""
    dev = _device.device_index() if device is None else int(device)
    _device.require()
    out = _vp()
    rc = library().qm_ctx_create(dev, _c.byref(out))
    if rc != QM_OK:
        raise QMError(rc, f"cannot create a context on device {dev}")
    return out.value


def destroy_context(ctx):
    if ctx:
        library().qm_ctx_destroy(ctx)


def call_ctx(ctx, name, *args):
    """Call an entry point on an explicit context; raise on error."""
    rc = getattr(library(), name)(ctx, *args)
    check(rc, ctx)
    return rc


def check(rc, ctx=None):
    if rc == QM_OK:
        return
    msg = library().qm_last_error(ctx).decode() if ctx else ""
    if rc == QM_ENOCONV:
        raise NotConverged(rc, msg)
    raise QMError(rc, msg)


def call(name, *args):
    """Call an entry point whose first argument is the context; raise on error."""
    ctx = context()
    rc = getattr(library(), name)(ctx, *args)
    check(rc, ctx)
    return rc


def geom(grid_or_tuple):
    if isinstance(grid_or_tuple, tuple):
        nx, ny, nz, h, o = grid_or_tuple
    else:
        nx, ny, nz, h, o = grid_or_tuple.geometry()
    g = Geom()
    g.nx, g.ny, g.nz = int(nx), int(ny), int(nz)
    g.step = float(h)
    g.origin[0], g.origin[1], g.origin[2] = (float(o[0]), float(o[1]), float(o[2]))
    return g


def dptr(t):
    return _vp(int(t.data_ptr()))


def stream():
    return _vp(_device.stream_handle())


def f64p(arr):
    return arr.ctypes.data_as(_pd)


def i32p(arr):
    return arr.ctypes.data_as(_pi)
