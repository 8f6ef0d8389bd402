"""Batched density + Hartree for many molecules of uneven size.

One call handles a list of (molecule, grid) pairs:
  * density: `qm_density_batch` (fused launches over packed chunks of all grids);
  * Hartree: seed - density per molecule, optional multipole boundaries, then
    `qm_jacobi_batch` (one launch per sweep for the whole batch, fixed blocks of
    sweeps captured in a hipGraph and replayed).
A `Lane` is one native context plus one torch stream; lanes are independent,
so batches on different lanes overlap on the device.
No reference counterpart: the reference processes one molecule per call.
"""
from __future__ import annotations

import ctypes

import numpy

from . import _device
from . import _native as nat
from .qmtools import device_molecule

_PREC = {"fp64": nat.QM_PREC_FP64, "mixed": nat.QM_PREC_F16X3, "f16": nat.QM_PREC_F16X3,
         "bf16": nat.QM_PREC_BF16X3}


class Lane:
    """A native context and a stream (default priority) for one stream of batches."""

    def __init__(self, device=None, default=False):
        t = _device.torch()
        dev = _device.require()
        if default:
            self.ctx = nat.context(device)
            self.stream = t.cuda.current_stream(dev)
            self._own = False
        else:
            self.ctx = nat.new_context(device)
            self.stream = t.cuda.Stream(dev)
            self._own = True

    def close(self):
        if self._own and self.ctx:
            self.stream.synchronize()
            nat.destroy_context(self.ctx)
        self.ctx = None

    @property
    def handle(self):
        return ctypes.c_void_p(self.stream.cuda_stream)


def _arrays(mols, grids):
    n = len(mols)
    handles = (ctypes.c_void_p * n)(*[device_molecule(m).handle for m in mols])
    geoms = (nat.Geom * n)(*[nat.geom(g) for g in grids])
    return handles, geoms


def density_batch(mols, grids, precision="mixed", lane=None, normalize=True):
    """Device densities (flat, x-fastest) of every molecule on its grid."""
    if len(mols) != len(grids):
        raise ValueError("one grid per molecule")
    lane = lane or Lane(default=True)
    t = _device.torch()
    f64 = precision == "fp64"
    dt = t.float64 if f64 else t.float32
    with t.cuda.stream(lane.stream):
        outs = [t.empty(int(g.npts), dtype=dt, device=_device.require()) for g in grids]
    handles, geoms = _arrays(mols, grids)
    optr = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
    nat.call_ctx(lane.ctx, "qm_density_batch", lane.handle, len(mols), handles, geoms, optr, _PREC[precision],
                 nat.QM_F64 if f64 else nat.QM_F32, 1 if normalize else 0)
    return outs


def jacobi_prepare(mols, grids, densities, precision="mixed", boundary="multipole", lane=None):
    """Per-molecule Jacobi state on the lane's stream: rho = seed - density, V holding the boundary
    (zero or multipole) and, in FP64, the sweep partner.  Returns one (rho, v, tmp) per molecule."""
    lane = lane or Lane(default=True)
    t = _device.torch()
    dev = _device.require()
    states = []
    with t.cuda.stream(lane.stream):
        for m, g, q in zip(mols, grids, densities):
            gg = nat.geom(g)
            seed = t.zeros(int(g.npts), dtype=t.float64, device=dev)
            nat.call_ctx(lane.ctx, "qm_seed", device_molecule(m).handle, lane.handle, ctypes.byref(gg), nat.QM_F64,
                         nat.dptr(seed))
            rho = seed - q.double()
            v = t.zeros(int(g.npts), dtype=t.float64, device=dev)
            if boundary == "multipole":
                nat.call_ctx(lane.ctx, "qm_multipole_boundary", lane.handle, ctypes.byref(gg), nat.QM_F64,
                             nat.dptr(rho), nat.dptr(v), None)
            elif boundary != "zero":
                raise ValueError("boundary must be 'zero' or 'multipole'")
            states.append((rho, v, v.clone() if precision == "fp64" else None))
    return states


def jacobi_block(precision):
    """Sweeps per replayed block: an FP64 block ends with the result in V, an FP32 block starts from
    the FP64 residual of V and ends by adding its correction to V, so V is the whole state between
    blocks and sweeps run in any number of calls of whole blocks give the same bits as one call."""
    return 100 if precision == "fp64" else 50


def jacobi_sweeps(grids, states, sweeps, precision="mixed", lane=None, use_graph=True):
    """`sweeps` Jacobi sweeps on every (rho, v, tmp) state, in one batched launch sequence."""
    lane = lane or Lane(default=True)
    n = len(states)
    geoms = (nat.Geom * n)(*[nat.geom(g) for g in grids])
    rp = (ctypes.c_void_p * n)(*[st[0].data_ptr() for st in states])
    vp = (ctypes.c_void_p * n)(*[st[1].data_ptr() for st in states])
    tp = (ctypes.c_void_p * n)(*[(st[2].data_ptr() if st[2] is not None else 0) for st in states])
    prec = nat.QM_PREC_FP64 if precision == "fp64" else nat.QM_PREC_FP32
    nat.call_ctx(lane.ctx, "qm_jacobi_batch", lane.handle, n, geoms, rp, vp, tp, prec, int(sweeps),
                 jacobi_block(precision), 1 if use_graph else 0)


def hartree_jacobi_batch(mols, grids, densities, sweeps=2000, precision="mixed", boundary="multipole", lane=None,
                         use_graph=True):
    """V = -phi/(4 pi) after `sweeps` Jacobi sweeps for every molecule (FP64 device arrays)."""
    lane = lane or Lane(default=True)
    states = jacobi_prepare(mols, grids, densities, precision, boundary, lane)
    jacobi_sweeps(grids, states, sweeps, precision, lane, use_graph)
    # keep the inputs alive until the stream has consumed them
    lane._keep = [(st[0], st[2]) for st in states]
    return [st[1] for st in states]


def hartree_fft_batch(mols, grids, densities, precision="mixed", mode="freespace", sigma=0.2, cutoff=6.0,
                      lane=None):
    """FFT Poisson potential of electrons + Gaussian nuclei for every molecule: volts, flat device arrays
    in the C order V[i, j, k] that `QMTools.Compute_hartree_fft` returns (reference qmtools.py:2637-2639).

    The transforms run one molecule after the other on the lane's stream; the
    context caches the twiddle tables and free-space Green's-function transforms
    per grid shape, so molecules on the same grid reuse them.
    """
    lane = lane or Lane(default=True)
    t = _device.torch()
    dev = _device.require()
    f64 = precision == "fp64"
    dt = t.float64 if f64 else t.float32
    m = {"periodic": nat.QM_FFT_PERIODIC, "freespace": nat.QM_FFT_FREESPACE}[mode]
    outs = []
    with t.cuda.stream(lane.stream):
        for mol, g, q in zip(mols, grids, densities):
            q = q.to(dt)
            out = t.empty(int(g.npts), dtype=dt, device=dev)
            nat.call_ctx(lane.ctx, "qm_hartree_fft", device_molecule(mol).handle, lane.handle,
                         ctypes.byref(nat.geom(g)), nat.QM_PREC_FP64 if f64 else nat.QM_PREC_FP32,
                         nat.QM_F64 if f64 else nat.QM_F32, nat.dptr(q), float(sigma), float(cutoff), m,
                         nat.dptr(out))
            outs.append(out)
    return outs


def run(mols, grids, precision="mixed", sweeps=2000, boundary="multipole", lane=None, use_graph=True,
        poisson="jacobi"):
    """Density + Hartree for a batch; returns (densities, potentials) as device tensors.

    poisson="jacobi": V = -phi/(4 pi) after `sweeps` sweeps (FP64 arrays);
    poisson="fft": free-space FFT potential in volts (FP64 or FP32 arrays, C order V[i, j, k]).
    """
    lane = lane or Lane(default=True)
    qs = density_batch(mols, grids, precision, lane)
    if poisson == "fft":
        vs = hartree_fft_batch(mols, grids, qs, precision, lane=lane)
    elif poisson == "jacobi":
        vs = hartree_jacobi_batch(mols, grids, qs, sweeps, precision, boundary, lane, use_graph)
    else:
        raise ValueError("poisson must be 'jacobi' or 'fft'")
    return qs, vs


def grid_points(grids):
    return int(numpy.sum([int(g.npts) for g in grids]))
