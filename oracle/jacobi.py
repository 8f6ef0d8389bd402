"""Jacobi Poisson relaxation on the grid (FP64).

Reference `qmtools/cuda/kernel_hartree.cu:85-171` and `qmtools/qmtools.py:2497-2573`:

  V'(x) = (sum_{6 neighbours} V - rho(x)/h) / 6          (kernel_hartree.cu:156-158)
  V' = 0 on the outer voxel layer, neighbours outside the grid read 0 (:143-147, :161-165)
  rho = seed - Q  (qmtools.py:2525),  V0 = 0 (:2535-2536)
  loop: V->T, T->V; every 100 iterations delta = sum|V - T|; stop when delta <= tol (:2550-2561)

The fixed point solves the 7-point Laplacian  lap_h V = rho / h^3, i.e. V = -phi/(4 pi)
with phi the electrostatic potential (a.u.) of the charges rho.

Multipole Dirichlet boundaries (no reference counterpart): the outer layer is
held at V_b = -phi_mp/(4 pi), where phi_mp is the monopole+dipole+quadrupole
expansion of the charges rho placed at the nodes origin + i h, expanded about
the box centre.  The initial guess then carries V_b on the outer layer.
"""
import numpy
from scipy import fft as sfft


def _view(a, nx, ny, nz):
    return numpy.asarray(a, dtype=numpy.float64).reshape(nz, ny, nx)


def sweep(v, rho, nx, ny, nz, h, bvals=None):
    """One Jacobi sweep; v, rho, bvals flat x-fastest.  Returns the new flat array."""
    V = _view(v, nx, ny, nz)
    R = _view(rho, nx, ny, nz)
    s = numpy.zeros_like(V)
    s[:, :, 1:] += V[:, :, :-1]
    s[:, :, :-1] += V[:, :, 1:]
    s[:, 1:, :] += V[:, :-1, :]
    s[:, :-1, :] += V[:, 1:, :]
    s[1:, :, :] += V[:-1, :, :]
    s[:-1, :, :] += V[1:, :, :]
    out = (s - R / h) / 6.0
    B = numpy.zeros_like(out) if bvals is None else _view(bvals, nx, ny, nz)
    for sl in _faces():
        out[sl] = B[sl]
    return out.ravel()


def _faces():
    return [(0, slice(None), slice(None)), (-1, slice(None), slice(None)),
            (slice(None), 0, slice(None)), (slice(None), -1, slice(None)),
            (slice(None), slice(None), 0), (slice(None), slice(None), -1)]


def boundary_mask(nx, ny, nz):
    m = numpy.zeros((nz, ny, nx), dtype=bool)
    for sl in _faces():
        m[sl] = True
    return m.ravel()


def jacobi(rho, nx, ny, nz, h, nsweeps, bvals=None, v0=None):
    """nsweeps plain sweeps from v0 (default: 0 inside, bvals on the boundary)."""
    if v0 is None:
        v = numpy.zeros(nx * ny * nz) if bvals is None else numpy.where(
            boundary_mask(nx, ny, nz), bvals, 0.0)
    else:
        v = numpy.asarray(v0, dtype=numpy.float64).copy()
    for _ in range(int(nsweeps)):
        v = sweep(v, rho, nx, ny, nz, h, bvals)
    return v


def jacobi_closed_form(rho, nx, ny, nz, h, nsweeps, bvals=None):
    """The same m = nsweeps sweeps from 0 (boundary held at bvals) without iterating.

    On the interior the sweep is V' = G V + c with G = (sum of the 6 neighbours) / 6
    under zero Dirichlet conditions and c = (held boundary neighbours - rho / h) / 6.
    DST-I diagonalises G, eigenvalue mu = (sum_axes 2 cos(pi k / (n_axis - 1))) / 6, so the
    m-th iterate from V = 0 is V_m = sum_{j<m} G^j c = IDST(DST(c) (1 - mu^m) / (1 - mu)).
    Costs a few transforms for any sweep count (2000 sweeps on 96^3 in seconds).
    """
    R = _view(rho, nx, ny, nz)
    B = numpy.zeros_like(R) if bvals is None else numpy.where(
        boundary_mask(nx, ny, nz).reshape(nz, ny, nx), _view(bvals, nx, ny, nz), 0.0)
    s = numpy.zeros_like(B)
    s[:, :, 1:] += B[:, :, :-1]
    s[:, :, :-1] += B[:, :, 1:]
    s[:, 1:, :] += B[:, :-1, :]
    s[:, :-1, :] += B[:, 1:, :]
    s[1:, :, :] += B[:-1, :, :]
    s[:-1, :, :] += B[1:, :, :]
    c = (s[1:-1, 1:-1, 1:-1] - R[1:-1, 1:-1, 1:-1] / h) / 6.0
    mz, my, mx = c.shape
    mu = numpy.zeros_like(c)
    for ax, m in zip((2, 1, 0), (mx, my, mz)):
        shape = [1, 1, 1]
        shape[ax] = m
        mu = mu + (2.0 * numpy.cos(numpy.pi * numpy.arange(1, m + 1) / (m + 1))).reshape(shape)
    mu = mu / 6.0
    vh = sfft.dstn(c, type=1) * (1.0 - mu ** int(nsweeps)) / (1.0 - mu)
    out = B.copy()
    out[1:-1, 1:-1, 1:-1] = sfft.idstn(vh, type=1)
    return out.ravel()


def hartree_iter(rho, nx, ny, nz, h, tolerance=1e-24, bvals=None, max_iterations=None):
    """The reference convergence loop (qmtools.py:2547-2561); returns (V, iterations, delta)."""
    v = numpy.zeros(nx * ny * nz) if bvals is None else numpy.where(
        boundary_mask(nx, ny, nz), bvals, 0.0)
    delta = tolerance + 1
    itr = 1
    while delta > tolerance:
        t = sweep(v, rho, nx, ny, nz, h, bvals)
        v = sweep(t, rho, nx, ny, nz, h, bvals)
        if itr % 100 == 0:
            delta = float(numpy.sum(numpy.abs(v - t)))
        if max_iterations is not None and itr >= max_iterations:
            break
        itr += 1
    return v, itr, delta


def moments(rho, nx, ny, nz, h, origin):
    """(q, p[3], Qtraceless[3,3], centre) of the node charges rho."""
    from .density import grid_points
    pts = grid_points(nx, ny, nz, h, origin, (0.0, 0.0, 0.0))
    c = numpy.asarray(origin, float) + 0.5 * h * (numpy.array([nx, ny, nz]) - 1)
    d = pts - c
    r = numpy.asarray(rho, float)
    q = r.sum()
    p = d.T @ r
    d2 = numpy.einsum("ij,ij->i", d, d)
    Q = 3.0 * (d.T * r) @ d - numpy.eye(3) * (r @ d2)
    return q, p, Q, c


def multipole_boundary(rho, nx, ny, nz, h, origin):
    """V_b on the outer layer (zero elsewhere), flat x-fastest."""
    from .density import grid_points
    q, p, Q, c = moments(rho, nx, ny, nz, h, origin)
    mask = boundary_mask(nx, ny, nz)
    pts = grid_points(nx, ny, nz, h, origin, (0.0, 0.0, 0.0))[mask]
    R = pts - c
    r2 = numpy.einsum("ij,ij->i", R, R)
    r = numpy.sqrt(r2)
    phi = q / r + (R @ p) / (r2 * r) + 0.5 * numpy.einsum("ij,jk,ik->i", R, Q, R) / (r2 * r2 * r)
    out = numpy.zeros(nx * ny * nz)
    out[mask] = -phi / (4.0 * numpy.pi)
    return out
