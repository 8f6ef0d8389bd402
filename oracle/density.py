"""Basis functions on points and the density grid (FP64).

Follows reference `qmtools/cuda/kernel_density.cu:14-155` and
`qmtools/qmtools.py:2381-2430`:

  phi_p(r) = S_{L,m}(r - R_A) * sum_k c_k exp(-alpha_k |r - R_A|^2)      (:104-131)
  Q(v)     = 2 h^3 n^-3 sum_{s in n^3} sum_{p>=q} (2 - delta_pq) P_pq phi_p(r_s) phi_q(r_s)
  r_s      = origin + (i + (s + 1/2)/n) h   per axis                       (:96-100)
  Q       <- Q * (sum Z / sum Q)                                          (qmtools.py:2420-2424)

Output layout is flat x-fastest: index i + nx*(j + ny*k) (:151-154).
"""
import numpy

from .system import lower_symmetric

# Real solid harmonic prefactors (kernel_density.cu:17-28)
C00 = 0.28209479177387814
C1 = 0.4886025119029199
C2A = 1.0925484305920792
C20 = 0.31539156525252005
C22 = 0.5462742152960396


def solid_harmonic(L, m, x, y, z):
    """r^L Y_Lm as in kernel_density.cu:14-31 (0 for anything else)."""
    if L == 0:
        return numpy.full_like(x, C00)
    if L == 1:
        return C1 * {-1: y, 0: z, 1: x}[m]
    if L == 2:
        if m == -2:
            return C2A * x * y
        if m == -1:
            return C2A * z * y
        if m == 0:
            return C20 * (2 * z * z - x * x - y * y)
        if m == 1:
            return C2A * x * z
        if m == 2:
            return C22 * (x * x - y * y)
    return numpy.zeros_like(x)


M_ORDER = {0: (0,), 1: (0, 1, -1), 2: (0, 1, -1, 2, -2)}


def basis_on_points(system, pts):
    """Phi[npts, nbf] in FP64 for points pts (npts, 3) in bohr."""
    pts = numpy.asarray(pts, dtype=numpy.float64)
    out = numpy.empty((pts.shape[0], system.nbf))
    c = 0
    for (a, L, al, co) in system.shells:
        d = pts - system.coords[a]
        r2 = numpy.einsum("ij,ij->i", d, d)
        radial = numpy.exp(-numpy.outer(r2, al)) @ co
        for m in M_ORDER[L]:
            out[:, c] = solid_harmonic(L, m, d[:, 0], d[:, 1], d[:, 2]) * radial
            c += 1
    return out


def grid_points(nx, ny, nz, h, origin, offset=(0.5, 0.5, 0.5)):
    """All grid positions in flat x-fastest order, r = origin + (i + offset) h."""
    i = numpy.arange(nx) + offset[0]
    j = numpy.arange(ny) + offset[1]
    k = numpy.arange(nz) + offset[2]
    K, J, I = numpy.meshgrid(k, j, i, indexing="ij")
    pts = numpy.stack([I.ravel(), J.ravel(), K.ravel()], axis=1) * h
    return pts + numpy.asarray(origin, dtype=numpy.float64)


def rho_on_points(system, pts, chunk=32768):
    """sum_pq Pl_pq phi_p phi_q at points (alpha-spin P, no factor 2)."""
    P = lower_symmetric(system.dm)
    out = numpy.empty(pts.shape[0])
    for s in range(0, pts.shape[0], chunk):
        ph = basis_on_points(system, pts[s:s + chunk])
        out[s:s + chunk] = numpy.einsum("ij,ij->i", ph @ P, ph)
    return out


def density_grid_raw(system, nx, ny, nz, h, origin, subgrid=1, chunk=32768):
    """Un-normalised electrons per voxel, 2 h^3 <rho>_subgrid (kernel_density.cu:74-147)."""
    n = int(subgrid)
    acc = numpy.zeros(nx * ny * nz)
    for sx in range(n):
        for sy in range(n):
            for sz in range(n):
                off = ((sx + 0.5) / n, (sy + 0.5) / n, (sz + 0.5) / n)
                pts = grid_points(nx, ny, nz, h, origin, off)
                acc += rho_on_points(system, pts, chunk)
    return 2.0 * acc * h ** 3 / n ** 3


def density_grid(system, nx, ny, nz, h, origin, subgrid=1, chunk=32768):
    """Normalised density grid as returned by QMTools.Compute_density (qmtools.py:2381-2430)."""
    q = density_grid_raw(system, nx, ny, nz, h, origin, subgrid, chunk)
    tot = numpy.sum(q)
    return q * (system.qtot / tot), tot
