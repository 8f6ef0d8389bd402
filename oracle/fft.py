"""FFT Poisson solve of the total charge (FP64, Angstrom / volt units).

Reference `qmtools/cuda/kernel_fft.cu:23-106`, `qmtools/qmtools.py:2578-2639`:

  rho_tot(r) = -Q/h^3 + sum_A Z_A (2 pi s^2)^-3/2 exp(-r^2 / 2 s^2)   [r^2 <= (cutoff s)^2]
  rhs        = -rho_tot / eps0,  eps0 = 0.0055263494 e/(V A)          (kernel_fft.cu:18,57-77)
  voxel centre (i + 1/2) h_A + origin_A; nuclei at coords * 0.529177211 (:42-54)
  rhs is read x-fastest and written C order [i, j, k]                  (:57, :77)
  periodic: V^ = rhs^ * (-1 / (4 pi^2 |k|^2)), k in cycles/A, V^(0) = 0 (:82-106)
  returned: Re(ifft(V^)) as [nx, ny, nz]                               (qmtools.py:2634-2639)

Free-space mode (no reference counterpart): Hockney-Eastwood convolution of the
voxel charges q_j = -rhs_j eps0 h^3 with G(d) = 1/(4 pi eps0 |d|), G(0) =
C0/(4 pi eps0 h), C0 = integral over the unit cube of 1/|u| (the self-cell
average), evaluated exactly by zero padding to 2n per axis.
"""
import math

import numpy

R_BOHR = 0.529177211
EPS0 = 0.0055263494
ANG2BOR = 1.8897259886


def cube_self_potential():
    """C0 = int_{[-1/2,1/2]^3} d^3u / |u| (closed form, see cube_self_potential_check)."""
    # 8 * (1/2)^2 * I1 with I1 = int_{[0,1]^3} d^3u/|u| = 3 ln(1+sqrt3) - 3 ln(2)/2 - pi/4 ... written
    # through the standard result I1 = 3*asinh(1/sqrt2) - pi/4
    i1 = 3.0 * math.asinh(1.0 / math.sqrt(2.0)) - math.pi / 4.0
    return 2.0 * i1


def total_charge_rhs(q, types, coords_bohr, nx, ny, nz, h_bohr, origin_bohr,
                     sigma=0.2, cutoff=6.0):
    """rhs in C order [i, j, k] (kernel_fft.cu:23-79)."""
    step = h_bohr / ANG2BOR
    o = numpy.asarray(origin_bohr, float) / ANG2BOR
    atoms = numpy.asarray(coords_bohr, float) * R_BOHR
    Qxyz = numpy.asarray(q, float).reshape(nz, ny, nx).transpose(2, 1, 0)
    rho = Qxyz / (-step ** 3)
    xs = o[0] + (numpy.arange(nx) + 0.5) * step
    ys = o[1] + (numpy.arange(ny) + 0.5) * step
    zs = o[2] + (numpy.arange(nz) + 0.5) * step
    X, Y, Z = numpy.meshgrid(xs, ys, zs, indexing="ij")
    norm = 1.0 / (sigma ** 3 * 2 * math.pi * math.sqrt(2 * math.pi))
    cut2 = (cutoff * sigma) ** 2
    for z, a in zip(types, atoms):
        r2 = (X - a[0]) ** 2 + (Y - a[1]) ** 2 + (Z - a[2]) ** 2
        rho = rho + numpy.where(r2 <= cut2, float(z) * norm * numpy.exp(-r2 / (2 * sigma * sigma)), 0.0)
    return -rho / EPS0


def poisson_periodic(rhs, h_ang):
    """kernel_fft.cu:82-106 with numpy's FFT (FP64)."""
    nx, ny, nz = rhs.shape
    kx = numpy.fft.fftfreq(nx, d=h_ang)
    ky = numpy.fft.fftfreq(ny, d=h_ang)
    kz = numpy.fft.fftfreq(nz, d=h_ang)
    K2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    with numpy.errstate(divide="ignore"):
        scale = -1.0 / (4.0 * math.pi ** 2 * K2)
    scale[0, 0, 0] = 0.0
    return numpy.fft.ifftn(numpy.fft.fftn(rhs) * scale).real


def green_padded(nx, ny, nz, h_ang):
    """Free-space kernel G(d)*h^3 on the 2n cyclic grid (volts per (e/A^3) of rhs*(-eps0))."""
    def axis(n):
        i = numpy.arange(2 * n)
        return numpy.where(i <= n, i, i - 2 * n) * h_ang
    X, Y, Z = numpy.meshgrid(axis(nx), axis(ny), axis(nz), indexing="ij")
    r = numpy.sqrt(X * X + Y * Y + Z * Z)
    with numpy.errstate(divide="ignore"):
        g = 1.0 / r
    g[0, 0, 0] = cube_self_potential() / h_ang
    return g * h_ang ** 3 / (4.0 * math.pi * EPS0)


def poisson_freespace(rhs, h_ang):
    """V = sum_j G(x_i - x_j) q_j with q_j = -rhs_j eps0 (charge density e/A^3 times h^3 in G)."""
    nx, ny, nz = rhs.shape
    g = green_padded(nx, ny, nz, h_ang)
    pad = numpy.zeros((2 * nx, 2 * ny, 2 * nz))
    pad[:nx, :ny, :nz] = -rhs * EPS0
    v = numpy.fft.ifftn(numpy.fft.fftn(pad) * numpy.fft.fftn(g)).real
    return v[:nx, :ny, :nz]


def poisson_freespace_direct(rhs, h_ang):
    """O(N^2) direct sum of the same convolution (small grids only)."""
    nx, ny, nz = rhs.shape
    q = (-rhs * EPS0).ravel() * h_ang ** 3 / (4.0 * math.pi * EPS0)
    I, J, K = numpy.meshgrid(numpy.arange(nx), numpy.arange(ny), numpy.arange(nz), indexing="ij")
    P = numpy.stack([I.ravel(), J.ravel(), K.ravel()], 1) * h_ang
    out = numpy.empty(P.shape[0])
    c0 = cube_self_potential() / h_ang
    for s in range(0, P.shape[0], 512):
        d = numpy.sqrt(((P[s:s + 512, None, :] - P[None, :, :]) ** 2).sum(-1))
        with numpy.errstate(divide="ignore"):
            g = numpy.where(d > 0, 1.0 / numpy.where(d > 0, d, 1.0), c0)
        out[s:s + 512] = g @ q
    return out.reshape(nx, ny, nz)


def hartree_fft(q, types, coords_bohr, nx, ny, nz, h_bohr, origin_bohr, sigma=0.2, cutoff=6.0,
                mode="periodic"):
    rhs = total_charge_rhs(q, types, coords_bohr, nx, ny, nz, h_bohr, origin_bohr, sigma, cutoff)
    h_ang = h_bohr / ANG2BOR
    if mode == "periodic":
        return poisson_periodic(rhs, h_ang)
    if mode == "freespace":
        return poisson_freespace(rhs, h_ang)
    raise ValueError(mode)
