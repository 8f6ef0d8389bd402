"""Analytic potential of the GTO density plus nuclei, at arbitrary points (FP64).

Reference `qmtools/cuda/kernel_hartree.cu:288-869`, `qmtools/qmtools.py:2436-2489`:

  V(C) = sum_A Z_A/|C - R_A|  (only |C - R_A| > 1e-8, :851-865)
       - sum_{p>=q} 2 (2 - delta_pq) P_pq  int phi_p(r) phi_q(r) / |r - C| d^3r   (:396-838)

with phi as in the density kernel (solid harmonics S_{L,m} times contracted
Gaussians).  The reference evaluates the integral with hand-derived closed
forms per (La, Lb) class; here the same integral is computed with
McMurchie-Davidson Hermite expansions and the Boys function (Helgaker,
Jorgensen, Olsen, "Molecular Electronic-Structure Theory", sec. 9.5 and 9.9):

  int x_A^i x_B^j ... e^{-a r_A^2} e^{-b r_B^2} / r_C = (2 pi/p) K sum_tuv E^ij_t E^kl_u E^mn_v R_tuv(p, P - C)

`potential_quadrature` computes the same integrals by an independent route
(1/r = 2/sqrt(pi) int_0^inf exp(-t^2 r^2) dt, inner Gaussian integrals by
Gauss-Hermite quadrature) and is used to pin this module.
"""
import math

import numpy
from scipy import special

from .density import C00, C1, C20, C22, C2A, M_ORDER
from .system import lower_symmetric

# Cartesian expansion of the kernel's solid harmonics: (coef, (i, j, k))
CART = {
    (0, 0): [(C00, (0, 0, 0))],
    (1, -1): [(C1, (0, 1, 0))],
    (1, 0): [(C1, (0, 0, 1))],
    (1, 1): [(C1, (1, 0, 0))],
    (2, -2): [(C2A, (1, 1, 0))],
    (2, -1): [(C2A, (0, 1, 1))],
    (2, 0): [(2 * C20, (0, 0, 2)), (-C20, (2, 0, 0)), (-C20, (0, 2, 0))],
    (2, 1): [(C2A, (1, 0, 1))],
    (2, 2): [(C22, (2, 0, 0)), (-C22, (0, 2, 0))],
}


def boys(nmax, T):
    """F_n(T), n = 0..nmax, for an array T >= 0 (incomplete gamma; series near 0)."""
    T = numpy.asarray(T, dtype=numpy.float64)
    out = numpy.empty((nmax + 1,) + T.shape)
    small = T < 1e-3
    Ts = numpy.where(small, 1.0, T)
    for n in range(nmax + 1):
        a = n + 0.5
        big = 0.5 * special.gamma(a) * special.gammainc(a, Ts) / Ts ** a
        # e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)), 12 terms suffice for T < 1e-3
        term = numpy.full_like(T, 1.0 / (2 * n + 1))
        ser = term.copy()
        for k in range(1, 12):
            term = term * 2 * T / (2 * n + 2 * k + 1)
            ser = ser + term
        out[n] = numpy.where(small, numpy.exp(-T) * ser, big)
    return out


def hermite_E(la, lb, a, b, XAB):
    """E[i][j][t] for one axis (Helgaker eq. 9.5.6-9.5.7), without the K prefactor."""
    p = a + b
    mu = a * b / p
    XPA = -b / p * XAB
    XPB = a / p * XAB
    E = numpy.zeros((la + 1, lb + 1, la + lb + 2))
    E[0, 0, 0] = 1.0
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = XPA * E[i - 1, j, t] + (t + 1) * E[i - 1, j, t + 1]
                    if t > 0:
                        v += E[i - 1, j, t - 1] / (2 * p)
                else:
                    v = XPB * E[i, j - 1, t] + (t + 1) * E[i, j - 1, t + 1]
                    if t > 0:
                        v += E[i, j - 1, t - 1] / (2 * p)
                E[i, j, t] = v
    return E, math.exp(-mu * XAB * XAB)


def hermite_R(L, p, PC):
    """R[t,u,v] arrays over points for t+u+v <= L (Helgaker eq. 9.9.18-9.9.20)."""
    X, Y, Z = PC[:, 0], PC[:, 1], PC[:, 2]
    T = p * (X * X + Y * Y + Z * Z)
    F = boys(L, T)
    R = {}
    for n in range(L + 1):
        R[(0, 0, 0, n)] = (-2 * p) ** n * F[n]
    for tot in range(1, L + 1):
        for n in range(L - tot + 1):
            for t in range(tot + 1):
                for u in range(tot - t + 1):
                    v = tot - t - u
                    if t > 0:
                        val = X * R[(t - 1, u, v, n + 1)]
                        if t > 1:
                            val = val + (t - 1) * R[(t - 2, u, v, n + 1)]
                    elif u > 0:
                        val = Y * R[(t, u - 1, v, n + 1)]
                        if u > 1:
                            val = val + (u - 1) * R[(t, u - 2, v, n + 1)]
                    else:
                        val = Z * R[(t, u, v - 1, n + 1)]
                        if v > 1:
                            val = val + (v - 1) * R[(t, u, v - 2, n + 1)]
                    R[(t, u, v, n)] = val
    return {k[:3]: val for k, val in R.items() if k[3] == 0}


def _bf_offsets(system):
    offs, c = [], 0
    for (_, L, _, _) in system.shells:
        offs.append(c)
        c += 2 * L + 1
    return offs


def pair_integrals(system, sa, sb, pts):
    """I[ma, mb, npts] = int phi_(sa,ma) phi_(sb,mb) / |r - C| for shells sa, sb."""
    (A_, La, ala, coa) = system.shells[sa]
    (B_, Lb, alb, cob) = system.shells[sb]
    A = system.coords[A_]
    B = system.coords[B_]
    AB = A - B
    out = numpy.zeros((2 * La + 1, 2 * Lb + 1, pts.shape[0]))
    for a, ca in zip(ala, coa):
        for b, cb in zip(alb, cob):
            p = a + b
            P = (a * A + b * B) / p
            Ex, Kx = hermite_E(La, Lb, a, b, AB[0])
            Ey, Ky = hermite_E(La, Lb, a, b, AB[1])
            Ez, Kz = hermite_E(La, Lb, a, b, AB[2])
            R = hermite_R(La + Lb, p, P - pts)
            pref = ca * cb * Kx * Ky * Kz * 2 * math.pi / p
            for ia, ma in enumerate(M_ORDER[La]):
                for ib, mb in enumerate(M_ORDER[Lb]):
                    acc = 0.0
                    for (fa, (i1, j1, k1)) in CART[(La, ma)]:
                        for (fb, (i2, j2, k2)) in CART[(Lb, mb)]:
                            for t in range(i1 + i2 + 1):
                                for u in range(j1 + j2 + 1):
                                    for v in range(k1 + k2 + 1):
                                        e = Ex[i1, i2, t] * Ey[j1, j2, u] * Ez[k1, k2, v]
                                        if e != 0.0:
                                            acc = acc + fa * fb * e * R[(t, u, v)]
                    out[ia, ib] += pref * acc
    return out


def potential(system, pts, nuclei=True, shells_a=None):
    """Total potential (a.u.) at points pts (npts, 3) bohr.

    shells_a restricts the outer shell index of the pair sum (sa >= sb) to an iterable of
    shells, so that disjoint subsets can be summed by separate processes (with nuclei once).
    """
    pts = numpy.asarray(pts, dtype=numpy.float64).reshape(-1, 3)
    Pt = 2.0 * lower_symmetric(system.dm)
    offs = _bf_offsets(system)
    V = numpy.zeros(pts.shape[0])
    ns = len(system.shells)
    for sa in (range(ns) if shells_a is None else shells_a):
        La = system.shells[sa][1]
        for sb in range(sa + 1):
            Lb = system.shells[sb][1]
            blk = Pt[offs[sa]:offs[sa] + 2 * La + 1, offs[sb]:offs[sb] + 2 * Lb + 1]
            if not numpy.any(blk):
                continue
            I = pair_integrals(system, sa, sb, pts)
            f = 1.0 if sa == sb else 2.0
            V -= f * numpy.einsum("ab,abn->n", blk, I)
    if nuclei:
        for z, R in zip(system.types, system.coords):
            r = numpy.sqrt(numpy.sum((pts - R) ** 2, axis=1))
            V += numpy.where(r > 1e-8, z / numpy.where(r > 1e-8, r, 1.0), 0.0)
    return V


def potential_quadrature(system, pts, nt=96, ngh=24, nuclei=True):
    """Same quantity by brute-force quadrature (independent of the Hermite algebra).

    1/|r-C| = (2/sqrt(pi)) int_0^inf exp(-t^2 |r-C|^2) dt.  For fixed t the
    integrand phi_a phi_b exp(-t^2|r-C|^2) is a polynomial times one Gaussian per
    axis, integrated exactly by ngh-point Gauss-Hermite.  The t integral uses
    t = s/(1-s), Gauss-Legendre in s on (0, 1).
    """
    pts = numpy.asarray(pts, dtype=numpy.float64).reshape(-1, 3)
    xg, wg = numpy.polynomial.hermite.hermgauss(ngh)
    sl, wl = numpy.polynomial.legendre.leggauss(nt)
    s = 0.5 * (sl + 1.0)
    ws = 0.5 * wl
    tq = s / (1.0 - s)
    wt = ws / (1.0 - s) ** 2
    Pt = 2.0 * lower_symmetric(system.dm)
    offs = _bf_offsets(system)
    V = numpy.zeros(pts.shape[0])
    for n, C in enumerate(pts):
        acc = 0.0
        for sa, (A_, La, ala, coa) in enumerate(system.shells):
            for sb, (B_, Lb, alb, cob) in enumerate(system.shells):
                blk = Pt[offs[sa]:offs[sa] + 2 * La + 1, offs[sb]:offs[sb] + 2 * Lb + 1]
                A = system.coords[A_]
                B = system.coords[B_]
                for a, ca in zip(ala, coa):
                    for b, cb in zip(alb, cob):
                        for t, w in zip(tq, wt):
                            g = a + b + t * t
                            Pc = (a * A + b * B + t * t * C) / g
                            K = math.exp(-(a * b * numpy.sum((A - B) ** 2) + a * t * t * numpy.sum((A - C) ** 2)
                                           + b * t * t * numpy.sum((B - C) ** 2)) / g)
                            # nodes per axis: r = Pc + x / sqrt(g)
                            r = Pc[None, :] + xg[:, None] / math.sqrt(g)
                            wx = wg / math.sqrt(g)
                            ra = r - A
                            rb = r - B
                            tot = 0.0
                            for ia, ma in enumerate(M_ORDER[La]):
                                for ib, mb in enumerate(M_ORDER[Lb]):
                                    if blk[ia, ib] == 0.0:
                                        continue
                                    val = 0.0
                                    for (fa, ea) in CART[(La, ma)]:
                                        for (fb, eb) in CART[(Lb, mb)]:
                                            prod = 1.0
                                            for d in range(3):
                                                prod *= numpy.dot(wx, ra[:, d] ** ea[d] * rb[:, d] ** eb[d])
                                            val += fa * fb * prod
                                    tot += blk[ia, ib] * val
                            acc += ca * cb * K * tot * w
        V[n] = -acc * 2.0 / math.sqrt(math.pi)
    if nuclei:
        for z, R in zip(system.types, system.coords):
            r = numpy.sqrt(numpy.sum((pts - R) ** 2, axis=1))
            V += numpy.where(r > 1e-8, z / numpy.where(r > 1e-8, r, 1.0), 0.0)
    return V


def potential_grid(system, nx, ny, nz, h, origin, nuclei=True):
    """Potential at grid nodes origin + (i, j, k) h (kernel_hartree.cu:342-345), flat x-fastest."""
    from .density import grid_points
    return potential(system, grid_points(nx, ny, nz, h, origin, (0.0, 0.0, 0.0)), nuclei)
