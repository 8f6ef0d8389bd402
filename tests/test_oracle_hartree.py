"""Pins of the oracle's Jacobi, FFT, GTO and V_ne functions (CPU)."""
import math

import numpy
import pytest
from scipy import fft as sfft
from scipy import special

from oracle import fft as off
from oracle import gto as og
from oracle import jacobi as oj
from oracle import vne as ov
from oracle.density import C00, M_ORDER, solid_harmonic
from oracle.system import lower_symmetric, make_system

# ----------------------------------------------------------------------------- Jacobi


def _dst_jacobi(c_int, m_sweeps):
    """Closed form of m Jacobi sweeps from 0 for V' = G V + c on the interior (DST-I diagonalises G)."""
    mz, my, mx = c_int.shape
    ch = sfft.dstn(c_int, type=1)
    mu = numpy.zeros_like(ch)
    for ax, m in zip((2, 1, 0), (mx, my, mz)):
        k = numpy.arange(1, m + 1)
        shape = [1, 1, 1]
        shape[ax] = m
        mu = mu + (2 * numpy.cos(numpy.pi * k / (m + 1))).reshape(shape)
    mu = mu / 6.0
    vh = ch * (1 - mu ** m_sweeps) / (1 - mu)
    return sfft.idstn(vh, type=1)


def test_jacobi_zero_dirichlet_closed_form():
    rng = numpy.random.default_rng(0)
    nx, ny, nz, h = 16, 24, 12, 0.3
    rho = rng.standard_normal(nx * ny * nz)
    v = oj.jacobi(rho, nx, ny, nz, h, 37)
    R = rho.reshape(nz, ny, nx)
    c = -R[1:-1, 1:-1, 1:-1] / (6 * h)
    want = _dst_jacobi(c, 37)
    V = v.reshape(nz, ny, nx)
    numpy.testing.assert_allclose(V[1:-1, 1:-1, 1:-1], want, rtol=0, atol=1e-12 * numpy.abs(want).max())
    assert numpy.all(oj.boundary_mask(nx, ny, nz).reshape(nz, ny, nx)[0])
    assert numpy.all(V[0] == 0) and numpy.all(V[:, :, -1] == 0)


def test_jacobi_fixed_boundary_closed_form():
    """Held boundary values enter as a source term on the interior neighbours."""
    rng = numpy.random.default_rng(1)
    nx, ny, nz, h = 12, 10, 14, 0.25
    rho = rng.standard_normal(nx * ny * nz)
    B = numpy.where(oj.boundary_mask(nx, ny, nz), rng.standard_normal(nx * ny * nz), 0.0)
    v = oj.jacobi(rho, nx, ny, nz, h, 23, bvals=B)
    R = rho.reshape(nz, ny, nx)
    Bv = B.reshape(nz, ny, nx)
    s = numpy.zeros_like(Bv)
    s[:, :, 1:] += Bv[:, :, :-1]
    s[:, :, :-1] += Bv[:, :, 1:]
    s[:, 1:, :] += Bv[:, :-1, :]
    s[:, :-1, :] += Bv[:, 1:, :]
    s[1:, :, :] += Bv[:-1, :, :]
    s[:-1, :, :] += Bv[1:, :, :]
    c = (s[1:-1, 1:-1, 1:-1] - R[1:-1, 1:-1, 1:-1] / h) / 6.0
    want = _dst_jacobi(c, 23)
    V = v.reshape(nz, ny, nx)
    numpy.testing.assert_allclose(V[1:-1, 1:-1, 1:-1], want, atol=1e-12 * numpy.abs(want).max())
    numpy.testing.assert_array_equal(V[0], Bv[0])


def test_jacobi_convergence_loop_reaches_fixed_point():
    rng = numpy.random.default_rng(2)
    nx, ny, nz, h = 8, 8, 8, 0.5
    rho = rng.standard_normal(nx * ny * nz)
    v, itr, delta = oj.hartree_iter(rho, nx, ny, nz, h, tolerance=1e-10)
    assert itr % 100 == 1 or delta <= 1e-10
    R = rho.reshape(nz, ny, nx)
    fixed = _dst_jacobi(-R[1:-1, 1:-1, 1:-1] / (6 * h), 10 ** 9)
    numpy.testing.assert_allclose(v.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1], fixed, atol=1e-9)


def test_multipole_moments_single_charge():
    nx, ny, nz, h = 16, 16, 16, 0.2
    origin = numpy.array([-1.0, 0.5, 2.0])
    rho = numpy.zeros(nx * ny * nz)
    i, j, k = 5, 11, 8
    rho[i + nx * (j + ny * k)] = 2.5
    q, p, Q, c = oj.moments(rho, nx, ny, nz, h, origin)
    d = origin + h * numpy.array([i, j, k]) - c
    assert q == 2.5
    numpy.testing.assert_allclose(p, 2.5 * d, rtol=1e-14)
    numpy.testing.assert_allclose(Q, 2.5 * (3 * numpy.outer(d, d) - d @ d * numpy.eye(3)), rtol=1e-13, atol=1e-14)


def test_multipole_boundary_vs_direct_sum():
    """Compact charges near the centre: expansion error is octupole order, (a/R)^3 of the potential."""
    rng = numpy.random.default_rng(3)
    nx = ny = nz = 32
    h = 0.25
    origin = numpy.zeros(3)
    rho = numpy.zeros(nx * ny * nz)
    for _ in range(12):
        i, j, k = rng.integers(14, 18, size=3)
        rho[i + nx * (j + ny * k)] += rng.uniform(0.1, 1.0)
    b = oj.multipole_boundary(rho, nx, ny, nz, h, origin)
    mask = oj.boundary_mask(nx, ny, nz)
    from oracle.density import grid_points
    pts = grid_points(nx, ny, nz, h, origin, (0, 0, 0))
    src = numpy.nonzero(rho)[0]
    direct = -numpy.array([numpy.sum(rho[src] / numpy.linalg.norm(pts[src] - x, axis=1))
                           for x in pts[mask]]) / (4 * math.pi)
    err = numpy.abs(b[mask] - direct).max() / numpy.abs(direct).max()
    assert err < 5 * (3 * h * math.sqrt(3) / (16 * h)) ** 3
    assert numpy.all(b[~mask] == 0)


def test_jacobi_point_charge_far_field():
    """V = -phi/(4 pi): a unit charge at the centre, multipole boundary, relaxed to the fixed point."""
    n, h = 24, 0.5
    rho = numpy.zeros(n ** 3)
    c = n // 2
    rho[c + n * (c + n * c)] = 1.0
    B = oj.multipole_boundary(rho, n, n, n, h, numpy.zeros(3))
    R = rho.reshape(n, n, n)
    Bv = B.reshape(n, n, n)
    s = numpy.zeros_like(Bv)
    s[:, :, 1:] += Bv[:, :, :-1]
    s[:, :, :-1] += Bv[:, :, 1:]
    s[:, 1:, :] += Bv[:, :-1, :]
    s[:, :-1, :] += Bv[:, 1:, :]
    s[1:, :, :] += Bv[:-1, :, :]
    s[:-1, :, :] += Bv[1:, :, :]
    fixed = _dst_jacobi((s[1:-1, 1:-1, 1:-1] - R[1:-1, 1:-1, 1:-1] / h) / 6.0, 10 ** 9)
    # along +x from the charge, 6 nodes away
    val = fixed[c - 1, c - 1, c - 1 + 6]
    assert val == pytest.approx(-1.0 / (4 * math.pi * 6 * h), rel=2e-2)


# ----------------------------------------------------------------------------- FFT


def test_total_charge_rhs_bruteforce():
    rng = numpy.random.default_rng(4)
    nx, ny, nz = 8, 6, 4
    hb = 0.35
    origin = numpy.array([-1.1, -0.7, -0.4])
    q = rng.standard_normal(nx * ny * nz)
    types = [6, 1]
    coords = numpy.array([[0.0, 0.1, 0.0], [0.4, -0.3, 0.2]])
    sigma, cutoff = 0.3, 2.0
    rhs = off.total_charge_rhs(q, types, coords, nx, ny, nz, hb, origin, sigma, cutoff)
    st = hb / off.ANG2BOR
    for (i, j, k) in [(0, 0, 0), (7, 5, 3), (3, 2, 1), (4, 3, 2)]:
        x = origin / off.ANG2BOR + (numpy.array([i, j, k]) + 0.5) * st
        v = -q[i + nx * j + nx * ny * k] / st ** 3
        for z, a in zip(types, coords * off.R_BOHR):
            r2 = numpy.sum((x - a) ** 2)
            if r2 <= (cutoff * sigma) ** 2:
                v += z / (sigma ** 3 * 2 * math.pi * math.sqrt(2 * math.pi)) * math.exp(-r2 / (2 * sigma ** 2))
        assert rhs[i, j, k] == pytest.approx(-v / off.EPS0, rel=1e-13)


def test_gaussian_nucleus_integrates_to_Z():
    n = 64
    st = 0.05
    hb = st * off.ANG2BOR
    origin = -numpy.full(3, 0.5 * n * hb)
    rhs = off.total_charge_rhs(numpy.zeros(n ** 3), [8], [[0.0, 0.0, 0.0]], n, n, n, hb, origin, 0.2, 6)
    # the Gaussian is cut at 6 sigma: retained fraction erf(6/sqrt2) - sqrt(2/pi) 6 e^-18
    kept = math.erf(6 / math.sqrt(2)) - math.sqrt(2 / math.pi) * 6 * math.exp(-18)
    assert numpy.sum(-rhs * off.EPS0) * st ** 3 == pytest.approx(8.0 * kept, rel=1e-9)


def test_periodic_poisson_single_mode():
    nx, ny, nz, h = 16, 8, 24, 0.1
    x = numpy.arange(nx) * h
    z = numpy.arange(nz) * h
    kx, kz = 3 / (nx * h), 2 / (nz * h)
    rhs = numpy.cos(2 * math.pi * kx * x)[:, None, None] * numpy.sin(2 * math.pi * kz * z)[None, None, :] \
        * numpy.ones((1, ny, 1))
    v = off.poisson_periodic(rhs, h)
    numpy.testing.assert_allclose(v, -rhs / (4 * math.pi ** 2 * (kx ** 2 + kz ** 2)), atol=1e-14)


def test_cube_self_potential_value():
    assert off.cube_self_potential() == pytest.approx(2.3800773639795532, rel=1e-15)


def test_freespace_fft_equals_direct_sum():
    rng = numpy.random.default_rng(6)
    rhs = rng.standard_normal((8, 6, 5))
    a = off.poisson_freespace(rhs, 0.3)
    b = off.poisson_freespace_direct(rhs, 0.3)
    numpy.testing.assert_allclose(a, b, atol=1e-12 * numpy.abs(b).max())


def test_freespace_gaussian_far_field():
    n = 48
    st = 0.1
    hb = st * off.ANG2BOR
    origin = -numpy.full(3, 0.5 * n * hb)
    rhs = off.total_charge_rhs(numpy.zeros(n ** 3), [1], [[0.0, 0.0, 0.0]], n, n, n, hb, origin, 0.3, 6)
    v = off.poisson_freespace(rhs, st)
    # voxel centre (i+1/2) st - n st/2; take i = n-1 on the x axis row nearest the nucleus
    i, j = n - 1, n // 2
    pos = numpy.array([(i + 0.5) * st - 0.5 * n * st, (j + 0.5) * st - 0.5 * n * st, (j + 0.5) * st - 0.5 * n * st])
    r = numpy.linalg.norm(pos)
    want = special.erf(r / (math.sqrt(2) * 0.3)) / (4 * math.pi * off.EPS0 * r)
    assert v[i, j, j] == pytest.approx(want, rel=1e-5)


# ----------------------------------------------------------------------------- GTO


def test_boys_against_series():
    T = numpy.array([0.0, 1e-9, 1e-4, 0.01, 0.7, 3.0, 12.0, 35.0, 80.0])
    F = og.boys(6, T)
    for n in range(7):
        for t, f in zip(T, F[n]):
            # e^{-T} sum (2T)^k / ((2n+1)(2n+3)...) with plenty of terms
            term = 1.0 / (2 * n + 1)
            s = term
            for k in range(1, 400):
                term *= 2 * t / (2 * n + 2 * k + 1)
                s += term
            assert f == pytest.approx(math.exp(-t) * s, rel=1e-13)


def test_ss_primitive_closed_form():
    a, b, ca, cb = 0.9, 1.7, 1.3, 0.6
    basis = {1: [(0, [a], [ca])], 2: [(0, [b], [cb])]}
    A = numpy.array([0.1, 0.2, -0.3])
    B = numpy.array([1.0, -0.4, 0.5])
    sysm = make_system([1, 2], [A, B], numpy.zeros((2, 2)), basis)
    sysm.dm[1, 0] = 0.8       # lower triangle: the only element read for the pair
    pts = numpy.array([[0.3, 0.3, 0.3], [2.0, -1.0, 0.0], [-0.6, 0.1, 0.9]])
    v = og.potential(sysm, pts, nuclei=False)
    p = a + b
    P = (a * A + b * B) / p
    K = math.exp(-a * b / p * numpy.sum((A - B) ** 2))
    for x, got in zip(pts, v):
        d = numpy.linalg.norm(P - x)
        want = -2 * 2 * 0.8 * ca * cb * C00 ** 2 * (math.pi / p) ** 1.5 * K * math.erf(math.sqrt(p) * d) / d
        assert got == pytest.approx(want, rel=1e-13)


def _small_spd_system():
    basis = {
        8: [(0, [3.1, 0.6], [0.4, 0.7]), (1, [2.2, 0.5], [0.5, 0.6]), (2, [1.1], [1.0])],
        1: [(0, [1.3], [1.0]), (1, [0.8], [0.9]), (2, [0.7], [0.8])],
    }
    rng = numpy.random.default_rng(7)
    A = [0.0, 0.0, 0.0]
    B = [0.9, -0.6, 0.7]
    sysm = make_system([8, 1], [A, B], numpy.zeros((18, 18)), basis)
    m = rng.standard_normal((18, 18))
    sysm.dm = m           # not symmetric on purpose
    return sysm


def test_gto_mcmurchie_davidson_vs_quadrature():
    """Hermite/Boys integrals against brute-force t-quadrature for all (La, Lb) classes up to D-D."""
    sysm = _small_spd_system()
    pts = numpy.array([[0.35, -0.2, 0.25], [1.5, 0.4, -0.8], [-0.7, 0.9, 0.3]])
    v = og.potential(sysm, pts, nuclei=False)
    w = og.potential_quadrature(sysm, pts, nt=160, ngh=10, nuclei=False)
    numpy.testing.assert_allclose(v, w, rtol=1e-9)


def _potential_solid_harmonic_quadrature(system, pts, nt=160):
    """V_H at pts by a route that shares nothing with the Hermite algebra or its Cartesian
    expansion: 1/|r - C| = (2/sqrt(pi)) int exp(-t^2 |r - C|^2) dt (t = s/(1-s), Gauss-Legendre
    in s); for fixed t the integrand is one Gaussian exp(-g |r - P_t|^2) times S_La(r - A)
    S_Lb(r - B), of degree <= 4 per axis, integrated exactly on a 3 x 3 x 3 Gauss-Hermite grid
    with the solid harmonics evaluated at the nodes."""
    xg, wg = numpy.polynomial.hermite.hermgauss(3)
    X, Y, Z = numpy.meshgrid(xg, xg, xg, indexing="ij")
    W = (wg[:, None, None] * wg[None, :, None] * wg[None, None, :]).ravel()
    nodes = numpy.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    sl, wl = numpy.polynomial.legendre.leggauss(nt)
    s = 0.5 * (sl + 1.0)
    tq = s / (1.0 - s)
    wt = 0.5 * wl / (1.0 - s) ** 2
    Pt = 2.0 * lower_symmetric(system.dm)
    offs = numpy.cumsum([0] + [2 * L + 1 for (_, L, _, _) in system.shells])
    out = numpy.zeros(len(pts))
    for n, C in enumerate(numpy.asarray(pts, dtype=numpy.float64)):
        acc = 0.0
        for sa, (ia, La, ala, coa) in enumerate(system.shells):
            A = system.coords[ia]
            for sb, (ib, Lb, alb, cob) in enumerate(system.shells):
                B = system.coords[ib]
                blk = Pt[offs[sa]:offs[sa + 1], offs[sb]:offs[sb + 1]]
                for a, ca in zip(ala, coa):
                    for b, cb in zip(alb, cob):
                        g = a + b + tq * tq                                          # (nt,)
                        Pc = (a * A + b * B + (tq * tq)[:, None] * C) / g[:, None]   # (nt, 3)
                        K = numpy.exp(-(a * b * numpy.sum((A - B) ** 2) + a * tq * tq * numpy.sum((A - C) ** 2)
                                        + b * tq * tq * numpy.sum((B - C) ** 2)) / g)
                        r = Pc[:, None, :] + nodes[None, :, :] / numpy.sqrt(g)[:, None, None]   # (nt, 27, 3)
                        ra, rb = r - A, r - B
                        fa = numpy.array([solid_harmonic(La, m, ra[..., 0], ra[..., 1], ra[..., 2]) for m in M_ORDER[La]])
                        fb = numpy.array([solid_harmonic(Lb, m, rb[..., 0], rb[..., 1], rb[..., 2]) for m in M_ORDER[Lb]])
                        integ = numpy.einsum("atn,btn,n->abt", fa, fb, W) * (K * wt * g ** -1.5)
                        acc += ca * cb * numpy.sum(blk[:, :, None] * integ)
        out[n] = -acc * 2.0 / math.sqrt(math.pi)
    return out


def test_gto_potential_vs_solid_harmonic_quadrature():
    """The Hermite/Boys potential against a quadrature built on the solid harmonics themselves
    (no shared Cartesian expansion), all (La, Lb) classes up to D-D, unsymmetric P."""
    sysm = _small_spd_system()
    pts = numpy.array([[0.35, -0.2, 0.25], [1.5, 0.4, -0.8], [-0.7, 0.9, 0.3]])
    v = og.potential(sysm, pts, nuclei=False)
    w = _potential_solid_harmonic_quadrature(sysm, pts)
    numpy.testing.assert_allclose(v, w, rtol=1e-9)


def test_gto_far_field_monopole(oracle_basis):
    """Far away, V_el -> -sum_pq Pt_pq S_pq / R (total electron count over R)."""
    from test_oracle_density import _overlap
    rng = numpy.random.default_rng(8)
    sysm = make_system([1, 1], [[0, 0, 0], [1.4, 0, 0]], numpy.eye(10), oracle_basis)
    m = rng.standard_normal((10, 10))
    sysm.dm = m @ m.T / 10
    S = _overlap(sysm)
    ne = 2 * numpy.sum(sysm.dm * S)
    centre = numpy.array([0.7, 0, 0])
    R = 400.0
    pts = centre + R * numpy.array([[0, 1.0, 0], [0, 0, 1.0], [0.6, 0.8, 0]])
    v = og.potential(sysm, pts, nuclei=False)
    numpy.testing.assert_allclose(v, -ne / R, rtol=2e-3)


# ----------------------------------------------------------------------------- V_ne


def test_vne_classic_bruteforce():
    types = [6, 1]
    coords = numpy.array([[0.31, 0.2, 0.1], [1.3, -0.4, 0.5]])
    nx, ny, nz, h = 8, 8, 8, 0.4
    origin = numpy.array([-1.5, -1.6, -1.4])
    nd = 2
    v = ov.vne_classic(types, coords, nx, ny, nz, h, origin, nd)
    for (i, j, k) in [(0, 0, 0), (3, 4, 3), (7, 7, 7)]:
        acc = 0.0
        cnt = 0
        for a in range(nd):
            for b in range(nd):
                for c in range(nd):
                    x = origin + numpy.array([i, j, k]) * h + (numpy.array([a, b, c]) + 0.5) * h / nd
                    for z, R in zip(types, coords):
                        d = numpy.linalg.norm(x - R)
                        if d != 0:
                            acc += z / d
                            cnt += 1
        assert v[i + nx * (j + ny * k)] == pytest.approx(acc / cnt / h ** 3, rel=1e-13)


def test_vne_classic_skips_coincident_subpoint():
    h, nd = 0.4, 2
    origin = numpy.zeros(3)
    R = origin + numpy.array([1, 1, 1]) * h + 0.25 * h   # exactly the first sub-point of voxel (1,1,1)
    v = ov.vne_classic([3], [R], 8, 8, 8, h, origin, nd)
    idx = 1 + 8 * (1 + 8 * 1)
    acc = 0.0
    for s in numpy.ndindex(2, 2, 2):
        x = origin + h + (numpy.array(s) + 0.5) * h / nd
        d = numpy.linalg.norm(x - R)
        if d != 0:
            acc += 3 / d
    assert v[idx] == pytest.approx(acc / 7 / h ** 3, rel=1e-13)


# ----------------------------------------------------------------------------- diffusion V_ne


def _vne_sweep_bruteforce(v, types, coords, h, origin, ads, dif):
    """Point-by-point restatement of one sweep with explicit periodic index wrap."""
    nz, ny, nx = v.shape
    out = numpy.zeros_like(v)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                s = 0.0
                for x in (-1, 0, 1):
                    for y in (-1, 0, 1):
                        for z in (-1, 0, 1):
                            d = abs(x) + abs(y) + abs(z)
                            if d == 0:
                                continue
                            s += dif * v[(k + z) % nz, (j + y) % ny, (i + x) % nx] / math.sqrt(d)
                c = v[k, j, i]
                s += (1.0 - dif * ov.VNE_SELF) * c
                s -= c * ads
                p = ((i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h)
                for zt, r in zip(types, coords):
                    d = math.sqrt(sum((p[a] - (r[a] - origin[a])) ** 2 for a in range(3)))
                    if d < math.sqrt(3.0) * h:
                        s += zt * math.exp(-d)
                if i in (0, nx - 1) or j in (0, ny - 1) or k in (0, nz - 1):
                    s = 0.0
                out[k, j, i] = s
    return out


def test_vne_sweep_bruteforce():
    rng = numpy.random.default_rng(3)
    v = rng.standard_normal((6, 5, 7))
    types = [6.0, 1.0]
    origin = (-0.3, 0.2, 0.1)
    h = 0.4
    coords = [(origin[0] + 2.5 * h + 0.05, origin[1] + 1.5 * h, origin[2] + 3.5 * h - 0.1),
              (origin[0] + 4.2 * h, origin[1] + 2.7 * h, origin[2] + 2.2 * h)]
    a = ov.vne_sweep(v, types, coords, h, origin, 0.1, 0.03)
    b = _vne_sweep_bruteforce(v, types, coords, h, origin, 0.1, 0.03)
    assert numpy.max(numpy.abs(a - b)) < 1e-13


def test_vne_constant_field_weights():
    """A constant interior field gains DIF*(sum of the 26 weights - 16.7947) - ADS per sweep."""
    v = numpy.full((7, 7, 7), 2.0)
    out = ov.vne_sweep(v, [], [], 0.3, (0, 0, 0), 0.05, 0.02)
    wsum = 6 + 12 / math.sqrt(2) + 8 / math.sqrt(3)
    expect = 2.0 * (1 + 0.02 * (wsum - 16.794682450997072) - 0.05)
    assert abs(out[3, 3, 3] - expect) < 1e-14
    assert out[0, 3, 3] == 0.0 and out[3, 6, 3] == 0.0


def test_vne_source_geometry_and_cutoff():
    """Atom on a voxel centre: Z there, Z e^{-h} on face neighbours, Z e^{-sqrt2 h} on edges, 0 on corners."""
    h = 0.25
    origin = (1.0, -2.0, 0.5)
    c = (origin[0] + 3.5 * h, origin[1] + 3.5 * h, origin[2] + 3.5 * h)   # centre of voxel (3,3,3)
    out = ov.vne_sweep(numpy.zeros((8, 8, 8)), [7.0], [c], h, origin, 0.1, 0.01)
    assert abs(out[3, 3, 3] - 7.0) < 1e-14
    assert abs(out[3, 3, 4] - 7.0 * math.exp(-h)) < 1e-14
    assert abs(out[4, 2, 3] - 7.0 * math.exp(-math.sqrt(2) * h)) < 1e-14
    assert out[4, 4, 4] == 0.0                     # distance sqrt(3) h is not < sqrt(3) h
    assert numpy.count_nonzero(out) == 1 + 6 + 12


def test_vne_pure_adsorption_geometric_series():
    """With DIF = 0 every interior voxel follows v <- (1 - a) v + S: v_m = S (1 - (1 - a)^m) / a."""
    h = 0.3
    origin = (0.0, 0.0, 0.0)
    c = (3.5 * h, 3.5 * h, 3.5 * h)
    a = 0.2
    res, n, delta, conv = ov.vne_diffusion([5.0], [c], 8, 8, 8, h, origin, adsorb=a, diff=0.0, tolerance=-1.0,
                                           maxiters=4)
    assert n == 5 and not conv
    m = 2 * n
    assert abs(res[3, 3, 3] - 5.0 * (1 - (1 - a) ** m) / a) < 1e-12
    s1 = 5.0 * math.exp(-h)
    assert abs(res[3, 3, 2] - s1 * (1 - (1 - a) ** m) / a) < 1e-12
    # whole field and the last pair's delta |v_m - v_{m-1}| / npts from the source S of one sweep
    S = ov.vne_sweep(numpy.zeros((8, 8, 8)), [5.0], [c], h, origin, a, 0.0)
    assert numpy.max(numpy.abs(res - S * (1 - (1 - a) ** m) / a)) < 1e-12
    assert abs(delta - float(numpy.sum(S)) * (1 - a) ** (m - 1) / 512) < 1e-14


def test_vne_loop_control():
    res, n, delta, conv = ov.vne_diffusion([1.0], [(0.5, 0.5, 0.5)], 6, 6, 6, 0.2, (0, 0, 0), tolerance=1.0)
    assert n == 0 and conv and not res.any()
    res, n, delta, conv = ov.vne_diffusion([1.0], [(0.5, 0.5, 0.5)], 6, 6, 6, 0.2, (0, 0, 0), tolerance=1e-3)
    assert conv and delta <= 1e-3 and n >= 1


def test_vne_loop_runs_one_pair_below_zero_cap():
    """The stop test follows the pair, so a negative cap still runs one pair (qmtools.py:2715-2741)."""
    res, n, delta, conv = ov.vne_diffusion([1.0], [(0.5, 0.5, 0.5)], 6, 6, 6, 0.2, (0, 0, 0), tolerance=-1.0,
                                           maxiters=-0.5)
    assert n == 1 and not conv and res.any()


@pytest.mark.parametrize("shape,sweeps,held", [((11, 9, 13), 0, True), ((11, 9, 13), 1, True),
                                               ((11, 9, 13), 17, True), ((16, 8, 10), 40, False)])
def test_jacobi_closed_form_equals_sweeps(shape, sweeps, held):
    """The spectral form used for 2000-sweep comparisons equals the plain sweep loop."""
    nx, ny, nz = shape
    rng = numpy.random.default_rng(sweeps + 3)
    rho = rng.standard_normal(nx * ny * nz)
    B = numpy.where(oj.boundary_mask(nx, ny, nz), rng.standard_normal(nx * ny * nz), 0.0) if held else None
    a = oj.jacobi(rho, nx, ny, nz, 0.3, sweeps, bvals=B)
    b = oj.jacobi_closed_form(rho, nx, ny, nz, 0.3, sweeps, bvals=B)
    assert numpy.abs(a - b).max() <= 1e-14 * max(numpy.abs(a).max(), 1.0)
