"""Pins of the oracle's basis, density, reduction and seed functions (CPU).

Each check is chosen so that a plausible slip in the oracle (wrong solid-harmonic
constant or axis, x/z layout swap, wrong sub-point offset, missing factor 2,
wrong seed weight) fails it.
"""
import math

import numpy
import pytest
import os

from conftest import BASIS, MOL29766

from oracle import density as od
from oracle import reduce as orr
from oracle import seed as osd
from oracle.basisfile import read_basis
from oracle.gto import hermite_E, CART
from oracle.system import M_ORDER, lower_symmetric, make_system


def test_basis_parsers_agree(basis, oracle_basis):
    """Product parser and the oracle's independent parser give the same shells."""
    assert set(basis.element_shells) == set(oracle_basis)
    for z, shells in oracle_basis.items():
        mine = basis.element_shells[z]
        assert len(mine) == len(shells)
        for (L1, a1, c1), (L2, a2, c2) in zip(mine, shells):
            assert L1 == L2
            numpy.testing.assert_array_equal(a1, a2)
            numpy.testing.assert_array_equal(c1, c2)
        # slot-strided reference tables
        first = basis.atomOffset[z]
        for s, (L, a, c) in enumerate(shells):
            off = basis.shellOffset[first + s]
            assert off == (first + s) * 20
            numpy.testing.assert_array_equal(basis.alphas64[off:off + len(a)], a)
            assert basis.Ls[first + s] == L


@pytest.mark.parametrize("z", [1, 6, 7, 8, 17, 35])
def test_contracted_shells_are_normalised(oracle_basis, z):
    """int phi^2 = sum_ij c_i c_j Gamma(L+3/2) / (2 (a_i+a_j)^(L+3/2)) = 1 with the kernel's
    orthonormal solid harmonics (this pins the coefficient convention of the file)."""
    for (L, a, c) in oracle_basis[z]:
        if L > 2:
            continue
        s = sum(ci * cj * math.gamma(L + 1.5) / (2 * (ai + aj) ** (L + 1.5))
                for ai, ci in zip(a, c) for aj, cj in zip(a, c))
        assert abs(s - 1.0) < 1e-6, (z, L, s)


def test_solid_harmonic_constants_are_orthonormal():
    """Monte-Carlo-free check: the prefactors equal sqrt((2L+1)/4pi)-type closed forms."""
    assert od.C00 == pytest.approx(0.5 / math.sqrt(math.pi), rel=1e-15)
    assert od.C1 == pytest.approx(math.sqrt(3 / (4 * math.pi)), rel=1e-15)
    assert od.C2A == pytest.approx(math.sqrt(15 / (4 * math.pi)), rel=1e-15)
    assert od.C20 == pytest.approx(math.sqrt(5 / (16 * math.pi)), rel=1e-15)
    assert od.C22 == pytest.approx(math.sqrt(15 / (16 * math.pi)), rel=1e-15)


def _overlap(system):
    """Analytic overlap matrix via Hermite E_0 (Obara-Saika / MD, textbook)."""
    n = system.nbf
    S = numpy.zeros((n, n))
    offs = []
    c = 0
    for sh in system.shells:
        offs.append(c)
        c += 2 * sh[1] + 1
    for sa, (A_, La, ala, coa) in enumerate(system.shells):
        for sb, (B_, Lb, alb, cob) in enumerate(system.shells):
            AB = system.coords[A_] - system.coords[B_]
            for a, ca in zip(ala, coa):
                for b, cb in zip(alb, cob):
                    p = a + b
                    E = [hermite_E(La, Lb, a, b, AB[d]) for d in range(3)]
                    K = E[0][1] * E[1][1] * E[2][1]
                    for ia, ma in enumerate(M_ORDER[La]):
                        for ib, mb in enumerate(M_ORDER[Lb]):
                            v = 0.0
                            for fa, ea in CART[(La, ma)]:
                                for fb, eb in CART[(Lb, mb)]:
                                    v += fa * fb * E[0][0][ea[0], eb[0], 0] * E[1][0][ea[1], eb[1], 0] \
                                        * E[2][0][ea[2], eb[2], 0]
                            S[offs[sa] + ia, offs[sb] + ib] += ca * cb * K * (math.pi / p) ** 1.5 * v
    return S


def _h2_system(oracle_basis, rng):
    types = [1, 1, 1]
    coords = numpy.array([[0.0, 0.0, 0.0], [1.4, 0.0, 0.0], [0.3, 1.2, -0.4]])
    nbf = 15
    a = rng.standard_normal((nbf, nbf))
    dm = a @ a.T / nbf
    return make_system(types, coords, dm, oracle_basis)


def test_overlap_diagonal_is_one(oracle_basis):
    sysm = _h2_system(oracle_basis, numpy.random.default_rng(0))
    S = _overlap(sysm)
    numpy.testing.assert_allclose(numpy.diag(S), 1.0, atol=1e-6)


def test_density_integrates_to_trace_PS(oracle_basis):
    """sum_v Q_raw = 2 Tr(P S) up to midpoint-rule error (smooth H functions only)."""
    sysm = _h2_system(oracle_basis, numpy.random.default_rng(1))
    S = _overlap(sysm)
    h = 0.15   # box of 24 bohr: the diffuse H tails are below 1e-14 at the edge
    n = 160
    origin = numpy.array([0.7, 0.6, -0.2]) - 0.5 * n * h
    q = od.density_grid_raw(sysm, n, n, n, h, origin)
    want = 2 * numpy.sum(lower_symmetric(sysm.dm) * S)
    assert abs(q.sum() - want) / abs(want) < 2e-6


def _overlap_gauss_hermite(system, harmonic=od.solid_harmonic):
    """S_pq exactly, with no Cartesian expansion: a product of two primitives is
    K exp(-p |r - P|^2) times S_La(r - A) S_Lb(r - B), a polynomial of degree <= La + Lb
    <= 4 per axis, which 3-point Gauss-Hermite per axis integrates exactly; the
    polynomial is evaluated through `harmonic` at the 27 nodes."""
    xg, wg = numpy.polynomial.hermite.hermgauss(3)
    X, Y, Z = numpy.meshgrid(xg, xg, xg, indexing="ij")
    W = (wg[:, None, None] * wg[None, :, None] * wg[None, None, :]).ravel()
    nodes = numpy.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    offs = numpy.cumsum([0] + [2 * L + 1 for (_, L, _, _) in system.shells])
    S = numpy.zeros((system.nbf, system.nbf))
    for sa, (ia, La, ala, coa) in enumerate(system.shells):
        A = system.coords[ia]
        for sb, (ib, Lb, alb, cob) in enumerate(system.shells):
            B = system.coords[ib]
            blk = numpy.zeros((2 * La + 1, 2 * Lb + 1))
            for a, ca in zip(ala, coa):
                for b, cb in zip(alb, cob):
                    p = a + b
                    r = (a * A + b * B) / p + nodes / math.sqrt(p)
                    K = math.exp(-a * b / p * numpy.sum((A - B) ** 2))
                    ra, rb = r - A, r - B
                    fa = numpy.array([harmonic(La, m, ra[:, 0], ra[:, 1], ra[:, 2]) for m in od.M_ORDER[La]])
                    fb = numpy.array([harmonic(Lb, m, rb[:, 0], rb[:, 1], rb[:, 2]) for m in od.M_ORDER[Lb]])
                    blk += ca * cb * K * p ** -1.5 * (fa * W) @ fb.T
            S[offs[sa]:offs[sa + 1], offs[sb]:offs[sb + 1]] = blk
    return S


def test_ccsd_density_matrix_trace_pins_basis_convention(oracle_basis):
    """The reference's own molecule (molecule_29766_0: Br, 2 Cl, 6 C, 3 H; 162 cc-pVDZ functions
    with s, p and d shells) ships a CCSD density matrix computed in the basis convention the
    reference's kernels use.  Its electron count 2 Tr(P S) must be sum Z = 108 exactly when the
    oracle's basis functions (order of m, solid-harmonic forms and signs, contraction
    normalisation) are those of the matrix.  Swapping the d functions m = +-1, or flipping the
    sign of d(x^2 - y^2), moves the count by 0.3 to 0.7 electrons."""
    xyz = numpy.loadtxt(os.path.join(MOL29766, "GEOM-B3LYP.xyz"))
    dm = numpy.load(os.path.join(MOL29766, "D-CCSD.npy"), allow_pickle=False)
    ang = 1.8897261246257702
    sysm = make_system(xyz[:, 0].astype(int), xyz[:, 1:] * ang, dm, oracle_basis)
    P = lower_symmetric(dm)
    nz = float(xyz[:, 0].sum())

    def count(harmonic):
        return 2.0 * numpy.sum(P * _overlap_gauss_hermite(sysm, harmonic))

    assert abs(count(od.solid_harmonic) - nz) < 1e-8 * nz

    def swapped(L, m, x, y, z):
        return od.solid_harmonic(L, -m if (L == 2 and abs(m) == 1) else m, x, y, z)

    def flipped(L, m, x, y, z):
        v = od.solid_harmonic(L, m, x, y, z)
        return -v if (L == 2 and m == 2) else v

    assert abs(count(swapped) - nz) > 0.1
    assert abs(count(flipped) - nz) > 0.1


def test_single_s_primitive_closed_form():
    """One normalised s primitive: phi(r) = C00 c exp(-a r^2)."""
    a, c = 0.8, 1.7
    s = make_system([1], [[0.1, -0.2, 0.3]], numpy.array([[1.0]]), {1: [(0, [a], [c])]})
    pts = numpy.random.default_rng(3).normal(size=(50, 3))
    r2 = numpy.sum((pts - s.coords[0]) ** 2, axis=1)
    numpy.testing.assert_allclose(od.basis_on_points(s, pts)[:, 0], od.C00 * c * numpy.exp(-a * r2), rtol=1e-14)


def test_density_grid_bruteforce_layout(oracle_basis):
    """Vectorised density vs an explicit per-voxel loop on a tiny grid, subgrid 2.

    Pins the x-fastest layout, the (i + (s+1/2)/n) h sub-point positions and the 2 h^3 / n^3 factor.
    """
    rng = numpy.random.default_rng(5)
    types = [8, 1]
    coords = numpy.array([[0.2, 0.1, -0.1], [1.6, 0.4, 0.3]])
    sysm = make_system(types, coords, numpy.eye(19), oracle_basis)
    a = rng.standard_normal((19, 19))
    sysm.dm = a  # not symmetric: the kernels read the lower triangle only
    nx, ny, nz, h = 8, 16, 8, 0.37
    origin = numpy.array([-1.3, -2.9, -1.4])
    n = 2
    q = od.density_grid_raw(sysm, nx, ny, nz, h, origin, subgrid=n)
    P = lower_symmetric(sysm.dm)
    for (i, j, k) in [(0, 0, 0), (3, 7, 2), (7, 15, 7), (5, 1, 6)]:
        acc = 0.0
        for sx in range(n):
            for sy in range(n):
                for sz in range(n):
                    r = origin + h * numpy.array([i + (sx + .5) / n, j + (sy + .5) / n, k + (sz + .5) / n])
                    ph = od.basis_on_points(sysm, r[None, :])[0]
                    for p_ in range(19):
                        for q_ in range(p_ + 1):
                            f = 1.0 if p_ == q_ else 2.0
                            acc += f * sysm.dm[p_, q_] * ph[p_] * ph[q_]
        want = 2 * acc * h ** 3 / n ** 3
        assert q[i + nx * (j + ny * k)] == pytest.approx(want, rel=1e-12, abs=1e-300)
    assert P.shape == (19, 19)


def test_normalised_density_sums_to_qtot(oracle_basis):
    sysm = make_system([8, 1, 1], [[0, 0, 0], [1.8, 0, 0], [-0.5, 1.7, 0]], numpy.eye(24) * 0.3,
                       oracle_basis)
    q, raw = od.density_grid(sysm, 16, 16, 16, 0.4, [-3.2, -3.2, -3.2])
    assert math.fsum(q) == pytest.approx(10.0, rel=1e-13)
    assert raw > 0


def test_reductions_closed_forms():
    n = 4096
    a = numpy.arange(n, dtype=float)
    assert orr.total(a) == n * (n - 1) / 2
    b = a + 0.25
    assert orr.absdiff(a, b) == 0.25 * n
    assert orr.reldiff(b, b + 1.0) == pytest.approx(100 * math.fsum(1.0 / (b + 1.0)), rel=1e-15)
    numpy.testing.assert_array_equal(orr.rescale(a, 2.0), 2 * a)


def test_seed_weights_closed_form():
    """Atom at fractional offsets (fx, fy, fz) of node i0: the lower node gets Z fx fy fz."""
    h = 0.5
    origin = numpy.array([-1.0, -2.0, -3.0])
    fx, fy, fz = 0.25, 0.5, 0.875
    i0 = numpy.array([3, 4, 5])
    R = origin + (i0 + [fx, fy, fz]) * h
    q = osd.seed_grid([7], [R], 8, 16, 24, h, origin)
    nx, ny = 8, 16
    for t in numpy.ndindex(2, 2, 2):
        w = 7.0
        for d, f in enumerate((fx, fy, fz)):
            w *= f if t[d] == 0 else (1 - f)
        idx = i0 + numpy.array(t)
        assert q[idx[0] + nx * (idx[1] + ny * idx[2])] == pytest.approx(w, rel=1e-14)
    assert math.fsum(q) == pytest.approx(7.0, rel=1e-14)
    assert numpy.count_nonzero(q) == 8
