"""Pins of the oracle's host-side restatement (oracle/geometry.py) and of libcfk's native prep.

What pins what:
* real Gaunt COO: labels and ic_map bit-exact, values to 1e-15 relative, against the reference's
  own precalc fixtures for every shipped L (2..12);
* Gaunt values: against an independent brute-force quadrature of Y_q1 Y_q2 Y_q3 over the sphere
  (textbook definition of the real Gaunt coefficient), so a wrong phase or transform fails;
* selection rules (triangle, parity, real |m3| rule) hold for every entry;
* 1-D rules: polynomial exactness (tests/TestQuadraturePolynomialReproduction.m:13-49 recipe);
* libcfk's native prep (C++) against the same fixtures and against the oracle.
"""
import numpy as np
import pytest

from conftest import load_precalc
from oracle import geometry as G

LS = [2, 4, 6, 8, 10, 12]


@pytest.mark.parametrize("L", LS)
def test_oracle_gaunt_matches_reference_fixture(L):
    lab, val, ic = G.real_gaunt_geometry(L)
    rl, rv, ric, _, _ = load_precalc(L, 1.0)
    assert np.array_equal(lab, rl)
    assert np.array_equal(ic, ric)
    assert np.max(np.abs(val - rv)) <= 1e-15 * max(1.0, np.max(np.abs(rv))) * 8


@pytest.mark.parametrize("L", [2, 4, 6])
def test_fixture_geometry_independent_of_gamma(L):
    a = load_precalc(L, 0.0)
    b = load_precalc(L, 1.0)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


def test_gaunt_values_against_bruteforce_quadrature():
    # G[q1,q2,q3] = int Y_q1 Y_q2 Y_q3 dOmega; integrand degree <= 3L: Gauss-Legendre in cos(theta)
    # with 3L/2+1 nodes and 3L+1 uniform phi nodes integrate it exactly.
    L = 4
    lab, val, _ = G.real_gaunt_geometry(L)
    mu, wmu = np.polynomial.legendre.leggauss(3 * L // 2 + 2)
    nphi = 3 * L + 2
    phi = 2 * np.pi * np.arange(nphi) / nphi
    TH, PH = np.meshgrid(np.arccos(mu), phi, indexing="ij")
    W = (wmu[:, None] * np.full(nphi, 2 * np.pi / nphi)[None, :]).ravel()
    Y = G.real_sh(L, TH.ravel(), PH.ravel())   # [npts, N_Q]
    Gfull = np.einsum("p,pa,pb,pc->abc", W, Y, Y, Y)
    nz = np.abs(Gfull) > 1e-12
    dense = np.zeros_like(Gfull)
    dense[lab[:, 0] - 1, lab[:, 1] - 1, lab[:, 2] - 1] = val
    assert np.array_equal(nz, dense != 0)
    assert np.max(np.abs(Gfull - dense)) < 1e-13


@pytest.mark.parametrize("L", [5, 8])
def test_selection_rules(L):
    lab, _, _ = G.real_gaunt_geometry(L)
    ls = G.q_to_l(lab)
    ms = lab - 1 - ls * ls - ls
    l1, l2, l3 = ls.T
    assert np.all((l3 >= np.abs(l1 - l2)) & (l3 <= l1 + l2))
    assert np.all((l1 + l2 + l3) % 2 == 0)
    a1, a2, a3 = np.abs(ms).T
    assert np.all((a3 == a1 + a2) | (a3 == np.abs(a1 - a2)))
    assert lab.min() >= 1 and lab.max() <= (L + 1) ** 2
    # the complex skeleton obeys m1 + m2 = m3 (tests/TestSphericalHarmonics.m:45-56)
    c = G.gaunt_generate_labels(L)
    cl = G.q_to_l(c)
    cm = c - 1 - cl * cl - cl
    assert np.all(cm[:, 0] + cm[:, 1] == cm[:, 2])


def test_quadrature_polynomial_exactness():
    # tests/TestQuadraturePolynomialReproduction.m:13-49: x^m, m <= 4 on [1.5, 4.5]
    x, w = G.gauss_legendre(5, 1.5, 4.5)
    for m in range(5):
        exact = (4.5 ** (m + 1) - 1.5 ** (m + 1)) / (m + 1)
        assert abs(np.sum(w * x ** m) - exact) < 1e-12 * max(1, abs(exact))
    # generalized Laguerre: int x^m x^a e^-x = Gamma(m+a+1), exact up to degree 2n-1
    from scipy.special import gamma as Gam
    for a in (0.0, 0.5, 1.0):
        x, w = G.gauss_gen_laguerre(12, a)
        for m in range(0, 23):
            exact = Gam(m + a + 1)
            assert abs(np.sum(w * x ** m) - exact) <= 1e-12 * exact


def test_native_prep_matches_fixtures_and_oracle():
    import cfk
    for L in (2, 8, 12):
        lab, val, ic = cfk.real_gaunt(L)
        rl, rv, ric, _, _ = load_precalc(L, 1.0)
        assert np.array_equal(lab.astype(np.int64), rl)
        assert np.array_equal(ic.astype(np.int64), ric)
        assert np.max(np.abs(val - rv)) <= 1e-14
    for n, a, b in ((1, 0, 1), (7, -1, 1), (48, 0, 1)):
        x, w = cfk.gauss_legendre(n, a, b)
        xo, wo = G.gauss_legendre(n, a, b)
        assert np.max(np.abs(x - xo)) <= 1e-15 and np.max(np.abs(w - wo)) <= 1e-15
    for n, al in ((1, 0.0), (29, 0.0), (48, 0.5)):
        x, w = cfk.gauss_gen_laguerre(n, al)
        xo, wo = G.gauss_gen_laguerre(n, al)
        assert np.max(np.abs(x - xo) / xo) <= 1e-13
        assert np.max(np.abs(w - wo) / np.max(wo)) <= 1e-13
    for args in ((4, 2, 20, 20), (9, 8, 40, 0), (11, 12, 0, 0)):
        assert cfk.sumfac_grid_sizes(*args) == G.sumfac_grid_sizes(*args)
