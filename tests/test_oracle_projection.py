"""Pins of the projection oracle (oracle/projection.py, restating
tutorial_collisions_spectral_properties.m:100-158) used to check the GPU projection kernel.

* the basis is orthonormal under the projection grid (L=8, K=10);
* the unit Maxwellian projects onto the mass coefficient alone;
* shifted and anisotropic Maxwellians: mass, momentum and energy coefficients equal their closed
  forms (mass n pi^{-3/4}; momentum -n s u_y, n s u_z, -n s u_x in the real, Condon-Shortley
  basis with s = RadialNorm(0,1) sqrt(3/(4 pi)); energy n RadialNorm(1,0) (3/2 - <v^2>)/sqrt(4 pi)
  with <v^2> = sum_d T_d/2 + u_d^2).
"""
import math

import numpy as np
import pytest

from oracle import geometry as geo
from oracle import projection as pj

L, K = 8, 10
NQ = (L + 1) ** 2
S1 = math.exp(-0.5 * (-math.log(2) + math.lgamma(2.5))) * math.sqrt(3 / (4 * math.pi))
E1 = math.exp(-0.5 * (-math.log(2) + math.lgamma(2.5) - math.lgamma(2.0))) / math.sqrt(4 * math.pi)


def test_basis_orthonormal_on_grid():
    v, W = pj.grid()
    Psi = geo.evaluate_basis(K, L, v)
    G = Psi.T @ (W[:, None] * Psi)
    assert np.max(np.abs(G - np.eye(K * NQ))) < 1e-12


def test_unit_maxwellian_is_the_mass_mode():
    c = pj.project(np.array([[[math.pi ** 0.75, 0, 0, 0, 1, 1, 1]]]), L, K)[:, 0]
    assert abs(c[0] - 1.0) < 1e-13
    assert np.max(np.abs(c[1:])) < 1e-13


@pytest.mark.parametrize("T,tol", [((1.0, 1.0, 1.0), 1e-12), ((0.7, 1.3, 1.1), 1e-9), ((0.5, 2.0, 1.1), 1e-5)])
def test_moments_closed_form(T, tol):
    n, u, T = 1.7, np.array([0.3, -0.2, 0.4]), np.array(T)
    c = pj.project(np.array([[[n, *u, *T]]]), L, K)[:, 0]
    assert abs(c[0] - n * math.pi ** -0.75) < tol
    assert np.allclose(c[1:4], [-n * S1 * u[1], n * S1 * u[2], -n * S1 * u[0]], rtol=0, atol=tol)
    assert abs(c[NQ] - n * E1 * (1.5 - np.sum(T / 2 + u ** 2))) < tol


def test_mixture_is_linear_in_components():
    a = np.array([[0.6, 0.1, 0.0, -0.2, 0.8, 1.2, 1.0], [0.0, 0, 0, 0, 1, 1, 1], [0.0, 0, 0, 0, 1, 1, 1]])
    b = np.array([[1.1, -0.3, 0.2, 0.1, 1.5, 0.6, 0.9], [0.0, 0, 0, 0, 1, 1, 1], [0.0, 0, 0, 0, 1, 1, 1]])
    ab = np.array([a[0], b[0], [0.0, 0, 0, 0, 1, 1, 1]])
    c = pj.project(np.stack([a, b, ab]), 4, 6)
    assert np.max(np.abs(c[:, 2] - c[:, 0] - c[:, 1])) < 1e-13 * np.max(np.abs(c))
