"""ORACLE (test infrastructure only) -- numpy restatement of the basis projection of Maxwellian
mixtures, the data side of the learned surrogate.

Only tests may import this module, and only as a checker; the product path never calls it.

Restates tutorial_collisions_spectral_properties.m:100-158 (and the projector of
tutorial_collisions_bwu_solution.m:237-269): a tensor grid of N_rad generalized Gauss-Laguerre
(alpha = 1/2) nodes in x = |v|^2 with weights w/2, N_pol Gauss-Legendre nodes in cos(theta) and
N_azi uniform azimuths with weight 2 pi / N_azi; then

    c = Psi^T (W * F),   F(v) = f(v) / e^{-|v|^2}

with Psi the basis of src/SpectralBasis.m:25-58 (``oracle.geometry.evaluate_basis``), evaluated
densely at every grid point (no factorization).

Pins (tests/test_oracle_projection.py): orthonormality of Psi under the grid; the unit Maxwellian
projects onto the mass coefficient alone; mass, momentum and energy coefficients of shifted and
anisotropic Maxwellians against their closed forms.
"""
import math

import numpy as np

from . import geometry as geo


def grid(n_rad=60, n_pol=30, n_azi=30):
    """Points v [N, 3] and weights W [N] (rad-major, then polar, then azimuth)."""
    x, w = geo.gauss_gen_laguerre(n_rad, 0.5)
    r = np.sqrt(x)
    wr = w / 2.0
    mu, wp = geo.gauss_legendre(n_pol, -1.0, 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - mu * mu))
    phi = 2.0 * math.pi * np.arange(n_azi) / n_azi
    wphi = 2.0 * math.pi / n_azi
    R, P, A = np.meshgrid(np.arange(n_rad), np.arange(n_pol), np.arange(n_azi), indexing="ij")
    R, P, A = R.ravel(), P.ravel(), A.ravel()
    v = np.stack([r[R] * st[P] * np.cos(phi[A]), r[R] * st[P] * np.sin(phi[A]), r[R] * mu[P]], axis=1)
    W = wr[R] * wp[P] * wphi
    return v, W


def mixture_F(params, v):
    """f(v) e^{|v|^2} [N] of one mixture, params [n_comp, 7] = (n, ux, uy, uz, Tx, Ty, Tz)."""
    v2 = np.sum(v * v, axis=1)
    F = np.zeros(v.shape[0])
    for comp in params:
        n = comp[0]
        if n <= 0.0:
            continue
        u, T = comp[1:4], comp[4:7]
        lc = math.log(n) - 1.5 * math.log(math.pi) - 0.5 * math.log(T[0] * T[1] * T[2])
        F += np.exp(lc + v2 - np.sum((v - u[None, :]) ** 2 / T[None, :], axis=1))
    return F


def project(params, L, K, n_rad=60, n_pol=30, n_azi=30):
    """Coefficients [K*N_Q, B] (row k*N_Q + q) of B mixtures params [B, n_comp, 7]."""
    v, W = grid(n_rad, n_pol, n_azi)
    Psi = geo.evaluate_basis(K, L, v)            # [N, K*N_Q]
    out = np.empty((K * (L + 1) ** 2, len(params)))
    for b, p in enumerate(params):
        out[:, b] = Psi.T @ (W * mixture_F(np.asarray(p, dtype=np.float64), v))
    return out
