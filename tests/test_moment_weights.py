"""The moment functionals of cfk_moment_weights (closed forms in csrc/basis.hip) against the oracle:
int e^{-|v|^2} phi_i(v) p(v) dv for the monomials p = 1, v_x, v_y, v_z, v_x^2, v_y^2, v_z^2, v_x v_y,
v_x v_z, v_y v_z, by quadrature of the oracle's basis (src/SpectralBasis.m:25-58 restated) on a tensor
grid exact for these polynomials.  Host only (no device)."""
import numpy as np
import pytest

import cfk
from oracle import geometry as geo
from oracle import projection as pj


def oracle_weights(L, K):
    v, W = pj.grid(n_rad=K + 6, n_pol=L + 6, n_azi=2 * L + 8)
    Psi = geo.evaluate_basis(K, L, v)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    mono = [np.ones_like(x), x, y, z, x * x, y * y, z * z, x * y, x * z, y * z]
    return np.stack([Psi.T @ (W * p) for p in mono])


@pytest.mark.parametrize("L,K", [(2, 2), (4, 6), (8, 10), (12, 14)])
def test_moment_weights_match_oracle_quadrature(L, K):
    M = cfk.moment_weights((L + 1) ** 2, K)
    Mo = oracle_weights(L, K)
    assert M.shape == Mo.shape
    assert np.max(np.abs(M - Mo)) <= 1e-13 * np.max(np.abs(Mo))
    # only l <= 2 (q < 9) at k = 0, and the mass mode at k = 1, carry weight: the rest is orthogonal to
    # every polynomial of degree <= 2
    nq = (L + 1) ** 2
    cols = np.nonzero(np.any(M != 0, axis=0))[0]
    assert set(cols.tolist()) <= set(range(9)) | {nq}


def test_moment_weights_reject_small_bases():
    with pytest.raises(cfk.CfkError):
        cfk.moment_weights(4, 2)    # L = 1: no second-order moments
    with pytest.raises(cfk.CfkError):
        cfk.moment_weights(9, 1)    # K = 1: no energy mode
