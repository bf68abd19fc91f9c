"""GPU basis evaluation (cfk_basis_evaluate, src/SpectralBasis.m:25-58) against the oracle's
evaluate_basis, and GPU moment diagnostics (cfk_moments) against closed-form moments of projected
Maxwellian mixtures and against the oracle-checked moment weights."""
import math

import numpy as np
import pytest

import cfk
from oracle import geometry as geo

pytestmark = pytest.mark.gpu


def _points(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3)) * rng.uniform(0.1, 2.5, size=(n, 1))
    special = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.3], [0.0, 0.0, -0.7], [0.9, 0.0, 0.0], [0.0, -1.1, 0.0],
                        [1e-9, 0.0, 2e-9], [3.0, 3.0, -3.0]])
    return np.vstack([special, v])


@pytest.mark.parametrize("K,L", [(10, 8), (14, 12), (3, 16), (1, 0)])
def test_basis_matches_oracle(torch_cuda, K, L):
    torch = torch_cuda
    v = _points(300, K + L)
    print("basis K=%d L=%d on %d points" % (K, L, v.shape[0]), flush=True)
    Psi = cfk.basis_evaluate(torch.from_numpy(v).cuda(), K, L).cpu().numpy()
    ref = geo.evaluate_basis(K, L, v)
    scale = np.maximum(np.max(np.abs(ref), axis=0), 1e-300)   # per basis function
    assert np.max(np.abs(Psi - ref) / scale) <= 1e-12


def _mixture_moments(p):
    """Closed-form moments of a mixture p [n_comp, 7] = (n, u, T) with component densities
    n (pi^3 Tx Ty Tz)^{-1/2} exp(-sum (v_d - u_d)^2 / T_d)."""
    n = p[:, 0].sum()
    mom = (p[:, :1] * p[:, 1:4]).sum(axis=0)
    u = mom / n
    second = np.zeros((3, 3))
    for c in p:
        second += c[0] * (np.diag(c[4:7] / 2.0) + np.outer(c[1:4], c[1:4]))
    P = second - n * np.outer(u, u)
    return n, u, P


def test_moments_of_maxwellian_mixtures(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    B, L, K = 40, 8, 10
    params = np.zeros((B, 3, 7))
    params[:, :, 4:7] = 1.0
    for b in range(B):
        for c in range(1 + b % 3):
            params[b, c] = [rng.uniform(0.3, 1.5), *rng.uniform(-0.25, 0.25, 3), *rng.uniform(0.85, 1.15, 3)]
    c = cfk.project_maxwellians(params, L, K)
    out = cfk.moments(c, B, (L + 1) ** 2, K, layout="batch_fastest").cpu().numpy()
    for b in range(B):
        n, u, P = _mixture_moments(params[b])
        assert abs(out[b, 0] - n) <= 1e-8 * n
        assert np.max(np.abs(out[b, 1:4] - u)) <= 1e-8
        Pg = np.array([[out[b, 5], out[b, 8], out[b, 9]], [out[b, 8], out[b, 6], out[b, 10]],
                       [out[b, 9], out[b, 10], out[b, 7]]])
        assert np.max(np.abs(Pg - P)) <= 1e-8 * n
        assert abs(out[b, 4] - 2 * np.trace(P) / (3 * n)) <= 1e-12


def test_moments_match_weights_and_unit_maxwellian(torch_cuda):
    torch = torch_cuda
    L, K, B = 6, 8, 33
    nq = (L + 1) ** 2
    rng = np.random.default_rng(1)
    c = rng.standard_normal((B, K * nq)) * 0.05
    c[:, 0] = 1.0 + rng.uniform(0, 0.5, B)          # positive density
    c[0] = 0.0
    c[0, 0] = 1.0                                     # the unit Maxwellian: n = pi^(3/4), u = 0, T = 1
    M = cfk.moment_weights(nq, K)                     # oracle-checked (tests/test_moment_weights.py)
    m = c @ M.T
    for layout, arr in (("natural", c), ("batch_fastest", np.ascontiguousarray(c.T))):
        out = cfk.moments(torch.from_numpy(arr).cuda(), B, nq, K, layout=layout).cpu().numpy()
        n = m[:, 0]
        assert np.max(np.abs(out[:, 0] - n) / n) <= 1e-14
        u = m[:, 1:4] / n[:, None]
        assert np.max(np.abs(out[:, 1:4] - u)) <= 1e-14
        pxx = m[:, 4] - n * u[:, 0] ** 2
        assert np.max(np.abs(out[:, 5] - pxx)) <= 1e-13
    assert abs(out[0, 0] - math.pi ** 0.75) <= 1e-14
    assert np.max(np.abs(out[0, 1:4])) == 0.0
    assert abs(out[0, 4] - 1.0) <= 1e-14
