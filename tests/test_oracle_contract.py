"""Pins of the oracle contraction kernels (oracle/cfk_oracle.c, restating src/mex/*_collision_kernel_mex.cpp).

None of these retype the kernels' formula:
* the four reference algorithms (naive, angular-first, radial-first, dense assembly) agree with
  each other although they loop differently (a transposed R index or a dropped term in one of
  them breaks agreement);
* physics known answers on the reference's own R fixtures:
  - the absolute Maxwellian is an equilibrium, Q(e_000) = 0;
  - the five collision invariants of Q(f) are exactly zero;
  - Maxwell molecules: the linearized spectrum equals the Wang Chang-Uhlenbeck closed form to
    1e-10 (tests/TestWCUEigenvalues.m:77-134), on the fixture R and on the reference's own
    test recipe through the oracle sumfac (K=2, L=2, pad 16);
  - hard spheres: the second-order Chapman-Enskog viscosity factor is 205/202
    (tutorial_chapman_enscog_decay_rates.m:75-77).
"""
import numpy as np
import pytest

import oracle
from cfk.inputs import perturbed_maxwellians
from conftest import load_precalc
from oracle import geometry as G


def q_of(variant, f, lab, val, ic, R, nq, K):
    Q = np.zeros(nq * K)
    return oracle.contract(variant, Q, f, lab, val, ic, R, nq, K)


@pytest.mark.parametrize("L,gamma", [(2, 0.0), (4, 1.0), (6, 1.0)])
def test_variants_agree(L, gamma):
    lab, val, ic, R, _ = load_precalc(L, gamma)
    nq, K = (L + 1) ** 2, 5
    f = perturbed_maxwellians(L, K, 1, seed=7)[0].ravel()
    ls, vs, ics = G.sort_t_q1(lab, val, ic)
    qn = q_of("naive", f, lab, val, ic, R, nq, K)
    qa = q_of("angular_first", f, ls, vs, ics, R, nq, K)
    qr = q_of("radial_first", f, ls, vs, ics, R, nq, K)
    # the angular-first kernel on UNSORTED input (shorter runs) must give the same Q
    qa_u = q_of("angular_first", f, lab, val, ic, R, nq, K)
    scale = np.max(np.abs(qn))
    for q in (qa, qr, qa_u):
        assert np.max(np.abs(q - qn)) <= 1e-13 * scale
    if L <= 4:
        N = nq * K
        C = G.assemble_dense(lab, val, ic, R, nq, K)
        qd = oracle.dense(np.zeros(N), f, C, N)
        assert np.max(np.abs(qd - qn)) <= 1e-13 * scale


@pytest.mark.parametrize("L,gamma", [(4, 1.0), (8, 1.0), (6, 0.0)])
def test_maxwellian_equilibrium_and_invariants(L, gamma):
    lab, val, ic, R, _ = load_precalc(L, gamma)
    nq, K = (L + 1) ** 2, 5
    ls, vs, ics = G.sort_t_q1(lab, val, ic)
    e1 = np.zeros(nq * K)
    e1[0] = 1.0
    q0 = q_of("angular_first", e1, ls, vs, ics, R, nq, K)
    assert np.max(np.abs(q0)) < 1e-12
    f = perturbed_maxwellians(L, K, 3, seed=1)
    for b in range(3):
        q = q_of("angular_first", f[b].ravel(), ls, vs, ics, R, nq, K)
        assert np.max(np.abs(q)) > 1e-3
        inv = [0, 1, 2, 3, nq]   # q=0..3 at k=0 and q=0 at k=1 (0-based, flat k*N_Q + q)
        assert np.all(q[inv] == 0.0)


def jacobian(lab, val, ic, R, nq, K):
    """J = dQ/dc at the absolute Maxwellian, column j = (Q(e1+e_j) - Q(e1-e_j)) / 2."""
    N = nq * K
    J = np.zeros((N, N))
    for j in range(N):
        c = np.zeros(N)
        c[0] = 1.0
        c[j] += 1.0
        qp = q_of("naive", c, lab, val, ic, R, nq, K)
        c[j] -= 2.0
        qm = q_of("naive", c, lab, val, ic, R, nq, K)
        J[:, j] = 0.5 * (qp - qm)
    return J


def wcu_spectrum(K, L):
    # tests/TestWCUEigenvalues.m:89-134
    mu, w = G.gauss_legendre(200, -1.0, 1.0)
    c = np.sqrt((1 + mu) / 2)
    s = np.sqrt((1 - mu) / 2)
    lam = []
    for l in range(L + 1):
        Pc = G.assoc_legendre_all(l, c)[0]
        Ps = G.assoc_legendre_all(l, s)[0]
        for k in range(K):
            val = 2 * np.pi * np.sum(w * (c ** (2 * k + l) * Pc + s ** (2 * k + l) * Ps - 1))
            if (k, l) in ((0, 0), (0, 1), (1, 0)):
                val = 0.0
            lam += [val] * (2 * l + 1)
    return np.sort(np.array(lam))[::-1]


def check_wcu(J, K, L):
    lam = np.sort(np.real(np.linalg.eigvals(J)))[::-1]
    ref = wcu_spectrum(K, L)
    ln = lam / abs(lam[5])
    lr = ref / abs(ref[5])
    assert np.all(np.abs(ln[:5]) < 1e-10)
    assert np.max(np.abs(ln - lr)) < 1e-10


def test_wcu_spectrum_on_fixture():
    lab, val, ic, R, _ = load_precalc(2, 0.0)
    check_wcu(jacobian(lab, val, ic, R, 9, 5), 5, 2)


def test_wcu_spectrum_reference_recipe_through_oracle_sumfac():
    # tests/TestWCUEigenvalues.m:19-31: K_max=2, L=2, gamma=0, generate_R_tensor_sumfac(16,16)
    K_max, L = 2, 2
    lab, val, ic = G.real_gaunt_geometry(L)
    grids = G.sumfac_grids(0.0, G.sumfac_grid_sizes(K_max, L, 16, 16))
    args = G.sumfac_prep(K_max, L, 0.0, lab, val, ic, grids)
    R = G.conservation_zero(oracle.sumfac(args), args["L_triplets"].astype(int), K_max + 1)
    check_wcu(jacobian(lab, val, ic, R, 9, K_max + 1), K_max + 1, L)


def test_chapman_enskog_viscosity_factor():
    # tutorial_chapman_enscog_decay_rates.m:34-91 on the HS K_max=4, L=4 fixture
    lab, val, ic, R, _ = load_precalc(4, 1.0)
    nq, K = 25, 5
    J = jacobian(lab, val, ic, R, nq, K)
    idx = np.arange(K) * nq + 6
    H = J[np.ix_(idx, idx)]
    H = H / H[0, 0]
    fmu = [np.linalg.inv(H[:k, :k])[0, 0] for k in range(1, K + 1)]
    assert abs(fmu[0] - 1.0) < 1e-12
    assert abs(fmu[1] - 205.0 / 202.0) < 2e-6
    assert all(fmu[i] < fmu[i + 1] < 1.016034 + 1e-5 for i in range(K - 1))
