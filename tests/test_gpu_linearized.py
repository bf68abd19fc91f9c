"""The linearized operator on the GPU (cfk_plan_linearize), pinned by the physics the reference's
own tests and tutorials use, at sizes beyond the reference's fixtures:
* against the oracle's Jacobian (restated naive MEX) on the reference's L=4 hard-sphere fixture;
* Maxwell molecules: the spectrum equals the Wang Chang-Uhlenbeck closed form
  (tests/TestWCUEigenvalues.m:77-134) at K=6, L=6 with R from the GPU sumfac (pad 16, 16);
* hard spheres at L=8, K=10 (R from the config-1 golden tensor): the Chapman-Enskog viscosity
  factor f_mu is 1 at one mode, 205/202 at two, and increases towards 1.016034
  (tutorial_chapman_enscog_decay_rates.m:75-77).
"""
import numpy as np
import pytest

import cfk
from conftest import GOLDEN, load_precalc
from oracle import geometry as G
from test_oracle_contract import check_wcu, jacobian

pytestmark = pytest.mark.gpu


def test_matches_oracle_jacobian_on_fixture(torch_cuda):
    lab, val, ic, R, _ = load_precalc(4, 1.0)
    nq, K = 25, 5
    ls, vs, ics = G.sort_t_q1(lab, val, ic)
    plan = cfk.CollisionPlan(ls, vs, ics, R, nq, K)
    J = plan.linearize()
    Jo = jacobian(lab, val, ic, R, nq, K)
    assert np.max(np.abs(J - Jo)) <= 1e-12 * np.max(np.abs(Jo))
    # about another state: J(f0) e_j = B(f0, e_j) + B(e_j, f0) -- check against finite differences of Q
    f0 = np.zeros(nq * K)
    f0[0] = 1.0
    f0[6] = 0.05
    J0 = plan.linearize(f0)
    import oracle
    h = np.zeros(nq * K)
    h[nq + 2] = 1.0
    qp = oracle.contract("naive", np.zeros(nq * K), f0 + h, lab, val, ic, R, nq, K)
    qm = oracle.contract("naive", np.zeros(nq * K), f0 - h, lab, val, ic, R, nq, K)
    assert np.max(np.abs(J0[:, nq + 2] - 0.5 * (qp - qm))) <= 1e-12 * np.max(np.abs(qp))


def test_wcu_spectrum_maxwell_molecules_k6_l6(torch_cuda):
    K_max, L = 5, 6
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    print("GPU sumfac gamma=0, K=6, L=6, pad 16", flush=True)
    R = cfk.build_rtensor(K_max, L, 0.0, cfk.sumfac_grid_sizes(K_max, L, 16, 16), geometry=(lab, val, ic))
    plan = cfk.CollisionPlan(ls, vs, ics, R, (L + 1) ** 2, K_max + 1)
    check_wcu(plan.linearize(), K_max + 1, L)


def test_chapman_enskog_factor_hard_spheres_l8_k10(torch_cuda):
    g = np.load(GOLDEN + "/sumfac_hs_l8_k10_g48.npz")
    L, K = 8, 10
    nq = (L + 1) ** 2
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    plan = cfk.CollisionPlan(ls, vs, ics, g["R"], nq, K)
    J = plan.linearize()
    idx = np.arange(K) * nq + 6          # the l=2, m=0 stress mode, radial modes 0..K-1
    H = J[np.ix_(idx, idx)]
    H = H / H[0, 0]
    fmu = [np.linalg.inv(H[:k, :k])[0, 0] for k in range(1, K + 1)]
    print("f_mu by number of radial modes:", ["%.6f" % x for x in fmu], flush=True)
    assert abs(fmu[0] - 1.0) < 1e-12
    assert abs(fmu[1] - 205.0 / 202.0) < 2e-6
    assert all(fmu[i] < fmu[i + 1] + 1e-9 for i in range(K - 1))
    assert abs(fmu[-1] - 1.016034) < 2e-5
