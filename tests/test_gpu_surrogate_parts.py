"""GPU pieces of the learned surrogate against the oracle:
* Maxwellian-mixture projection kernel (cfk_project_maxwellians) vs oracle/projection.py at 1e-12;
* the linear blocks J_l of cfk_surrogate_create vs the dense linearization of the oracle
  angular-first contraction (restated reference MEX) at 1e-12, on the reference's L=4 fixture;
* the linear-only surrogate n^2 J d vs numpy with those blocks (1e-12) and vs the exact operator
  for a small perturbation (relative difference of the order of the perturbation);
* the anchored linear-only surrogate Q(c_M) + n^2 J g vs numpy with those blocks, c_M the projection of
  each cell's Maxwellian (moments from cfk.moments) and Q(c_M) from the exact operator (1e-12), and vs
  the exact operator near the Maxwellian (relative difference of the order of the deviation).
"""
import math

import numpy as np
import pytest

import cfk
import oracle
from conftest import load_precalc
from oracle import geometry as G
from oracle import projection as pj

pytestmark = pytest.mark.gpu


def test_projection_matches_oracle(torch_cuda):
    L, K = 8, 10
    params = cfk.maxwellian_mixture_params(6, seed=21)
    print("projection L=8 K=10, 6 mixtures", flush=True)
    c = cfk.project_maxwellians(params, L, K).cpu().numpy()
    co = pj.project(params, L, K)
    assert np.max(np.abs(c - co)) <= 1e-12 * np.max(np.abs(co))
    assert np.max(np.abs(c[0] - 1.0)) < 1e-5        # densities scaled so c_000 = 1


def _fixture_plan(L=4, K=5):
    lab, val, ic, R, _ = load_precalc(L, 1.0)
    ls, vs, ics = G.sort_t_q1(lab, val, ic)
    return (ls, vs, ics, R), cfk.CollisionPlan(ls, vs, ics, R, (L + 1) ** 2, K)


def test_linear_blocks_match_oracle(torch_cuda):
    L, K = 4, 5
    nq = (L + 1) ** 2
    NK = nq * K
    (ls, vs, ics, R), plan = _fixture_plan(L, K)
    sur = cfk.Surrogate(plan, width=64, depth=1, seed=0)
    JL = sur.linear_blocks()
    # dense J from the oracle: column j = (Q(e1 + e_j) - Q(e1 - e_j)) / 2, MATLAB order q + k*N_Q
    J = np.zeros((NK, NK))
    for j in range(NK):
        fp = np.zeros(NK)
        fp[0] = 1.0
        fm = fp.copy()
        fp[j] += 1.0
        fm[j] -= 1.0
        qp = oracle.contract("angular_first", np.zeros(NK), fp, ls, vs, ics, R, nq, K)
        qm = oracle.contract("angular_first", np.zeros(NK), fm, ls, vs, ics, R, nq, K)
        J[:, j] = 0.5 * (qp - qm)
    scale = np.max(np.abs(J))
    model = np.zeros_like(J)
    for q in range(nq):
        l = int(math.isqrt(q))
        idx = q + nq * np.arange(K)
        model[np.ix_(idx, idx)] = JL[l]
    assert np.max(np.abs(J - model)) <= 1e-12 * scale


def test_linear_surrogate(torch_cuda):
    torch = torch_cuda
    L, K = 4, 5
    nq = (L + 1) ** 2
    (ls, vs, ics, R), plan = _fixture_plan(L, K)
    sur = cfk.Surrogate(plan, width=64, depth=1, seed=0)
    JL = sur.linear_blocks()
    rng = np.random.default_rng(5)
    B = 37
    f = np.zeros((K * nq, B))
    f[0] = 1.0 + 0.1 * rng.random(B)
    f += 1e-4 * rng.standard_normal(f.shape) * f[0]
    fd = torch.from_numpy(f).cuda()
    Q = sur.apply(fd, use_network=False).cpu().numpy()
    n = f[0]
    d = f / n - np.eye(K * nq)[:, [0]]
    ref = np.zeros_like(f)
    for q in range(nq):
        l = int(math.isqrt(q))
        idx = q + nq * np.arange(K)
        ref[idx] = n ** 2 * (JL[l] @ d[idx])
    ref[[0, 1, 2, 3, nq]] = 0.0
    assert np.max(np.abs(Q - ref)) <= 1e-12 * np.max(np.abs(ref))
    # against the exact operator: the remainder is quadratic in the 1e-4 perturbation
    Qx = torch.zeros_like(fd)
    plan.execute(Qx, fd, B, precision="fp64", layout="batch_fastest")
    Qx = Qx.cpu().numpy()
    Qx[[0, 1, 2, 3, nq]] = 0.0
    assert np.linalg.norm(Q - Qx) <= 1e-3 * np.linalg.norm(Qx)


def test_anchored_linear_surrogate(torch_cuda):
    torch = torch_cuda
    L, K = 4, 5
    nq = (L + 1) ** 2
    (ls, vs, ics, R), plan = _fixture_plan(L, K)
    sur = cfk.Surrogate(plan, width=64, depth=1, seed=0, anchored=True)
    JL = sur.linear_blocks()
    # single Maxwellians close to the unit one, so that the dropped 2 B(d_M, g) term stays small
    params = cfk.maxwellian_mixture_params(29, seed=9, max_components=1, drift_max=0.05, temp_range=(0.95, 1.05))
    cM = cfk.project_maxwellians(params, L, K)
    rng = np.random.default_rng(6)
    pert = np.zeros((K * nq, 29))
    pert[4:] = 1e-4 * rng.standard_normal((K * nq - 4, 29))   # keep mass and momentum of the Maxwellian
    pert[nq] = 0.0                                            # and its energy coefficient
    fd = cM + torch.from_numpy(pert).cuda()
    Q = sur.apply(fd, use_network=False).cpu().numpy()
    # the library's anchors: c_M against the 3-D projection of the Maxwellian with f's moments (the
    # Funk-Hecke form integrates the azimuth exactly; the projection uses 30 nodes), Q(c_M) against the
    # symmetrized FP64 operator
    cMl, QMl = sur.anchors(fd)
    mom = cfk.moments(fd, 29, nq, K, layout="batch_fastest").cpu().numpy()
    par = np.zeros((29, 1, 7))
    par[:, 0, 0], par[:, 0, 1:4], par[:, 0, 4:7] = mom[:, 0], mom[:, 1:4], mom[:, 4:5]
    cMr = cfk.project_maxwellians(par, L, K)
    assert float((cMl - cMr).abs().max()) <= 1e-12 * float(cMr.abs().max())
    QMr = torch.zeros_like(cMl)
    plan.execute(QMr, cMl, 29, precision="fp64_sym", layout="batch_fastest")
    assert float((QMl - QMr).abs().max()) <= 1e-13 * float(QMr.abs().max() + 1e-300)
    # the anchor carries the cell's mass, momentum and temperature
    momM = cfk.moments(cMl, 29, nq, K, layout="batch_fastest").cpu().numpy()
    assert np.max(np.abs(momM[:, :5] - mom[:, :5])) <= 1e-8
    # the anchored linear part Q(c_M) + n^2 J g against numpy with the plan's blocks
    f, cMh, QM = fd.cpu().numpy(), cMl.cpu().numpy(), QMl.cpu().numpy()
    n = f[0]
    g = (f - cMh) / n
    ref = QM.copy()
    for q in range(nq):
        l = int(math.isqrt(q))
        idx = q + nq * np.arange(K)
        ref[idx] += n ** 2 * (JL[l] @ g[idx])
    ref[[0, 1, 2, 3, nq]] = 0.0
    assert np.max(np.abs(Q - ref)) <= 1e-12 * np.max(np.abs(ref))
    # against the exact operator: the dropped remainder 2 B(d_M, g) + Q(g) is small against the
    # first-order change Q(f) - Q(c_M) for |d_M| ~ 0.05
    Qx = torch.zeros_like(fd)
    plan.execute(Qx, fd, 29, precision="fp64", layout="batch_fastest")
    Qx = Qx.cpu().numpy()
    Qx[[0, 1, 2, 3, nq]] = 0.0
    QM[[0, 1, 2, 3, nq]] = 0.0
    assert np.linalg.norm(Q - Qx) <= 0.5 * np.linalg.norm(Qx - QM)


@pytest.mark.parametrize("L,K", [(8, 10), (12, 14)], ids=["L8_K10", "L12_K14"])
def test_anchor_coefficients_of_drifting_maxwellians(torch_cuda, L, K):
    """Funk-Hecke anchors against the 3-D projection of the Maxwellian with the same moments, for drifts
    in every direction (including none and along -z), with the temperatures and drifts of the training
    mixtures; Q(c_M) against the symmetrized FP64 operator."""
    torch = torch_cuda
    nq = (L + 1) ** 2
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    R = cfk.build_rtensor(K - 1, L, 1.0, cfk.sumfac_grid_sizes(K - 1, L, 0, 0), geometry=(lab, val, ic))
    plan = cfk.CollisionPlan(ls, vs, ics, R, nq, K)
    sur = cfk.Surrogate(plan, width=64, depth=1, seed=0)
    params = cfk.maxwellian_mixture_params(64, seed=4)
    params[0, :, 1:4] = 0.0                       # no drift
    params[1, :, 1:4] = (0.0, 0.0, -0.4)          # along -z
    f = cfk.project_maxwellians(params, L, K)     # mixtures: their own anchors differ from them
    cM, QM = sur.anchors(f)
    mom = cfk.moments(f, 64, nq, K, layout="batch_fastest").cpu().numpy()
    par = np.zeros((64, 1, 7))
    par[:, 0, 0], par[:, 0, 1:4], par[:, 0, 4:7] = mom[:, 0], mom[:, 1:4], mom[:, 4:5]
    cMr = cfk.project_maxwellians(par, L, K)
    err = float((cM - cMr).abs().max()) / float(cMr.abs().max())
    print("anchor vs 3-D projection, L=%d K=%d: %.2e" % (L, K, err), flush=True)
    assert err <= 1e-12
    QMr = torch.zeros_like(cM)
    plan.execute(QMr, cM, 64, precision="fp64_sym", layout="batch_fastest")
    cfk.conservation_project(QMr, 64, nq, K, layout="batch_fastest")
    assert float((QM - QMr).abs().max()) <= 1e-13 * float(QMr.abs().max())
