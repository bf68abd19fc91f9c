"""Every radial size the kernels are instantiated for: hard spheres at L=4 with K = 1..16 (odd and even,
the matrix-core kernel's K in {6,8,10,12,14} on both kernels), FP64, symmetrized FP64 and mixed
precision against the oracle (the C restatement of angular_first_collision_kernel_mex.cpp:6-76) on a
ragged batch of 70 distributions, from batch-fastest and natural-layout input.  FP64 paths at 1e-12,
mixed at 1e-6 (normwise per distribution), the five invariants of the mixed path exactly 0.
"""
import numpy as np
import pytest

import cfk
import oracle
from cfk.inputs import perturbed_maxwellians

pytestmark = pytest.mark.gpu

L = 4
NQ = (L + 1) ** 2
B = 70


@pytest.fixture(scope="module")
def geometry(torch_cuda):
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    return lab, val, ic, ls, vs, ics


def rel_rows(Q, Qref):
    return float(np.max(np.max(np.abs(Q - Qref), axis=1) / np.max(np.abs(Qref), axis=1)))


@pytest.mark.parametrize("K", list(range(1, 17)), ids=lambda k: "K%d" % k)
def test_every_k_matches_oracle(geometry, torch_cuda, K):
    torch = torch_cuda
    lab, val, ic, ls, vs, ics = geometry
    R = cfk.build_rtensor(K - 1, L, 1.0, cfk.sumfac_grid_sizes(K - 1, L, 0, 0), geometry=(lab, val, ic))
    plan = cfk.CollisionPlan(ls, vs, ics, R, NQ, K)
    f = perturbed_maxwellians(L, K, B, seed=500 + K)
    ref = oracle.angular_first_batch(np.zeros(B * NQ * K), f, ls, vs, ics, R, NQ, K, nthreads=8).reshape(B, -1)
    fd = torch.from_numpy(f).cuda()
    fT = fd.reshape(B, -1).t().contiguous()
    kernels = [None] + ([0, 1] if K in (6, 8, 10, 12, 14) else [])
    for prec, tol in (("fp64", 1e-12), ("fp64_sym", 1e-12), ("mixed", 1e-6)):
        for pf in (kernels if prec == "mixed" else [None]):
            if pf is not None:
                plan.set_tile("mixed", 0, 0, pf)
            print("K=%d %s pf=%s" % (K, prec, pf), flush=True)
            Qn = torch.zeros(B * NQ * K, dtype=torch.float64, device="cuda")
            plan.execute(Qn, fd, B, precision=prec)
            QT = torch.zeros_like(fT)
            plan.execute(QT, fT, B, precision=prec, layout="batch_fastest")
            torch.cuda.synchronize()
            plan.set_tile("mixed")
            Qn = Qn.cpu().numpy().reshape(B, -1)
            Qb = QT.t().contiguous().cpu().numpy()
            assert rel_rows(Qn, ref) <= tol, (prec, pf)
            assert rel_rows(Qb, ref) <= tol, (prec, pf)
            if prec == "mixed":
                inv = [0, 1, 2, 3] + ([NQ] if K > 1 else [])
                assert float(np.abs(Qn[:, inv]).max()) == 0.0 and float(np.abs(Qb[:, inv]).max()) == 0.0
    plan.close()
