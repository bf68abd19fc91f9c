"""Batches that span several column groups of the angular-first launches (the grid runs groups of
1024 distributions for the symmetrized FP64 and mixed kernels, 4096 for the plain FP64 kernels, each
group's rows heaviest first), with a partial last group: B = 1100 (1024 + 76) and B = 5000 (4 x 1024 +
904; 4096 + 904 for plain FP64).  Config-0 geometry (gamma=0, L=4, K=6) with its golden R; every
precision, tile shape (waves per workgroup, row split, distributions per lane, VALU or matrix-core
radial stage) and layout against the oracle's angular-first kernel
(src/mex/angular_first_collision_kernel_mex.cpp:6-76): FP64 paths at 1e-12, mixed at 1e-6 (normwise
per distribution), the mixed path's five invariants exactly 0."""
import os

import numpy as np
import pytest

import cfk
import oracle
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

L, K = 4, 6
NQ = (L + 1) ** 2

# (precision, tolerance, pf, waves per workgroup, row split, distributions per lane)
CASES = [
    ("fp64", 1e-12, 0, 1, 0, 1),
    ("fp64", 1e-12, 0, 2, 0, 2),
    ("fp64_sym", 1e-12, 0, 1, 0, 1),
    ("fp64_sym", 1e-12, 0, 2, 1, 1),
    ("fp64_sym", 1e-12, 0, 4, 0, 2),
    ("mixed", 1e-6, 0, 1, 0, 1),
    ("mixed", 1e-6, 0, 2, 1, 1),
    ("mixed", 1e-6, 0, 4, 0, 2),
    ("mixed", 1e-6, 1, 1, 0, 1),
    ("mixed", 1e-6, 1, 2, 1, 1),
]


@pytest.fixture(scope="module")
def setup(torch_cuda):
    d = np.load(os.path.join(GOLDEN, "config0_maxwell_l4_k6.npz"))
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    plan = cfk.CollisionPlan(ls, vs, ics, d["R"], NQ, K)
    yield dict(torch=torch_cuda, plan=plan, lab=ls, val=vs, ic=ics, R=d["R"])
    plan.close()


def rel_rows(Q, ref):
    return float(np.max(np.max(np.abs(Q - ref), axis=1) / np.max(np.abs(ref), axis=1)))


@pytest.mark.parametrize("B", [1100, 5000])
def test_column_groups_match_oracle(setup, B):
    torch, plan = setup["torch"], setup["plan"]
    rng = np.random.default_rng(7000 + B)
    f = rng.normal(0.0, 0.05, size=(B, K, NQ))
    f[:, 0, 0] += 1.0
    ref = np.zeros(B * NQ * K)
    oracle.angular_first_batch(ref, f.reshape(-1), setup["lab"], setup["val"], setup["ic"], setup["R"], NQ, K,
                               nthreads=8)
    ref = ref.reshape(B, -1)
    fd = torch.tensor(f.reshape(-1), device="cuda")
    fT = fd.reshape(B, -1).t().contiguous()
    inv = [0, 1, 2, 3, NQ]
    try:
        for prec, tol, pf, wpb, split, nb in CASES:
            print("column groups: B=%d %s pf=%d waves=%d split=%d nb=%d" % (B, prec, pf, wpb, split, nb), flush=True)
            plan.set_tile(prec, 0, nb, pf)
            plan.set_waves(prec, wpb)
            plan.set_split(prec, split)
            Qn = torch.zeros(B * NQ * K, dtype=torch.float64, device="cuda")
            plan.execute(Qn, fd, B, precision=prec)
            Qb = torch.zeros_like(fT)
            plan.execute(Qb, fT, B, precision=prec, layout="batch_fastest")
            torch.cuda.synchronize()
            Qn = Qn.cpu().numpy().reshape(B, -1)
            Qb = Qb.t().contiguous().cpu().numpy()
            assert rel_rows(Qn, ref) <= tol, (B, prec, pf, wpb, split, nb, "natural")
            assert rel_rows(Qb, ref) <= tol, (B, prec, pf, wpb, split, nb, "batch_fastest")
            if prec == "mixed":
                assert float(np.abs(Qn[:, inv]).max()) == 0.0 and float(np.abs(Qb[:, inv]).max()) == 0.0
    finally:
        for prec in ("fp64", "fp64_sym", "mixed"):
            plan.set_tile(prec)
            plan.set_waves(prec, 0)
            plan.set_split(prec, -1)
