"""cfk_plan_execute_static (the entry point that allocates nothing) against cfk_plan_execute.

Both run the same kernels on the same inputs, so the results must be identical bit for bit, on two
calls in a row (the second with a smaller batch, reusing the fixed staging area), for every precision
and layout, on both mixed-precision kernels; a batch whose staging exceeds the fixed area is rejected
before anything runs."""
import os

import numpy as np
import pytest

import cfk
from cfk.inputs import perturbed_maxwellians
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plan(torch_cuda):
    L, K = 8, 10
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    R = np.load(os.path.join(GOLDEN, "sumfac_hs_l8_k10_g48.npz"))["R"]
    p = cfk.CollisionPlan(ls, vs, ics, R, (L + 1) ** 2, K)
    yield p
    p.close()


def _inputs(torch, L, K, B, layout, seed):
    f = torch.from_numpy(perturbed_maxwellians(L, K, B, seed=seed)).cuda()
    if layout == "batch_fastest":
        f = f.reshape(B, -1).t().contiguous()
    base = torch.from_numpy(np.random.default_rng(seed).standard_normal(f.numel())).cuda().reshape(f.shape)
    return f, base


@pytest.mark.parametrize("precision,pf", [("mixed", 0), ("mixed", 1), ("fp64", 0), ("fp64_sym", 0)])
@pytest.mark.parametrize("layout", ["natural", "batch_fastest"])
def test_static_matches_allocating_entry_point(plan, torch_cuda, precision, pf, layout):
    torch = torch_cuda
    if precision == "mixed":
        plan.set_tile("mixed", 0, 1, pf)
    try:
        for B, seed in ((300, 1), (130, 2)):   # second call smaller: the area keeps the first call's data
            f, base = _inputs(torch, 8, 10, B, layout, seed)
            Q_ref = base.clone()
            plan.execute(Q_ref, f, B, precision=precision, layout=layout)
            Q = base.clone()
            print("static execute %s pf=%d %s B=%d" % (precision, pf, layout, B), flush=True)
            plan.execute_static(Q, f, B, precision=precision, layout=layout)
            torch.cuda.synchronize()
            assert torch.equal(Q, Q_ref), (precision, pf, layout, B)
    finally:
        plan.set_tile("mixed")


def test_static_rejects_a_batch_beyond_the_fixed_area(plan, torch_cuda):
    torch = torch_cuda
    # mixed precision, natural layout: 81*12 FP32 + 81*10 FP64 staging values per distribution
    # (10,368 bytes), so 8000 distributions need 83 MB > 64 MiB
    B = 8000
    f = torch.zeros(B * 810, dtype=torch.float64, device="cuda")
    Q = torch.full_like(f, 0.5)
    with pytest.raises(cfk.CfkError) as e:
        plan.execute_static(Q, f, B, precision="mixed", layout="natural")
    assert e.value.code == -1 and "fixed area" in str(e.value)
    torch.cuda.synchronize()
    assert bool((Q == 0.5).all())   # nothing ran
    # the batch-fastest FP64 operator needs no staging at all
    fT = torch.from_numpy(perturbed_maxwellians(8, 10, 64, seed=3)).cuda().reshape(64, -1).t().contiguous()
    Q1 = torch.zeros_like(fT)
    Q2 = torch.zeros_like(fT)
    plan.execute(Q1, fT, 64, precision="fp64", layout="batch_fastest")
    plan.execute_static(Q2, fT, 64, precision="fp64", layout="batch_fastest")
    torch.cuda.synchronize()
    assert torch.equal(Q1, Q2)
