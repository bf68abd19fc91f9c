"""Edge cases of the batched contraction and the sumfac entry point on the GPU: empty batches,
batches of 1 and around the 64-lane wave (63, 65), the batch-size limit (f must stay below 4 GiB,
rejected before any launch), and an empty x-node range.  Config-0 geometry (γ=0, L=4, K=6) with
its golden R; every size against the oracle's angular-first kernel
(src/mex/angular_first_collision_kernel_mex.cpp:6-76)."""
import ctypes
import os

import numpy as np
import pytest

import cfk
import oracle
from conftest import GOLDEN
from oracle import geometry as G

pytestmark = pytest.mark.gpu

L, K = 4, 6
NQ = (L + 1) ** 2


@pytest.fixture(scope="module")
def setup(torch_cuda):
    d = np.load(os.path.join(GOLDEN, "config0_maxwell_l4_k6.npz"))
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    plan = cfk.CollisionPlan(ls, vs, ics, d["R"], NQ, K)
    return dict(torch=torch_cuda, plan=plan, lab=ls, val=vs, ic=ics, R=d["R"])


def batch(B, seed):
    rng = np.random.default_rng(seed)
    f = rng.normal(0.0, 0.05, size=(B, K, NQ))
    f[:, 0, 0] += 1.0
    return f


@pytest.mark.parametrize("B", [1, 63, 65])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp64_sym", 1e-12), ("mixed", 1e-6)])
def test_small_batches_match_oracle(setup, B, precision, tol):
    torch = setup["torch"]
    f = batch(B, B)
    ref = np.zeros(B * NQ * K)
    oracle.angular_first_batch(ref, f.reshape(-1), setup["lab"], setup["val"], setup["ic"], setup["R"], NQ, K)
    print("edges: B=%d %s" % (B, precision), flush=True)
    fd = torch.tensor(f.reshape(-1), device="cuda")
    Qd = torch.zeros(B * NQ * K, dtype=torch.float64, device="cuda")
    setup["plan"].execute(Qd, fd, B, precision=precision)
    torch.cuda.synchronize()
    Q = Qd.cpu().numpy().reshape(B, -1)
    ref = ref.reshape(B, -1)
    err = np.max(np.max(np.abs(Q - ref), axis=1) / np.max(np.abs(ref), axis=1))
    assert err <= tol, (B, precision, err)


def test_empty_batch_leaves_q_unchanged(setup):
    torch = setup["torch"]
    Qd = torch.full((NQ * K,), 3.5, dtype=torch.float64, device="cuda")
    fd = torch.ones(NQ * K, dtype=torch.float64, device="cuda")
    for precision in ("fp64", "fp64_sym", "mixed"):
        setup["plan"].execute(Qd, fd, 0, precision=precision)
    torch.cuda.synchronize()
    assert bool((Qd == 3.5).all())


def test_batch_above_the_limit_is_rejected_before_launch(setup):
    """B beyond min((2^32-1)/(8 N_Q K), (2^31-1)/(4 N_Q KP)) distributions (KP = K rounded up to
    4): CFK_EINVAL from the argument checks; the buffers passed are tiny and never touched."""
    torch = setup["torch"]
    limit = min((2 ** 32 - 1) // (8 * NQ * K), (2 ** 31 - 1) // (4 * NQ * ((K + 3) // 4 * 4)))
    Qd = torch.zeros(NQ * K, dtype=torch.float64, device="cuda")
    fd = torch.zeros(NQ * K, dtype=torch.float64, device="cuda")
    for prec in (0, 1, 2):
        rc = cfk.lib().cfk_plan_execute(setup["plan"]._h, 0, prec, 1, ctypes.c_void_p(Qd.data_ptr()),
                                        ctypes.c_void_p(fd.data_ptr()), limit + 1, None)
        assert rc == -1 and "exceeds" in cfk.lib().cfk_last_error().decode()
    torch.cuda.synchronize()
    assert bool((Qd == 0).all())


def test_sumfac_empty_x_range_gives_zero_tensor():
    lab, val, ic = G.real_gaunt_geometry(2)
    args = G.sumfac_prep(2, 2, 0.0, lab, val, ic, G.sumfac_grids(0.0, (4, 3, 3, 3, 3, 4, 5)))
    prhs = [args[k] for k in oracle.SUMFAC_ORDER]
    R = cfk.compute_rtensor_sumfac_mex(*prhs, x_range=(2, 2))
    assert R.size == 27 * int(args["N_L"]) and not np.any(R)
