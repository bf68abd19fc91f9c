"""Dense contraction (dense_tensor_kernel_mex) and its producer (assemble_full_tensor) on the GPU.

* cfk_dense_assemble vs the oracle's assemble_dense (src/GeneralCollisionTensor.m:283-311): bit for bit
  (each entry is one product R g);
* cfk_dense_execute vs the oracle's dense kernel (dense_tensor_kernel_mex.cpp:14-29) at 1e-12 relative,
  for even N (16-byte column pairs) and odd N (single columns), batches that need two passes, and N above
  1024 (one distribution per pass) against a float64 matrix-vector product of torch on the GPU.
"""
import numpy as np
import pytest

import cfk
import oracle
from oracle import geometry as G
from cfk.inputs import perturbed_maxwellians

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _operator(L, K, seed):
    lab, val, ic = cfk.real_gaunt(L)
    n_l = int(np.max(ic))
    R = np.random.default_rng(seed).standard_normal(n_l * K ** 3)
    return np.asarray(lab).astype(np.int64), np.asarray(val), np.asarray(ic).astype(np.int64), R


@pytest.mark.parametrize("L,K", [(2, 5), (4, 6)])
def test_dense_assemble_matches_oracle_bitwise(torch_cuda, L, K):
    torch = torch_cuda
    lab, val, ic, R = _operator(L, K, seed=L)
    nq = (L + 1) ** 2
    N = nq * K
    C = torch.full((N ** 3,), float("nan"), dtype=torch.float64, device="cuda")
    cfk.dense_assemble(C, lab, val, ic, R, nq, K)
    torch.cuda.synchronize()
    Cref = G.assemble_dense(lab, val, ic, R, nq, K)
    assert np.array_equal(C.cpu().numpy(), Cref)


@pytest.mark.parametrize("L,K,B", [(2, 5, 5), (4, 6, 1), (6, 8, 2)])
def test_dense_execute_matches_oracle(torch_cuda, L, K, B):
    torch = torch_cuda
    lab, val, ic, R = _operator(L, K, seed=10 + L)
    nq = (L + 1) ** 2
    N = nq * K
    print("dense: L=%d K=%d N=%d B=%d" % (L, K, N, B), flush=True)
    C = torch.empty(N ** 3, dtype=torch.float64, device="cuda")
    cfk.dense_assemble(C, lab, val, ic, R, nq, K)
    f = perturbed_maxwellians(L, K, B, seed=3).reshape(B, N)
    Q0 = np.random.default_rng(4).standard_normal((B, N))
    Q = torch.from_numpy(Q0.copy()).cuda()
    cfk.dense_execute(Q, torch.from_numpy(f).cuda(), C, N, B)
    torch.cuda.synchronize()
    Cref = C.cpu().numpy()
    Qg = Q.cpu().numpy()
    for b in range(B):
        qref = oracle.dense(Q0[b].copy(), f[b].copy(), Cref, N)
        assert np.max(np.abs(Qg[b] - qref)) <= TOL * np.max(np.abs(qref - Q0[b])), b


@pytest.mark.parametrize("N", [1026, 1027])
def test_dense_execute_large_n(torch_cuda, N):
    """N above 1024: one distribution per pass; compared with C^T w, w[k N + j] = f[j] f[k], in float64."""
    torch = torch_cuda
    B = 2
    print("dense: N=%d B=%d" % (N, B), flush=True)
    g = torch.Generator(device="cuda").manual_seed(N)
    C = torch.randn(N ** 3, dtype=torch.float64, device="cuda", generator=g)
    f = torch.randn(B, N, dtype=torch.float64, device="cuda", generator=g)
    Q = torch.zeros(B, N, dtype=torch.float64, device="cuda")
    cfk.dense_execute(Q, f, C, N, B)
    Cm = C.view(N * N, N)
    for b in range(B):
        w = torch.outer(f[b], f[b]).reshape(-1)   # w[k*N + j] = f[k] f[j]
        ref = Cm.t() @ w
        assert float((Q[b] - ref).abs().max()) <= TOL * float(ref.abs().max()) * 10, b
