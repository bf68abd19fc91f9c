"""The two GPU quadrature kernels of compute_rtensor_sumfac (src/mex/compute_rtensor_sumfac_mex.cpp:80-398):
the matrix-core form (gain-point sums once per SH degree on FP64 MFMA, then the P_gain slot
contraction and a matrix-core fold over units) and the per-channel form (the reference's summation
order), each against the oracle at 1e-12 relative to max |R|.

Ragged grids put the gain-point count off the 12-point batch and the radial count at the edges of
the 16-row matrix tiles (N_K = 1, 3, 16); degrees up to 12 give SH blocks of 2 column tiles."""
import os

import numpy as np
import pytest

import cfk
import oracle
from oracle import geometry as G

pytestmark = pytest.mark.gpu

TOL64 = 1e-12

CASES = [
    # (L, K_max, gamma, (N_x, N_u1, N_t1, N_y2, N_t2, N_chi, N_eps))
    (2, 0, 0.0, (3, 3, 2, 3, 2, 3, 5)),
    (3, 2, 1.0, (4, 3, 4, 3, 2, 5, 7)),
    (6, 7, 1.0, (3, 5, 4, 4, 3, 7, 5)),
    (2, 15, 0.0, (3, 3, 3, 3, 3, 4, 6)),
    (9, 5, 1.0, (2, 3, 2, 3, 2, 9, 4)),
]


def run_both(L, K_max, gamma, sizes, x_range=None):
    lab, val, ic = G.real_gaunt_geometry(L)
    grids = G.sumfac_grids(gamma, sizes)
    args = G.sumfac_prep(K_max, L, gamma, lab, val, ic, grids)
    prhs = [args[k] for k in oracle.SUMFAC_ORDER]
    out = {}
    try:
        for kind in ("per_channel", "matrix_core"):
            print("sumfac kernel %s: L=%d K_max=%d gamma=%g sizes=%s" % (kind, L, K_max, gamma, sizes), flush=True)
            cfk.set_sumfac_kernel(kind)
            out[kind] = cfk.compute_rtensor_sumfac_mex(*prhs, x_range=x_range).ravel(order="F")
    finally:
        cfk.set_sumfac_kernel("matrix_core")
    a, b = (0, sizes[0]) if x_range is None else x_range
    ref = oracle.sumfac(args, a, b, nthreads=min(16, os.cpu_count() or 1))
    return out, ref


@pytest.mark.parametrize("L,K_max,gamma,sizes", CASES)
def test_both_kernels_match_oracle_on_ragged_grids(L, K_max, gamma, sizes):
    out, ref = run_both(L, K_max, gamma, sizes)
    scale = np.max(np.abs(ref))
    assert scale > 0
    for kind, R in out.items():
        err = np.max(np.abs(R - ref)) / scale
        print("  %s: max |R - R_oracle| / max |R_oracle| = %.3e" % (kind, err), flush=True)
        assert err <= TOL64, kind
    assert np.max(np.abs(out["matrix_core"] - out["per_channel"])) <= TOL64 * scale


def test_kernels_agree_at_l12_k12_on_two_x_nodes():
    """Tensor generation shape of the L=12, K=12 case (48 nodes per axis), x nodes 30..31."""
    L, K_max = 12, 11
    lab, val, ic = G.real_gaunt_geometry(L)
    grids = G.sumfac_grids(1.0, (48,) * 7)
    args = G.sumfac_prep(K_max, L, 1.0, lab, val, ic, grids)
    prhs = [args[k] for k in oracle.SUMFAC_ORDER]
    out = {}
    try:
        for kind in ("per_channel", "matrix_core"):
            print("L=12 K=12 kernel %s" % kind, flush=True)
            cfk.set_sumfac_kernel(kind)
            out[kind] = cfk.compute_rtensor_sumfac_mex(*prhs, x_range=(30, 32)).ravel(order="F")
    finally:
        cfk.set_sumfac_kernel("matrix_core")
    scale = np.max(np.abs(out["per_channel"]))
    assert scale > 0
    assert np.max(np.abs(out["matrix_core"] - out["per_channel"])) <= TOL64 * scale


def test_unknown_kernel_is_rejected():
    with pytest.raises(KeyError):
        cfk.set_sumfac_kernel("tensor")
    with pytest.raises(cfk.CfkError):
        cfk._lib.check(cfk.lib().cfk_sumfac_set_kernel(7))
