"""The GPU quadrature kernels of compute_rtensor_sumfac (src/mex/compute_rtensor_sumfac_mex.cpp:80-398)
against the oracle at 1e-12 relative to max |R|: the matrix-core form (gain-point sums once per SH
degree on FP64 MFMA, then the P_gain slot contraction and a matrix-core fold over units), the
per-channel form (the reference's summation order) and the exact-sphere form (the gain sum of a unit
step as its exact surface integral, where the (chi, eps) rule is exact for the integrand's degree).

Ragged grids put the gain-point count off the 12-point batch and the radial count at the edges of
the 16-row matrix tiles (N_K = 1, 3, 16); degrees up to 12 give SH blocks of 2 column tiles.  The
exact-sphere cases sit at the smallest (chi, eps) rule exact for the gain integrand's degree on the
sphere, L + K_max, and at the sizes of src/GeneralCollisionTensor.m:63-64; one node fewer must
make the exact-sphere kernel refuse."""
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
        cfk.set_sumfac_kernel("auto")
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
        cfk.set_sumfac_kernel("auto")
    scale = np.max(np.abs(out["per_channel"]))
    assert scale > 0
    assert np.max(np.abs(out["matrix_core"] - out["per_channel"])) <= TOL64 * scale


def sphere_sizes(L, K_max, rule):
    """(N_chi, N_eps): "min", the smallest product rule exact for degree L + K_max (the gain
    integrand's degree on the sphere); "ref", the reference's sizes (src/GeneralCollisionTensor.m:63-64)."""
    if rule == "ref":
        return K_max + (L + 1) // 2 + 1, 2 * K_max + L + 1
    d = L + K_max
    return (d + 2) // 2, d + 1


SPHERE_CASES = [
    # (L, K_max, gamma, (N_x, N_u1, N_t1, N_y2, N_t2), rule)
    (2, 0, 0.0, (3, 3, 2, 3, 2), "min"),
    (3, 2, 1.0, (4, 3, 4, 3, 2), "min"),
    (6, 4, 0.0, (3, 4, 3, 4, 3), "ref"),
    (9, 5, 1.0, (2, 3, 2, 3, 2), "min"),
    (2, 15, 0.0, (3, 3, 3, 3, 3), "min"),
    (12, 11, 1.0, (2, 3, 2, 3, 2), "ref"),
    (12, 11, 0.0, (2, 3, 2, 3, 2), "min"),
]


@pytest.mark.parametrize("L,K_max,gamma,base,rule", SPHERE_CASES)
def test_exact_sphere_kernel_matches_oracle(L, K_max, gamma, base, rule):
    sizes = base + sphere_sizes(L, K_max, rule)
    lab, val, ic = G.real_gaunt_geometry(L)
    args = G.sumfac_prep(K_max, L, gamma, lab, val, ic, G.sumfac_grids(gamma, sizes))
    prhs = [args[k] for k in oracle.SUMFAC_ORDER]
    out = {}
    try:
        for kind in ("exact_sphere", "auto", "matrix_core"):
            print("sumfac kernel %s: L=%d K_max=%d gamma=%g sizes=%s" % (kind, L, K_max, gamma, sizes), flush=True)
            cfk.set_sumfac_kernel(kind)
            out[kind] = cfk.compute_rtensor_sumfac_mex(*prhs).ravel(order="F")
            used = cfk.sumfac_last_timing()["kernel"]
            assert used == ("exact_sphere" if kind == "auto" else kind)
    finally:
        cfk.set_sumfac_kernel("auto")
    ref = oracle.sumfac(args, nthreads=min(16, os.cpu_count() or 1))
    scale = np.max(np.abs(ref))
    assert scale > 0
    err = np.max(np.abs(out["exact_sphere"] - ref)) / scale
    print("  exact_sphere: max |R - R_oracle| / max |R_oracle| = %.3e" % err, flush=True)
    assert err <= TOL64
    assert np.array_equal(out["auto"], out["exact_sphere"])
    assert np.max(np.abs(out["exact_sphere"] - out["matrix_core"])) <= TOL64 * scale


@pytest.mark.parametrize("short", ["chi", "eps"])
def test_exact_sphere_needs_an_exact_rule(short):
    """One chi node or one eps node below the exact rule: the exact-sphere kernel refuses and "auto"
    runs the matrix-core quadrature (still equal to the oracle)."""
    L, K_max, gamma = 3, 2, 1.0
    n_chi, n_eps = sphere_sizes(L, K_max, "min")
    if short == "chi":
        n_chi -= 1
    else:
        n_eps -= 1
    sizes = (3, 3, 2, 3, 2, n_chi, n_eps)
    lab, val, ic = G.real_gaunt_geometry(L)
    args = G.sumfac_prep(K_max, L, gamma, lab, val, ic, G.sumfac_grids(gamma, sizes))
    prhs = [args[k] for k in oracle.SUMFAC_ORDER]
    try:
        cfk.set_sumfac_kernel("exact_sphere")
        with pytest.raises(cfk.CfkError):
            cfk.compute_rtensor_sumfac_mex(*prhs)
        cfk.set_sumfac_kernel("auto")
        R = cfk.compute_rtensor_sumfac_mex(*prhs).ravel(order="F")
        assert cfk.sumfac_last_timing()["kernel"] == "matrix_core"
    finally:
        cfk.set_sumfac_kernel("auto")
    ref = oracle.sumfac(args, nthreads=min(16, os.cpu_count() or 1))
    assert np.max(np.abs(R - ref)) <= TOL64 * np.max(np.abs(ref))


def test_unknown_kernel_is_rejected():
    with pytest.raises(KeyError):
        cfk.set_sumfac_kernel("tensor")
    with pytest.raises(cfk.CfkError):
        cfk._lib.check(cfk.lib().cfk_sumfac_set_kernel(7))
