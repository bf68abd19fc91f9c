"""Tensor generation at L=12, K=12 radial modes, 48 nodes on every quadrature axis, gamma in {0,1}:
the GPU quadrature through the 26-argument MEX entry point against the oracle on one x node (the
sum over x nodes is additive, tests/test_oracle_sumfac.py), at 1e-12 relative."""
import os

import numpy as np
import pytest

import cfk
import oracle
from oracle import geometry as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def geometry12():
    return G.real_gaunt_geometry(12)


@pytest.mark.parametrize("gamma", [0.0, 1.0])
def test_config3_sumfac_matches_oracle_on_one_x_node(geometry12, gamma):
    """The exact-sphere kernel (the default for this rule) and the matrix-core quadrature against the
    oracle on one x node."""
    L, K_max = 12, 11
    lab, val, ic = geometry12
    grids = G.sumfac_grids(gamma, (48,) * 7)
    args = G.sumfac_prep(K_max, L, gamma, lab, val, ic, grids)
    prhs = [args[k] for k in oracle.SUMFAC_ORDER]
    xi = 20
    out = {}
    try:
        for kind in ("auto", "matrix_core"):
            print("config3 gamma=%g: GPU x node %d, kernel %s" % (gamma, xi, kind), flush=True)
            cfk.set_sumfac_kernel(kind)
            out[kind] = cfk.compute_rtensor_sumfac_mex(*prhs, x_range=(xi, xi + 1)).ravel(order="F")
            if kind == "auto":
                assert cfk.sumfac_last_timing()["kernel"] == "exact_sphere"
    finally:
        cfk.set_sumfac_kernel("auto")
    print("config3 gamma=%g: oracle" % gamma, flush=True)
    R_ref = oracle.sumfac(args, xi, xi + 1, nthreads=min(16, os.cpu_count() or 1))
    scale = np.max(np.abs(R_ref))
    assert scale > 0
    for kind, R in out.items():
        err = np.max(np.abs(R - R_ref)) / scale
        print("  %s: %.3e" % (kind, err), flush=True)
        assert err <= 1e-12, kind


@pytest.fixture
def chunk_cap():
    """Restores the automatic chunking after a test that caps the x nodes per chunk."""
    yield cfk.set_sumfac_chunk_nodes
    cfk.set_sumfac_chunk_nodes(0)


@pytest.mark.parametrize("kind", ["matrix_core", "per_channel", "exact_sphere"])
def test_chunked_x_nodes_match_oracle(chunk_cap, kind):
    """The x nodes split into 3 chunks (each its own stage-A and fold launches over the unit buffers,
    compute_rtensor_sumfac_mex.cpp:126-134 sums them in one loop) against the oracle at 1e-12, and
    against the same tensor computed in one chunk."""
    L, K_max, gamma = 4, 3, 1.0
    sizes = (7, 4, 3, 4, 3, 6, 11)   # (chi, eps) rule exact for the gain integrand (degree L + K_max = 7)
    lab, val, ic = G.real_gaunt_geometry(L)
    args = G.sumfac_prep(K_max, L, gamma, lab, val, ic, G.sumfac_grids(gamma, sizes))
    prhs = [args[k] for k in oracle.SUMFAC_ORDER]
    ref = oracle.sumfac(args, nthreads=min(16, os.cpu_count() or 1))
    scale = np.max(np.abs(ref))
    try:
        cfk.set_sumfac_kernel(kind)
        one = cfk.compute_rtensor_sumfac_mex(*prhs).ravel(order="F")
        assert cfk.sumfac_last_chunks() == 1
        chunk_cap(3)
        print("sumfac %s, 7 x nodes in chunks of <= 3" % kind, flush=True)
        three = cfk.compute_rtensor_sumfac_mex(*prhs).ravel(order="F")
        assert cfk.sumfac_last_chunks() == 3
    finally:
        cfk.set_sumfac_kernel("auto")
    assert np.max(np.abs(three - ref)) <= 1e-12 * scale
    assert np.max(np.abs(three - one)) <= 1e-13 * scale


def test_build_rtensor_l12_k12_device_tables_match_oracle(geometry12, chunk_cap):
    """cfk_build_rtensor at L=12, K=12 (P tables built on the device from the Gaunt COO, matrix-core
    quadrature, conservation zeroing) against the oracle's prep (src/GeneralCollisionTensor.m:43-192
    restated), sumfac and zeroing on a reduced grid, with the 3 x nodes in 3 chunks."""
    L, K_max, gamma = 12, 11, 1.0
    sizes = (3, 6, 4, 5, 3, 8, 7)
    lab, val, ic = geometry12
    args = G.sumfac_prep(K_max, L, gamma, lab, val, ic, G.sumfac_grids(gamma, sizes))
    print("L=12 K=12 oracle prep + sumfac on %s" % (sizes,), flush=True)
    ref = G.conservation_zero(oracle.sumfac(args, nthreads=min(16, os.cpu_count() or 1)),
                              args["L_triplets"].astype(int), K_max + 1)
    chunk_cap(1)
    print("L=12 K=12 build_rtensor (device P tables), chunks of 1 x node", flush=True)
    R = cfk.build_rtensor(K_max, L, gamma, sizes, geometry=cfk.real_gaunt(L))
    assert cfk.sumfac_last_chunks() == 3
    scale = np.max(np.abs(ref))
    assert scale > 0
    assert np.max(np.abs(R - ref)) <= 1e-12 * scale
