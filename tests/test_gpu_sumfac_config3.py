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
    L, K_max = 12, 11
    lab, val, ic = geometry12
    grids = G.sumfac_grids(gamma, (48,) * 7)
    args = G.sumfac_prep(K_max, L, gamma, lab, val, ic, grids)
    prhs = [args[k] for k in oracle.SUMFAC_ORDER]
    xi = 20
    print("config3 gamma=%g: GPU x node %d" % (gamma, xi), flush=True)
    R_gpu = cfk.compute_rtensor_sumfac_mex(*prhs, x_range=(xi, xi + 1)).ravel(order="F")
    print("config3 gamma=%g: oracle" % gamma, flush=True)
    R_ref = oracle.sumfac(args, xi, xi + 1, nthreads=min(16, os.cpu_count() or 1))
    scale = np.max(np.abs(R_ref))
    assert scale > 0
    assert np.max(np.abs(R_gpu - R_ref)) <= 1e-12 * scale
