"""Pins of the oracle sumfac (oracle/cfk_oracle.c: oracle_sumfac, restating
compute_rtensor_sumfac_mex.cpp:80-398) and of the prep restatement (oracle/geometry.py:
sumfac_prep, GeneralCollisionTensor.m:43-192).

* End to end against the reference's own R fixtures: precompute_collision_operator.m:19-25
  records pad 20 for the gamma=0 files; the gamma=1 files reproduce with pad 20 as well.  The
  chain geometry -> grids -> P tables -> quadrature -> conservation zeroing must reproduce R.
* Additivity over x nodes (the bounded-sample decomposition used by the GPU parity tests).
* The WCU test through the sumfac lives in test_oracle_contract.py.
"""
import numpy as np
import pytest

import oracle
from conftest import load_precalc
from oracle import geometry as G


def build(L, gamma, K_max=4, pad=20):
    lab, val, ic = G.real_gaunt_geometry(L)
    grids = G.sumfac_grids(gamma, G.sumfac_grid_sizes(K_max, L, pad, pad))
    return G.sumfac_prep(K_max, L, gamma, lab, val, ic, grids)


@pytest.mark.parametrize("L,gamma", [(2, 0.0), (2, 1.0), (4, 0.0)])
def test_sumfac_reproduces_reference_fixture(L, gamma):
    args = build(L, gamma)
    R = G.conservation_zero(oracle.sumfac(args), args["L_triplets"].astype(int), 5)
    Rf = load_precalc(L, gamma)[3]
    assert np.max(np.abs(R - Rf)) <= 1e-13 * np.max(np.abs(Rf))


def test_sumfac_additive_over_x_nodes():
    args = build(2, 1.0, K_max=3, pad=4)
    n_x = args["x"].size
    full = oracle.sumfac(args, nthreads=1)
    parts = oracle.sumfac(args, 0, 5, nthreads=1) + oracle.sumfac(args, 5, n_x, nthreads=2)
    assert np.max(np.abs(full - parts)) <= 1e-13 * np.max(np.abs(full))
