"""Golden radial tensors computed by the CPU oracle (oracle/cfk_oracle.c restates
compute_rtensor_sumfac_mex.cpp:80-398; oracle/geometry.py restates the MATLAB prep
src/GeneralCollisionTensor.m:43-192).

    sumfac_hs_l8_k10_g48.npz : gamma=1, L=8, K_max=9 (10 radial modes), 48 nodes on every axis
                               (config "GL 32 points per dimension, padding 1.5x"), R after
                               conservation zeroing, plus the grid sizes used.

Run once (about 10 minutes on 8 cores); the result is committed.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

import oracle  # noqa: E402
from oracle import geometry as G  # noqa: E402


def main():
    L, K_max, gamma = 8, 9, 1.0
    sizes = (48,) * 7
    lab, val, ic = G.real_gaunt_geometry(L)
    grids = G.sumfac_grids(gamma, sizes)
    args = G.sumfac_prep(K_max, L, gamma, lab, val, ic, grids)
    t0 = time.time()
    R = oracle.sumfac(args, nthreads=8)
    R = G.conservation_zero(R, args["L_triplets"].astype(int), K_max + 1)
    print("sumfac oracle %.1fs" % (time.time() - t0))
    np.savez_compressed(os.path.join(HERE, "sumfac_hs_l8_k10_g48.npz"), R=R, sizes=np.array(sizes),
                        meta=np.array([L, K_max, gamma]))


if __name__ == "__main__":
    main()
