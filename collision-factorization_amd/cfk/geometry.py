"""Host-side preparation exposed from libcfk (native restatement of the reference's MATLAB host code).

    real_gaunt          src/GeneralCollisionTensor.m:251-281 (+ src/SHL/gaunt_*.m, complex2real.m)
    sort_t_q1           benchmark_collision_contraction_scaling.m:82-85
    gauss_legendre      src/Gauss.m:38-105
    gauss_gen_laguerre  src/Gauss.m:128-173
    sumfac_grid_sizes   src/GeneralCollisionTensor.m:55-64
    build_rtensor       src/GeneralCollisionTensor.m:43-192 (host prep, GPU sumfac, conservation zeroing)
    set_sumfac_kernel   choice of the GPU quadrature kernel ("matrix_core" or "per_channel")
"""
import ctypes

import numpy as np

from ._lib import _i32, _i64, check, dptr, f64, lib


def real_gaunt(L):
    """(labels [N_G,3] float64 1-based, vals [N_G], ic_map [N_G] float64 1-based), MATLAB find order."""
    n = _i64(0)
    check(lib().cfk_real_gaunt(int(L), ctypes.byref(n), None, None, None))
    N = n.value
    lab = np.zeros(3 * N)
    val = np.zeros(N)
    ic = np.zeros(N)
    check(lib().cfk_real_gaunt(int(L), ctypes.byref(n), dptr(lab), dptr(val), dptr(ic)))
    return lab.reshape(3, N).T.copy(), val, ic


def sort_t_q1(labels, vals, ic_map):
    """Stable sortrows([ic_map, q1]) of the COO (what the reference benchmark feeds the fast kernels)."""
    labels = np.asarray(labels)
    order = np.lexsort((labels[:, 0], np.asarray(ic_map)))
    return labels[order], np.asarray(vals)[order], np.asarray(ic_map)[order]


def gauss_legendre(n, a=-1.0, b=1.0):
    x = np.zeros(int(n))
    w = np.zeros(int(n))
    check(lib().cfk_gauss_legendre(int(n), float(a), float(b), dptr(x), dptr(w)))
    return x, w


def gauss_gen_laguerre(n, alpha=0.0):
    x = np.zeros(int(n))
    w = np.zeros(int(n))
    check(lib().cfk_gauss_gen_laguerre(int(n), float(alpha), dptr(x), dptr(w)))
    return x, w


def sumfac_grid_sizes(K_max, L, radial_pad=40, angular_pad=0):
    s = (_i32 * 7)()
    check(lib().cfk_sumfac_grid_sizes(int(K_max), int(L), int(radial_pad), int(angular_pad), s))
    return tuple(int(v) for v in s)


SUMFAC_KERNELS = {"matrix_core": 0, "per_channel": 1, "exact_sphere": 2, "auto": 3}


def set_sumfac_kernel(kind):
    """Select the GPU quadrature kernel for later sumfac calls (process-wide).

    "auto" (default): "exact_sphere" where the (chi, eps) rule is exact for the gain integrand, else
    "matrix_core".  "exact_sphere": the gain sum of each unit step as the exact surface integral
    (SH at one direction per side and 1-D Gauss-Legendre integrals); raises where the rule is not
    exact.  "matrix_core": gain-point sums once per SH degree on FP64 matrix cores, then the P_gain
    slot contraction; "per_channel": one lane per t_chan in the reference's summation order
    (compute_rtensor_sumfac_mex.cpp:219-234).  All FP64; they agree to rounding."""
    check(lib().cfk_sumfac_set_kernel(SUMFAC_KERNELS[kind]))


def set_sumfac_chunk_nodes(max_nodes=0):
    """Cap the x nodes per sumfac chunk (process-wide; 0 = sized by the unit buffers' memory cap)."""
    check(lib().cfk_sumfac_set_chunk_nodes(int(max_nodes)))


def sumfac_last_chunks():
    """Number of x-node chunks the last sumfac call ran."""
    n = _i64(0)
    check(lib().cfk_sumfac_last_chunks(ctypes.byref(n)))
    return n.value


def sumfac_last_timing():
    """Device time of the last sumfac call: {"kernel", "units", "units_ms", "fold_ms"}."""
    k, n, a, b = _i32(0), _i64(0), ctypes.c_double(0.0), ctypes.c_double(0.0)
    check(lib().cfk_sumfac_last_timing(ctypes.byref(k), ctypes.byref(n), ctypes.byref(a), ctypes.byref(b)))
    names = {v: kk for kk, v in SUMFAC_KERNELS.items()}
    return {"kernel": names[k.value], "units": n.value, "units_ms": a.value, "fold_ms": b.value}


def build_rtensor(K_max, L, alpha, grid_sizes, geometry=None):
    """R [N_K^3 * N_L] (flat, MATLAB column-major) built on the GPU, conservation rows zeroed.

    ``grid_sizes`` = (N_x, N_u1, N_t1, N_y2, N_t2, N_chi, N_eps); ``geometry`` = real_gaunt(L) output.
    """
    lab, val, ic = geometry if geometry is not None else real_gaunt(L)
    n_g = np.asarray(val).size
    n_l = int(np.max(ic))
    nk = int(K_max) + 1
    R = np.zeros(nk ** 3 * n_l)
    gs = (_i32 * 7)(*[int(v) for v in grid_sizes])
    check(lib().cfk_build_rtensor(dptr(R), int(K_max), int(L), float(alpha), gs, dptr(f64(lab)), dptr(f64(val)),
                                  dptr(f64(ic)), n_g))
    return R
