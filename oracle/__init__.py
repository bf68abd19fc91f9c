"""ORACLE -- CPU restatement of the reference's hot path, used ONLY as a checker.

Allowed importers: ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py``.  The product library never links, loads or calls anything here.

* ``oracle.geometry``  -- numpy restatement of the MATLAB host preparation (Gaunt COO,
  quadrature rules, sumfac P tables, conservation zeroing, dense assembly, basis evaluation).
* ``liboracle_cfk.so`` -- plain-C restatement of the five reference MEX kernels
  (built from ``cfk_oracle.c`` by ``make -C oracle``), wrapped below with numpy in/out.

Parity status per function (what pins it) is listed in the test files
``tests/test_oracle_*.py``; every function here is pinned by at least one check that does not
retype its own formula.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle_cfk.so in place (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle_cfk.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_cfk.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        i64, i32 = ctypes.c_int64, ctypes.c_int
        for name in ("oracle_naive", "oracle_angular_first", "oracle_radial_first"):
            fn = getattr(L, name)
            fn.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, i64, i32, i32]
            fn.restype = None
        L.oracle_dense.argtypes = [_dp, _dp, _dp, i64]
        L.oracle_dense.restype = None
        L.oracle_angular_first_batch.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, i64, i32, i32, i64, i32]
        L.oracle_angular_first_batch.restype = None
        L.oracle_sumfac.argtypes = ([_dp, i32, i32, i32, ctypes.c_double] + [_dp, _dp, i32] * 7
                                    + [_dp] * 8 + [i32, i32, i32])
        L.oracle_sumfac.restype = None
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


def contract(variant, Q, f, labels, vals, ic_map, R, N_Q, K):
    """Q (float64, N_Q*K, column-major) += Q(f) through the named reference kernel.

    ``labels`` [N_G,3] 1-based, ``vals`` [N_G], ``ic_map`` [N_G] 1-based, ``R`` flat column-major.
    ``variant`` in {"naive", "angular_first", "radial_first"}.
    """
    assert Q.dtype == np.float64 and Q.flags.c_contiguous and Q.size == N_Q * K
    labels = np.asarray(labels)
    n_g = labels.shape[0]
    lab = _f64(labels)   # column-major [N_G x 3]
    fn = getattr(lib(), "oracle_" + variant)
    f = _f64(f)
    v, ic, r = _f64(vals), _f64(ic_map), _f64(R)
    fn(_p(Q), _p(f), _p(lab), _p(v), _p(ic), _p(r), n_g, int(N_Q), int(K))
    return Q


def angular_first_batch(Q, f, labels, vals, ic_map, R, N_Q, K, nthreads=1):
    """Batched angular-first over B distributions laid out [B, K, N_Q] (each column-major)."""
    labels = np.asarray(labels)
    n_g = labels.shape[0]
    B = Q.size // (N_Q * K)
    lab = _f64(labels)
    f = np.ascontiguousarray(f, dtype=np.float64)
    lib().oracle_angular_first_batch(_p(Q), _p(f), _p(lab), _p(_f64(vals)), _p(_f64(ic_map)), _p(_f64(R)),
                                     n_g, int(N_Q), int(K), int(B), int(nthreads))
    return Q


def dense(Q, f, C, N):
    f = _f64(f)
    C = np.ascontiguousarray(C, dtype=np.float64)
    lib().oracle_dense(_p(Q), _p(f), _p(C), int(N))
    return Q


SUMFAC_ORDER = ["K_max", "N_L", "N_Q", "alpha", "x", "W_x", "u1", "W_u1", "t1", "W_t1", "y2", "W_y2",
                "t2", "W_t2", "mu_chi", "W_chi", "eps", "W_eps", "RadialNorm", "SH_Norm", "L_triplets",
                "qi_valid_mat", "P_loss_p1", "P_gain_p1", "P_loss_p2", "P_gain_p2"]


def sumfac(args, x_begin=0, x_end=None, nthreads=None):
    """R [N_K^3 * N_L] (column-major) from the 26 reference inputs (dict), x nodes [x_begin, x_end)."""
    nk = int(args["K_max"]) + 1
    n_l = int(args["N_L"])
    R = np.zeros(nk ** 3 * n_l)
    arrs = {k: _f64(args[k]) for k in SUMFAC_ORDER[4:]}
    n_x = arrs["x"].size
    if x_end is None:
        x_end = n_x
    if nthreads is None:
        nthreads = min(8, os.cpu_count() or 1)
    grid = []
    for a, w in (("x", "W_x"), ("u1", "W_u1"), ("t1", "W_t1"), ("y2", "W_y2"), ("t2", "W_t2"),
                 ("mu_chi", "W_chi"), ("eps", "W_eps")):
        grid += [_p(arrs[a]), _p(arrs[w]), arrs[a].size]
    tabs = [_p(arrs[k]) for k in SUMFAC_ORDER[18:]]
    lib().oracle_sumfac(_p(R), int(args["K_max"]), n_l, int(args["N_Q"]), float(args["alpha"]),
                        *grid, *tabs, int(x_begin), int(x_end), int(nthreads))
    return R
