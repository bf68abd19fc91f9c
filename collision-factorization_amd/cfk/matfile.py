"""MATLAB v7.3 .mat (HDF5) reading through libcfk's native reader (cfk_h5_read_f64).

    read_mat_dataset  one float64 dataset, in MATLAB's column-major order
    load_precalc      the reference's precalc TensorObj files (src/precalc, written by
                      precompute_collision_operator.m:61): R, Gaunt COO and metadata

A TensorObj saved with -v7.3 keeps its properties as MCOS datasets under ``/#refs#``; in the
reference's precalc files they are t = R_tensor [N_K,N_K,N_K,N_L], u = gaunt_labels [N_G,3],
v = gaunt_vals, w = ic_map, and the scalars q, r, s, j = K_max, L_max, D_max, alpha.  load_precalc
checks those against each other (R's size against ic_map, the labels against L_max, the integrality
of the index arrays) before returning them.
"""
import ctypes

import numpy as np

from ._lib import _i32, _i64, CfkError, CFK_EINVAL, check, dptr, lib


def read_mat_dataset(path, name):
    """(data, dims): a float64 dataset of an HDF5/.mat v7.3 file; ``dims`` in MATLAB order (the HDF5
    dimensions reversed) and ``data`` flat in MATLAB's column-major order."""
    n, r, dims = _i64(0), _i32(0), (_i64 * 8)()
    p, nm = str(path).encode(), str(name).encode()
    check(lib().cfk_h5_read_f64(p, nm, None, 0, ctypes.byref(n), ctypes.byref(r), dims))
    out = np.zeros(n.value)
    check(lib().cfk_h5_read_f64(p, nm, dptr(out), n.value, ctypes.byref(n), ctypes.byref(r), dims))
    return out, tuple(int(d) for d in reversed(dims[:r.value]))


def load_precalc(path):
    """The reference's precalc TensorObj: dict(R flat column-major [N_K^3 N_L], labels [N_G,3] and
    ic_map [N_G] as 1-based float64, vals [N_G], K_max, L_max, D_max, alpha)."""
    R, rd = read_mat_dataset(path, "/#refs#/t")
    lab, ld = read_mat_dataset(path, "/#refs#/u")
    vals, _ = read_mat_dataset(path, "/#refs#/v")
    ic, _ = read_mat_dataset(path, "/#refs#/w")
    K_max, L_max, D_max, alpha = (float(read_mat_dataset(path, "/#refs#/" + c)[0][0]) for c in "qrsj")
    n_g = vals.size
    if ld != (n_g, 3) or ic.size != n_g:
        raise CfkError(CFK_EINVAL, "%s: gaunt_labels %s do not match %d values" % (path, ld, n_g))
    labels = lab.reshape(3, n_g).T.copy()  # column-major [N_G x 3]
    nk = int(K_max) + 1
    nq = (int(L_max) + 1) ** 2
    if not (np.all(labels == np.round(labels)) and np.all(ic == np.round(ic))):
        raise CfkError(CFK_EINVAL, "%s: non-integral index arrays" % path)
    if labels.min() < 1 or labels.max() > nq:
        raise CfkError(CFK_EINVAL, "%s: labels outside 1..%d for L_max=%d" % (path, nq, int(L_max)))
    if len(rd) != 4 or rd[:3] != (nk,) * 3 or int(ic.max()) != rd[3]:
        raise CfkError(CFK_EINVAL, "%s: R_tensor dims %s vs K_max=%d, max(ic_map)=%d" % (path, rd, int(K_max),
                                                                                          int(ic.max())))
    return dict(R=R, labels=labels, vals=vals, ic_map=ic, K_max=int(K_max), L_max=int(L_max),
                D_max=D_max, alpha=alpha)
