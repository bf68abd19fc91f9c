"""MATLAB v7.3 .mat (HDF5) reading and writing through libcfk's native code (cfk_h5_read_f64,
cfk_mat_write_f64).

    read_mat_dataset  one float64 dataset, in MATLAB's column-major order
    load_precalc      the reference's precalc TensorObj files (src/precalc, written by
                      precompute_collision_operator.m:61): R, Gaunt COO and metadata
    save_mat          float64 arrays as the variables of a v7.3 .mat file MATLAB's load reads
    save_operator     a collision operator (R, Gaunt COO, K_max, L_max, alpha) as such a file
    load_operator     the same variables read back

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


def save_mat(path, variables):
    """Write {name: float64 array} as a MATLAB v7.3 .mat file (variables in MATLAB dimensions: a numpy
    array of shape (m, n, ...) becomes an m x n x ... MATLAB array; scalars become 1 x 1)."""
    names, ranks, dims, keep = [], [], [], []
    for name, a in variables.items():
        a = np.asarray(a, dtype=np.float64)
        shape = a.shape if a.ndim >= 1 else (1, 1)
        keep.append(np.asfortranarray(a.reshape(shape)).ravel(order="F"))
        names.append(str(name).encode())
        ranks.append(len(shape))
        dims.extend(int(d) for d in shape)
    n = len(names)
    c_names = (ctypes.c_char_p * n)(*names)
    c_ranks = (_i32 * n)(*ranks)
    c_dims = (_i64 * len(dims))(*dims)
    c_data = (ctypes.POINTER(ctypes.c_double) * n)(*[dptr(k) for k in keep])
    check(lib().cfk_mat_write_f64(str(path).encode(), n, c_names, c_ranks, c_dims, c_data))


OPERATOR_VARIABLES = ("R_tensor", "gaunt_labels", "gaunt_vals", "ic_map", "K_max", "L_max", "alpha")


def save_operator(path, R, labels, vals, ic_map, K_max, L_max, alpha):
    """The collision operator's arrays -- R_tensor [N_K,N_K,N_K,N_L], gaunt_labels [N_G,3] (1-based),
    gaunt_vals, ic_map, K_max, L_max, alpha -- as variables of a MATLAB v7.3 file (the TensorObj
    properties of src/GeneralCollisionTensor.m:9-12 that precompute_collision_operator.m:61 saves)."""
    nk = int(K_max) + 1
    Rf = np.asarray(R, dtype=np.float64).ravel(order="F")
    n_l = Rf.size // nk ** 3
    save_mat(path, {"R_tensor": Rf.reshape((nk, nk, nk, n_l), order="F"),
                     "gaunt_labels": np.asarray(labels, dtype=np.float64),
                     "gaunt_vals": np.asarray(vals, dtype=np.float64).reshape(-1, 1),
                     "ic_map": np.asarray(ic_map, dtype=np.float64).reshape(-1, 1),
                     "K_max": float(K_max), "L_max": float(L_max), "alpha": float(alpha)})


def load_operator(path):
    """The variables of save_operator: dict(R flat column-major, labels [N_G,3], vals, ic_map, K_max,
    L_max, alpha)."""
    R, _ = read_mat_dataset(path, "R_tensor")
    lab, ld = read_mat_dataset(path, "gaunt_labels")
    vals, _ = read_mat_dataset(path, "gaunt_vals")
    ic, _ = read_mat_dataset(path, "ic_map")
    sc = {k: float(read_mat_dataset(path, k)[0][0]) for k in ("K_max", "L_max", "alpha")}
    return dict(R=R, labels=lab.reshape(ld, order="F"), vals=vals, ic_map=ic, K_max=int(sc["K_max"]),
                L_max=int(sc["L_max"]), alpha=sc["alpha"])
