"""Host mirror of the reference's MEX interface, backed by libcfk (HIP, gfx950).

Same names, argument order and meaning, and the same in-place accumulation as the reference MEX
files; MATLAB arrays map to numpy float64 arrays in column-major (Fortran) order, and 1-based
index arrays are passed unchanged (as doubles, as MATLAB passes them).

    angular_first_collision_kernel_mex  src/mex/angular_first_collision_kernel_mex.cpp:5-16
    radial_first_collision_kernel_mex   src/mex/radial_first_collision_kernel_mex.cpp:4-17
    naive_collision_kernel_mex          src/mex/naive_collision_kernel_mex.cpp:3-16
    dense_tensor_kernel_mex             src/mex/dense_tensor_kernel_mex.cpp:3-11
    compute_rtensor_sumfac_mex          src/mex/compute_rtensor_sumfac_mex.cpp:80-118

Unlike the reference, index arrays are validated (1..N_Q, 1..N_L) and a bad call raises
``CfkError`` instead of reading out of bounds.
"""
import numpy as np

from ._lib import CfkError, check, dptr, f64, lib


class MexError(CfkError):
    """Mirror of ``mexErrMsgIdAndTxt(identifier, message)``."""

    def __init__(self, identifier, message):
        RuntimeError.__init__(self, "%s: %s" % (identifier, message))
        self.identifier = identifier
        self.code = -1


def _inplace_view(Q):
    """Flat column-major float64 view of Q (must be writable and share memory)."""
    if not isinstance(Q, np.ndarray) or Q.dtype != np.float64:
        raise TypeError("Q_out must be a float64 numpy array")
    v = Q.reshape(-1, order="F") if Q.flags.f_contiguous else None
    if v is None or not np.shares_memory(v, Q):
        raise ValueError("Q_out must be Fortran-contiguous (MATLAB layout) to be updated in place")
    return v


def _sparse(fn_name, Q_out, f, labels, vals, ic_map, R_tensor, N_Q, K_len):
    labels = np.asarray(labels)
    if labels.ndim != 2 or labels.shape[1] != 3:
        raise ValueError("labels must be N_G x 3")
    n_g = labels.shape[0]
    q = _inplace_view(Q_out)
    N_Q, K_len = int(N_Q), int(K_len)
    if q.size != N_Q * K_len:
        raise ValueError("Q_out has %d elements, expected N_Q*K = %d" % (q.size, N_Q * K_len))
    ff = f64(f)
    if ff.size != N_Q * K_len:
        raise ValueError("f has %d elements, expected N_Q*K = %d" % (ff.size, N_Q * K_len))
    lab, v, ic = f64(labels), f64(vals), f64(ic_map)
    if v.size != n_g or ic.size != n_g:
        raise ValueError("vals and ic_map must have N_G = %d entries" % n_g)
    R = f64(R_tensor)
    out = np.ascontiguousarray(q)  # contiguous work copy (q itself is contiguous, so no copy)
    check(getattr(lib(), fn_name)(dptr(out), dptr(ff), dptr(lab), dptr(v), dptr(ic), dptr(R), n_g, R.size, N_Q,
                                  K_len))
    if out is not q:
        q[...] = out
    return Q_out


def angular_first_collision_kernel_mex(Q_out, f, labels_sorted, vals_sorted, ic_map_sorted, R_tensor, N_Q, K_len):
    """Q_out += Q(f), angular-first (q1-sliced) contraction; COO sorted by (t, q1)."""
    return _sparse("cfk_angular_first_collision_kernel", Q_out, f, labels_sorted, vals_sorted, ic_map_sorted,
                   R_tensor, N_Q, K_len)


def radial_first_collision_kernel_mex(Q_out, f, labels_sorted, vals_sorted, ic_map_sorted, R_tensor, N_Q, K_len):
    """Q_out += Q(f), radial-first contraction; COO sorted by t."""
    return _sparse("cfk_radial_first_collision_kernel", Q_out, f, labels_sorted, vals_sorted, ic_map_sorted,
                   R_tensor, N_Q, K_len)


def naive_collision_kernel_mex(Q_out, f, gaunt_labels, gaunt_vals, ic_map, R_tensor, N_Q, K_len):
    """Q_out += Q(f), naive sparse contraction (any COO order)."""
    return _sparse("cfk_naive_collision_kernel", Q_out, f, gaunt_labels, gaunt_vals, ic_map, R_tensor, N_Q, K_len)


def dense_tensor_kernel_mex(Q_out, f, C_dense, N_terms):
    """Q_out[i] += sum_{j,k} C[i + j N + k N^2] f[j] f[k]."""
    q = _inplace_view(Q_out)
    N = int(N_terms)
    ff = f64(f)
    C = f64(C_dense)
    if q.size != N or ff.size != N or C.size != N ** 3:
        raise ValueError("dense_tensor_kernel_mex: sizes do not match N_terms=%d" % N)
    check(lib().cfk_dense_tensor_kernel(dptr(q), dptr(ff), dptr(C), N))
    return Q_out


_SUMFAC_ARGS = ["K_max", "N_L", "N_Q", "alpha", "x_nodes", "W_x", "u1_nodes", "W_u1", "t1_nodes", "W_t1",
                "y2_nodes", "W_y2", "t2_nodes", "W_t2", "mu_chi", "W_chi", "eps_vec", "W_eps", "RadialNorm",
                "SH_Norm", "L_triplets", "qi_valid_mat", "P_loss_p1", "P_gain_p1", "P_loss_p2", "P_gain_p2"]


def compute_rtensor_sumfac_mex(*prhs, x_range=None):
    """R = compute_rtensor_sumfac_mex(K_max, N_L, N_Q, alpha, x, W_x, ..., P_gain_p2)  (26 inputs).

    Returns R as a Fortran-ordered [N_K, N_K, N_K, N_L] float64 array (no conservation zeroing,
    as the MEX).  ``x_range=(a, b)`` restricts the outer sum to x nodes [a, b) (a partial sum;
    the full tensor is the sum over all nodes).
    """
    if len(prhs) != 26:
        raise MexError("R_tensor:InvalidInput", "Exactly 26 inputs required.")
    K_max, N_L, N_Q = int(prhs[0]), int(prhs[1]), int(prhs[2])
    alpha = float(prhs[3])
    arrs = [f64(a) for a in prhs[4:]]
    nk = K_max + 1
    R = np.zeros(nk ** 3 * N_L)
    grid = []
    for i in range(7):
        grid += [dptr(arrs[2 * i]), dptr(arrs[2 * i + 1]), arrs[2 * i].size]
        if arrs[2 * i].size != arrs[2 * i + 1].size:
            raise MexError("R_tensor:InvalidInput", "%s and %s differ in length" % (_SUMFAC_ARGS[4 + 2 * i],
                                                                                     _SUMFAC_ARGS[5 + 2 * i]))
    n_x = arrs[0].size
    a, b = (0, n_x) if x_range is None else x_range
    tabs = [dptr(t) for t in arrs[14:]]
    check(lib().cfk_compute_rtensor_sumfac(dptr(R), K_max, N_L, N_Q, alpha, *grid, *tabs, int(a), int(b)))
    return R.reshape((nk, nk, nk, N_L), order="F")
