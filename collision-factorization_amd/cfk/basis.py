"""Basis evaluation at arbitrary velocities and moment diagnostics on the GPU (libcfk).

``basis_evaluate`` mirrors ``Basis.evaluate(v)`` (src/SpectralBasis.m:25-58); ``moments`` reads the
density, bulk velocity, temperature and pressure tensor out of coefficient vectors, the quantities
the reference's tutorials follow (tutorial_collisions_spectral_properties.m:138-144,
tutorial_stress_relaxation.m:160-168).
"""
import ctypes

import numpy as np

from ._lib import LAYOUT, check, lib
from .plan import _check_dev_f64, _ptr, _stream_ptr

N_MOMENTS = 11
MOMENT_NAMES = ("n", "u_x", "u_y", "u_z", "T", "P_xx", "P_yy", "P_zz", "P_xy", "P_xz", "P_yz")


def basis_evaluate(v, K, L, stream=None):
    """Psi [n_pts, K (L+1)^2] (column k*N_Q + q) at velocities v [n_pts, 3], both float64 on the GPU."""
    import torch
    if not isinstance(v, torch.Tensor) or v.dim() != 2 or v.shape[1] != 3:
        raise ValueError("v must be a [n_pts, 3] torch tensor")
    n = int(v.shape[0])
    vc = v.t().contiguous()   # MATLAB's n x 3 column-major = [3][n]
    _check_dev_f64(vc, 3 * n, "v")
    nt = int(K) * (int(L) + 1) ** 2
    out = torch.empty(nt, n, dtype=torch.float64, device=v.device)   # column-major n x N_terms
    check(lib().cfk_basis_evaluate(_ptr(out), _ptr(vc), n, int(K), int(L), _stream_ptr(stream)))
    return out.t()


def moments(c, B, N_Q, K, layout="batch_fastest", stream=None):
    """[B, 11] moments (n, u_x, u_y, u_z, T, P_xx, P_yy, P_zz, P_xy, P_xz, P_yz) of B coefficient vectors."""
    import torch
    _check_dev_f64(c, int(B) * int(N_Q) * int(K), "c")
    out = torch.empty(int(B), N_MOMENTS, dtype=torch.float64, device=c.device)
    check(lib().cfk_moments(_ptr(out), _ptr(c), int(B), int(N_Q), int(K), LAYOUT[layout], _stream_ptr(stream)))
    return out


def moment_weights(N_Q, K):
    """The ten linear functionals behind ``moments`` as a dense [10, N_Q*K] matrix (host; index k*N_Q + q):
    int f, int v_x f, int v_y f, int v_z f, int v_x^2 f, v_y^2, v_z^2, v_x v_y, v_x v_z, v_y v_z."""
    idx = (ctypes.c_int32 * 40)()
    w = (ctypes.c_double * 40)()
    check(lib().cfk_moment_weights(int(N_Q), int(K), idx, w))
    M = np.zeros((10, int(N_Q) * int(K)))
    for j in range(10):
        for t in range(4):
            if idx[4 * j + t] >= 0:
                M[j, idx[4 * j + t]] += w[4 * j + t]
    return M
