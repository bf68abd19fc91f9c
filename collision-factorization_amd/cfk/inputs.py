"""Synthetic distributions used by the benchmark and the tests.

Perturbed Maxwellians as specified for the hot-path workloads: c_000 = 1 and every other
coefficient c_{k,l,m} ~ N(0, (0.1 * 0.5**(k+l))**2), drawn from numpy's PCG64 generator with a
fixed seed in MATLAB column-major (q fastest, then k) order, one distribution after another.
"""
import numpy as np


def l_of_q(N_Q):
    """Degree l of each 0-based angular index q (q = l^2 + l + m)."""
    return np.floor(np.sqrt(np.arange(N_Q))).astype(np.int64)


def perturbed_maxwellians(L, K, B, seed=0):
    """float64 array [B, K, N_Q]: B distributions, each an N_Q x K column-major MATLAB array."""
    nq = (L + 1) ** 2
    rng = np.random.Generator(np.random.PCG64(seed))
    scale = 0.1 * 0.5 ** (np.arange(K)[:, None] + l_of_q(nq)[None, :])   # [K, N_Q]
    f = rng.standard_normal((B, K, nq)) * scale[None]
    f[:, 0, 0] = 1.0
    return f
