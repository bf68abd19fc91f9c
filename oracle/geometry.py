"""ORACLE (test infrastructure only) -- numpy restatement of the reference's host-side preparation.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may import this
module, and only as a checker.  The product path never calls it.

Restated reference functions (all MATLAB, read as text, never run):

* ``gaunt_generate_labels``   -- src/SHL/gaunt_generate_labels.m:1-42
* ``gaunt_compute_values``    -- src/SHL/gaunt_compute_values.m:1-91 (log-factorial Racah 3j)
* ``complex2real``            -- src/SHL/complex2real.m:1-40
* ``real_gaunt_geometry``     -- src/GeneralCollisionTensor.m:251-281 (U (x) U (x) conj(U), |G|>1e-12,
                                 MATLAB ``find`` order, ``ic_map = unique(l-rows)``)
* ``gauss_legendre``          -- src/Gauss.m:38-105 (von Winckel Newton iteration)
* ``gauss_gen_laguerre``      -- src/Gauss.m:128-173 (Golub-Welsch seed + Newton + derivative weights)
* ``sumfac_grid_sizes``       -- src/GeneralCollisionTensor.m:54-64
* ``radial_norm``/``sh_norm`` -- src/GeneralCollisionTensor.m:81-104
* ``sumfac_prep``             -- src/GeneralCollisionTensor.m:106-174 (P_gain/P_loss tables)
* ``conservation_zero``       -- src/GeneralCollisionTensor.m:183-191
* ``assemble_dense``          -- src/GeneralCollisionTensor.m:283-311
* ``real_sh``                 -- src/SphericalHarmonics.m:15-50 (MATLAB ``legendre`` with Condon-Shortley phase)
* ``evaluate_basis``          -- src/SpectralBasis.m:25-58

Pins (tests/test_oracle_*.py): labels/ic_map bit-exact and values against the reference fixtures
for L in {2,4,6,8,10,12}; Gaunt values against brute-force quadrature of Y*Y*Y; quadrature
polynomial exactness; the sumfac prep + oracle sumfac against the fixture R (gamma=0, pad 20).
"""
import math

import numpy as np
from scipy.special import gammaln

# ----------------------------------------------------------------------------------------------
# Gaunt geometry
# ----------------------------------------------------------------------------------------------


def q_index(l, m):
    """1-based angular index q = l^2 + l + m + 1 (src/SphericalHarmonics.m:52-58)."""
    return l * l + l + m + 1


def q_to_l(q):
    """l of a 1-based q (``q2l`` in src/GeneralCollisionTensor.m:120)."""
    return np.floor(np.sqrt(np.asarray(q, dtype=np.float64) - 1.0)).astype(np.int64)


def gaunt_generate_labels(L):
    """Complex-Gaunt skeleton: triangle, parity and m3 = m1 + m2 (gaunt_generate_labels.m:1-33)."""
    out = []
    for l1 in range(L + 1):
        for l2 in range(L + 1):
            l3_min = abs(l1 - l2)
            l3_max = min(L, l1 + l2)
            start = l3_min if (l1 + l2 + l3_min) % 2 == 0 else l3_min + 1
            for l3 in range(start, l3_max + 1, 2):
                for m1 in range(-l1, l1 + 1):
                    q1 = q_index(l1, m1)
                    for m2 in range(-l2, l2 + 1):
                        m3 = m1 + m2
                        if abs(m3) <= l3:
                            out.append((q1, q_index(l2, m2), q_index(l3, m3)))
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def _log_factorials(n):
    # lf[n] = log(n!), built by a sequential cumulative sum of log(1..n) as in
    # gaunt_compute_values.m:3
    lf = np.zeros(n + 1)
    acc = 0.0
    for i in range(1, n + 1):
        acc += math.log(i)
        lf[i] = acc
    return lf


def _w3j_zero(l1, l2, l3, lf):
    # gaunt_compute_values.m:49-57
    J = l1 + l2 + l3
    if J % 2:
        return 0.0
    g = J // 2
    lw = 0.5 * (lf[2 * g - 2 * l1] + lf[2 * g - 2 * l2] + lf[2 * g - 2 * l3] - lf[2 * g + 1]) \
        + lf[g] - (lf[g - l1] + lf[g - l2] + lf[g - l3])
    return (-1.0) ** g * math.exp(lw)


def _w3j_magnetic(l1, l2, l3, m1, m2, m3, lf):
    # gaunt_compute_values.m:59-91 (Racah sum in log space, sequential accumulation)
    if abs(m1 + m2 + m3) > 1e-14:
        return 0.0
    log_delta = 0.5 * (lf[l1 + l2 - l3] + lf[l1 - l2 + l3] + lf[-l1 + l2 + l3] - lf[l1 + l2 + l3 + 1])
    t_min = max(0, l2 - l3 - m1, l1 - l3 + m2)
    t_max = min(l1 - m1, l2 + m2, l1 + l2 - l3)
    log_mag = 0.5 * (lf[l1 + m1] + lf[l1 - m1] + lf[l2 + m2] + lf[l2 - m2] + lf[l3 + m3] + lf[l3 - m3])
    s = 0.0
    for t in range(t_min, t_max + 1):
        term = lf[t] + lf[l1 + l2 - l3 - t] + lf[l1 - m1 - t] + lf[l2 + m2 - t] \
            + lf[l3 - l2 + m1 + t] + lf[l3 - l1 - m2 + t]
        s += (-1.0) ** t * math.exp(-term)
    phase = (-1.0) ** (l1 - l2 - m3)
    return phase * math.exp(log_delta + log_mag) * s


def gaunt_compute_values(L):
    """Complex Gaunt <Y1 Y2 Y3*> on the skeleton (gaunt_compute_values.m:1-47)."""
    lf = _log_factorials(3 * L + 2)
    labels = gaunt_generate_labels(L)
    ls = q_to_l(labels)
    ms = labels - 1 - ls * ls - ls
    vals = np.zeros(labels.shape[0])
    cache = {}
    for i in range(labels.shape[0]):
        l1, l2, l3 = (int(v) for v in ls[i])
        m1, m2, m3 = (int(v) for v in ms[i])
        key = (l1, l2, l3)
        if key not in cache:
            cache[key] = _w3j_zero(l1, l2, l3, lf)
        w0 = cache[key]
        if w0 == 0.0:
            continue
        pref = (-1.0) ** m3 * math.sqrt((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / (4 * math.pi)) * w0
        vals[i] = pref * _w3j_magnetic(l1, l2, l3, m1, m2, -m3, lf)
    return labels, vals


def complex2real(L):
    """Block-diagonal complex->real SH transform U (complex2real.m:1-40)."""
    n = (L + 1) ** 2
    U = np.zeros((n, n), dtype=np.complex128)
    inv = 1.0 / math.sqrt(2.0)
    off = 0
    for l in range(L + 1):
        d = 2 * l + 1
        if l == 0:
            U[off, off] = 1.0
            off += d
            continue
        m = np.arange(1, l + 1)
        diag = np.concatenate([1j * np.ones(l), [1.0], (-1.0) ** m]) * inv
        diag[l] = 1.0
        adiag = np.concatenate([-1j * ((-1.0) ** m[::-1]), [1.0], np.ones(l)]) * inv
        adiag[l] = 0.0
        blk = np.diag(diag) + np.fliplr(np.diag(adiag))
        U[off:off + d, off:off + d] = blk
        off += d
    return U


def real_gaunt_geometry(L):
    """Real Gaunt COO: labels [N_G,3] (1-based), vals [N_G], ic_map [N_G] (1-based).

    GeneralCollisionTensor.m:251-281.  Order is MATLAB ``find`` order on G_real[q1,q2,q3]
    (column-major: q1 fastest, then q2, then q3).
    """
    c_labels, c_vals = gaunt_compute_values(L)
    n = (L + 1) ** 2
    G = np.zeros((n, n, n), dtype=np.complex128)
    G[c_labels[:, 0] - 1, c_labels[:, 1] - 1, c_labels[:, 2] - 1] = c_vals
    U = complex2real(L)
    # G_real[a,b,c] = sum U[a,i] U[b,j] conj(U[c,k]) G[i,j,k]
    G = np.tensordot(U, G, axes=([1], [0]))                       # [a,j,k]
    G = np.tensordot(U, G, axes=([1], [1])).transpose(1, 0, 2)    # [a,b,k]
    G = np.tensordot(G, np.conj(U), axes=([2], [1]))              # [a,b,c]
    Gr = np.real(G)
    # find() in column-major order == lexicographic (q3, q2, q1)
    idx = np.nonzero(np.abs(Gr.transpose(2, 1, 0)) > 1e-12)
    q3, q2, q1 = idx
    labels = np.stack([q1, q2, q3], axis=1) + 1
    vals = Gr[q1, q2, q3]
    ls = q_to_l(labels)
    _, ic = np.unique(ls, axis=0, return_inverse=True)
    return labels.astype(np.int64), vals, ic.reshape(-1).astype(np.int64) + 1


def sort_t_q1(labels, vals, ic_map):
    """Stable ``sortrows([ic_map, q1])`` (benchmark_collision_contraction_scaling.m:82-85)."""
    order = np.lexsort((labels[:, 0], ic_map))
    return labels[order], vals[order], ic_map[order]


def l_triplets(labels, ic_map):
    """L_triplets[t] from the first COO entry of each t (GeneralCollisionTensor.m:116-122)."""
    n_l = int(ic_map.max())
    out = np.zeros((n_l, 3), dtype=np.int64)
    for t in range(1, n_l + 1):
        first = np.nonzero(ic_map == t)[0][0]
        out[t - 1] = q_to_l(labels[first])
    return out


# ----------------------------------------------------------------------------------------------
# 1-D quadrature (Gauss.m)
# ----------------------------------------------------------------------------------------------

def gauss_legendre(N, a=-1.0, b=1.0):
    """Gauss-Legendre nodes/weights on [a,b], ascending (Gauss.m:38-105)."""
    N = N - 1
    N1 = N + 1
    N2 = N + 2
    xu = np.linspace(-1.0, 1.0, N1)
    y = np.cos((2.0 * np.arange(N + 1) + 1.0) * np.pi / (2 * N + 2)) + (0.27 / N1) * np.sin(np.pi * xu * N / N2)
    Lm = np.zeros((N1, N2))
    y0 = 2.0
    Lp = None
    while np.max(np.abs(y - y0)) > np.finfo(float).eps:
        Lm[:, 0] = 1.0
        Lm[:, 1] = y
        for k in range(2, N1 + 1):
            Lm[:, k] = ((2 * k - 1) * y * Lm[:, k - 1] - (k - 1) * Lm[:, k - 2]) / k
        Lp = N2 * (Lm[:, N1 - 1] - y * Lm[:, N2 - 1]) / (1.0 - y ** 2)
        y0 = y
        y = y0 - Lm[:, N2 - 1] / Lp
    x = ((a * (1.0 - y) + b * (1.0 + y)) / 2.0)[::-1]
    w = ((b - a) / ((1.0 - y ** 2) * Lp ** 2) * (N2 / N1) ** 2)[::-1]
    return x.copy(), w.copy()


def gauss_gen_laguerre(N, alpha=0.0):
    """Generalized Gauss-Laguerre for weight x^alpha e^-x (Gauss.m:128-173)."""
    i = np.arange(1, N + 1, dtype=np.float64)
    a = 2.0 * i - 1.0 + alpha
    bb = np.sqrt(i[:-1] * (i[:-1] + alpha))
    J = np.diag(a) + np.diag(bb, 1) + np.diag(bb, -1)
    x = np.sort(np.linalg.eigvalsh(J))
    Lp1 = None
    for _ in range(15):
        L0 = np.ones(N)
        L1 = 1.0 + alpha - x
        Lp0 = np.zeros(N)
        Lp1 = -np.ones(N)
        for k in range(2, N + 1):
            L2 = ((2 * k - 1 + alpha - x) * L1 - (k - 1 + alpha) * L0) / k
            Lp2 = ((2 * k - 1 + alpha - x) * Lp1 - L1 - (k - 1 + alpha) * Lp0) / k
            L0, L1 = L1, L2
            Lp0, Lp1 = Lp1, Lp2
        dx = L1 / Lp1
        x = x - dx
        if np.max(np.abs(dx)) < 1e-15:
            break
    ln_const = gammaln(N + alpha + 1.0) - gammaln(N + 1.0)
    w = np.exp(ln_const) / (x * Lp1 ** 2)
    return x, w


# ----------------------------------------------------------------------------------------------
# Spherical harmonics and basis
# ----------------------------------------------------------------------------------------------

def assoc_legendre_all(l, x):
    """P_l^m(x), m=0..l, with Condon-Shortley phase (MATLAB ``legendre(l,x)`` convention)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros((l + 1,) + x.shape)
    s = np.sqrt(np.maximum((1.0 - x) * (1.0 + x), 0.0))
    for m in range(l + 1):
        pmm = np.ones_like(x)
        fact = 1.0
        for _ in range(m):
            pmm = -pmm * fact * s
            fact += 2.0
        if l == m:
            out[m] = pmm
            continue
        pm1 = x * (2 * m + 1) * pmm
        if l == m + 1:
            out[m] = pm1
            continue
        pll = pm1
        for ll in range(m + 2, l + 1):
            pll = (x * (2 * ll - 1) * pm1 - (ll + m - 1) * pmm) / (ll - m)
            pmm, pm1 = pm1, pll
        out[m] = pll
    return out


def real_sh(L, theta, phi):
    """Real SH Y[N, (L+1)^2] (SphericalHarmonics.m:15-50)."""
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64)).ravel()
    phi = np.atleast_1d(np.asarray(phi, dtype=np.float64)).ravel()
    x = np.cos(theta)
    Y = np.zeros((theta.size, (L + 1) ** 2))
    for l in range(L + 1):
        P = assoc_legendre_all(l, x)
        for m in range(-l, l + 1):
            am = abs(m)
            nf = math.sqrt(((2 * l + 1) / (4 * math.pi)) * math.exp(gammaln(l - am + 1) - gammaln(l + am + 1)))
            idx = q_index(l, m) - 1
            if m > 0:
                Y[:, idx] = math.sqrt(2.0) * nf * P[am] * np.cos(am * phi)
            elif m == 0:
                Y[:, idx] = nf * P[0]
            else:
                Y[:, idx] = math.sqrt(2.0) * nf * P[am] * np.sin(am * phi)
    return Y


def radial_norm(K, L):
    """RadialNorm[N_K, L+1] (GeneralCollisionTensor.m:82-90)."""
    out = np.zeros((K, L + 1))
    for k in range(K):
        for l in range(L + 1):
            ln_m = -math.log(2.0) + gammaln(k + l + 0.5 + 1.0) - gammaln(k + 1.0)
            out[k, l] = math.exp(-0.5 * ln_m)
    return out


def sh_norm(L):
    """SH_Norm[N_Q] (GeneralCollisionTensor.m:92-104)."""
    out = np.zeros((L + 1) ** 2)
    for l in range(L + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            base = math.sqrt(((2 * l + 1) / (4 * math.pi)) * math.exp(gammaln(l - am + 1) - gammaln(l + am + 1)))
            out[q_index(l, m) - 1] = base if m == 0 else math.sqrt(2.0) * base
    return out


def gen_laguerre_val(k, alpha, x):
    """L_k^(alpha)(x) by the three-term recursion (SpectralBasis.m:143-162)."""
    x = np.asarray(x, dtype=np.float64)
    if k == 0:
        return np.ones_like(x)
    if k == 1:
        return 1.0 + alpha - x
    p2 = np.ones_like(x)
    p1 = 1.0 + alpha - x
    v = p1
    for i in range(1, k):
        v = ((2 * i + 1 + alpha - x) * p1 - (i + alpha) * p2) / (i + 1)
        p2, p1 = p1, v
    return v


def evaluate_basis(K, L, v):
    """Psi[N_pts, K*(L+1)^2], column index k*N_Q + q (SpectralBasis.m:25-58)."""
    v = np.asarray(v, dtype=np.float64)
    nq = (L + 1) ** 2
    r2 = np.sum(v * v, axis=1)
    r = np.sqrt(r2) + np.finfo(float).eps
    theta = np.arccos(v[:, 2] / r)
    phi = np.arctan2(v[:, 1], v[:, 0])
    Y = real_sh(L, theta, phi)
    Psi = np.zeros((v.shape[0], K * nq))
    for l in range(L + 1):
        cols = np.arange(l * l, (l + 1) ** 2)
        solid = (r ** l)[:, None] * Y[:, cols]
        for k in range(K):
            lk = gen_laguerre_val(k, l + 0.5, r2)
            ln_m = -math.log(2.0) + gammaln(k + l + 1.5) - gammaln(k + 1.0)
            Psi[:, k * nq + cols] = math.exp(-0.5 * ln_m) * lk[:, None] * solid
    return Psi


# ----------------------------------------------------------------------------------------------
# Sum-factorized quadrature preparation
# ----------------------------------------------------------------------------------------------

def sumfac_grid_sizes(K_max, L, radial_pad, angular_pad):
    """(N_x, N_u1, N_t1, N_y2, N_t2, N_chi, N_eps) (GeneralCollisionTensor.m:55-64)."""
    N_x = radial_pad + int(math.ceil((3 * K_max + 1.5 * L + 3) / 2.0))
    N_u1 = radial_pad + 4 * K_max + 3 * L + 4
    N_t1 = angular_pad + K_max + int(math.ceil(1.5 * L)) + 1
    N_y2 = angular_pad + 4 * K_max + 3 * L + 4
    N_t2 = radial_pad + 3 * K_max + int(math.floor(1.5 * L)) + 3
    N_chi = K_max + int(math.ceil(0.5 * L)) + 1
    N_eps = 2 * K_max + L + 1
    return N_x, N_u1, N_t1, N_y2, N_t2, N_chi, N_eps


def sumfac_grids(alpha, sizes):
    """The 1-D rules handed to the sumfac (GeneralCollisionTensor.m:66-79)."""
    N_x, N_u1, N_t1, N_y2, N_t2, N_chi, N_eps = sizes
    g = {}
    g["x"], g["W_x"] = gauss_gen_laguerre(N_x, alpha / 2.0)
    g["u1"], g["W_u1"] = gauss_legendre(N_u1, 0.0, 1.0)
    g["t1"], g["W_t1"] = gauss_legendre(N_t1, 0.0, 1.0)
    g["y2"], g["W_y2"] = gauss_legendre(N_y2, 0.0, 1.0)
    g["t2"], g["W_t2"] = gauss_legendre(N_t2, 0.0, 1.0)
    g["mu_chi"], g["W_chi"] = gauss_legendre(N_chi, -1.0, 1.0)
    g["eps"] = 2.0 * np.pi * np.arange(N_eps) / N_eps
    g["W_eps"] = (2.0 * np.pi / N_eps) * np.ones(N_eps)
    return g


def sumfac_prep(K_max, L, alpha, labels, vals, ic_map, grids):
    """All 26 sumfac inputs as a dict (GeneralCollisionTensor.m:50-174).

    ``labels``/``vals``/``ic_map`` are the unsorted (MATLAB find order) COO, 1-based.
    Arrays are returned in MATLAB column-major memory order (flat, Fortran-ordered ndarray).
    """
    nk = K_max + 1
    nq = (L + 1) ** 2
    n_l = int(ic_map.max())
    u1, t1, y2 = grids["u1"], grids["t1"], grids["y2"]
    N_u1, N_t1, N_y2 = u1.size, t1.size, y2.size
    trip = np.zeros((n_l, 3))
    qi_valid = -np.ones((nq, n_l))
    P_loss_p1 = np.zeros((N_t1, N_u1, n_l), order="F")
    P_gain_p1 = np.zeros((N_t1, N_u1, nq, n_l), order="F")
    P_loss_p2 = np.zeros((N_y2, n_l), order="F")
    P_gain_p2 = np.zeros((N_y2, nq, n_l), order="F")
    # SH evaluated at phi=0 on the patch-1 (t,u) grid and the patch-2 y grid
    yy1 = u1[None, :] * t1[:, None]                         # [N_t1, N_u1]
    mu1 = np.clip(1.0 - 2.0 * yy1 ** 2, -1.0, 1.0)
    Y1 = real_sh(L, np.arccos(mu1).ravel(order="F"), np.zeros(mu1.size)).reshape(N_t1, N_u1, nq, order="F")
    mu2 = np.clip(1.0 - 2.0 * y2 ** 2, -1.0, 1.0)
    Y2 = real_sh(L, np.arccos(mu2), np.zeros(N_y2))          # [N_y2, nq]
    for t in range(1, n_l + 1):
        g_idx = np.nonzero(ic_map == t)[0]
        l_i, l_j, l_k = (int(v) for v in q_to_l(labels[g_idx[0]]))
        trip[t - 1] = (l_i, l_j, l_k)
        q_i0 = l_i * l_i + l_i + 1
        q_j0 = l_j * l_j + l_j + 1
        Y_i0 = math.sqrt((2 * l_i + 1) / (4 * math.pi))
        Y_j0 = math.sqrt((2 * l_j + 1) / (4 * math.pi))
        D_full = 0.0
        for g in g_idx:
            D_full += vals[g] ** 2
        SCALE = (8 * math.pi ** 2 * Y_j0) / D_full
        qi_list = []
        for g in g_idx:
            q_i, q_j, q_k = (int(v) for v in labels[g])
            if q_j != q_j0:
                continue
            G = vals[g]
            if q_i not in qi_list:
                qi_list.append(q_i)
            P_gain_p1[:, :, q_i - 1, t - 1] += G * Y1[:, :, q_k - 1]
            P_gain_p2[:, q_i - 1, t - 1] += G * Y2[:, q_k - 1]
            if q_i == q_i0:
                P_loss_p1[:, :, t - 1] += G * Y_i0 * Y1[:, :, q_k - 1]
                P_loss_p2[:, t - 1] += G * Y_i0 * Y2[:, q_k - 1]
        P_gain_p1[:, :, :, t - 1] *= SCALE
        P_loss_p1[:, :, t - 1] *= SCALE
        P_gain_p2[:, :, t - 1] *= SCALE
        P_loss_p2[:, t - 1] *= SCALE
        qi_valid[:len(qi_list), t - 1] = qi_list
    return dict(K_max=K_max, N_L=n_l, N_Q=nq, alpha=float(alpha), **grids,
                RadialNorm=np.asfortranarray(radial_norm(nk, L)), SH_Norm=sh_norm(L),
                L_triplets=np.asfortranarray(trip), qi_valid_mat=np.asfortranarray(qi_valid),
                P_loss_p1=P_loss_p1, P_gain_p1=P_gain_p1, P_loss_p2=P_loss_p2, P_gain_p2=P_gain_p2)


def conservation_zero(R, trip, nk):
    """Zero the conserved rows of R[k1,k2,k3,t] (GeneralCollisionTensor.m:183-191).

    ``R`` is flat in MATLAB column-major order; returns a zeroed copy.
    """
    n_l = trip.shape[0]
    Rv = np.array(R, dtype=np.float64).reshape(n_l, nk, nk, nk).copy()   # [t,k3,k2,k1]
    for t in range(n_l):
        if trip[t, 0] == 0:
            Rv[t, :, :, 0] = 0.0
            Rv[t, :, :, 1] = 0.0
        elif trip[t, 0] == 1:
            Rv[t, :, :, 0] = 0.0
    return Rv.reshape(-1)


def assemble_dense(labels, vals, ic_map, R, nq, nk):
    """Dense C[i + j*N + k*N^2], N = nq*nk, index k*nq + q (GeneralCollisionTensor.m:283-311)."""
    N = nq * nk
    n_l = int(ic_map.max())
    Rv = np.asarray(R).reshape(n_l, nk, nk, nk)   # [t,k3,k2,k1]
    C = np.zeros((N, N, N))                        # C[k,j,i] so that C.ravel() is column-major
    ks = np.arange(nk)
    for z in range(labels.shape[0]):
        q1, q2, q3 = (int(v) - 1 for v in labels[z])
        t = int(ic_map[z]) - 1
        i1 = ks * nq + q1
        i2 = ks * nq + q2
        i3 = ks * nq + q3
        C[np.ix_(i3, i2, i1)] = Rv[t] * vals[z]
    return C.reshape(-1)
