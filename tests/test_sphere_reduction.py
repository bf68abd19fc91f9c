"""The reduction behind the exact-sphere sumfac kernel, checked on the host against brute force.

For one unit step and side the reference sums the gain integrand over the product rule
(mu_chi, W_chi) x (eps, W_eps) at v' = c + rho R omega (compute_rtensor_sumfac_mex.cpp:186-234).
On the sphere |v'|^2 = |c|^2 + rho^2 + 2 rho c . R omega is linear in omega, so the integrand agrees
there with a polynomial of degree l + k; with a rule exact for that degree the sum equals
    2 pi S_q(c_hat) int_{-1}^{1} L_k^{(l+1/2)}(s(t)) Z_l(t) dt,
Z_l the zonal solid harmonic about c, which the kernel integrates by Gauss-Legendre with
ceil((L + N_K) / 2) nodes.  Here both sides are evaluated independently (scipy's Laguerre and
associated Legendre functions, the rotation and the rule of the reference's loops), and one eps node
fewer than the exact rule must break the equality."""
import numpy as np
from scipy.special import eval_genlaguerre, lpmv


def s_raw(l, m, ct, cp, sp):
    """Unnormalized real SH: P_l^|m|(cos theta) cos(m phi) for m >= 0, sin(|m| phi) for m < 0."""
    phi = np.arctan2(sp, cp)
    P = lpmv(abs(m), l, ct)
    return P * np.cos(m * phi) if m >= 0 else P * np.sin(-m * phi)


def quadrature(l, k, m, c, rho, zx, zz, n_chi, n_eps):
    """sum_p W L_k(|v'|^2) |v'|^l Y_raw(v') over the reference's product rule."""
    mu, wc = np.polynomial.legendre.leggauss(n_chi)
    eps = 2 * np.pi * np.arange(n_eps) / n_eps
    we = np.full(n_eps, 2 * np.pi / n_eps)
    M, E = np.meshgrid(mu, eps, indexing="ij")
    W = np.outer(wc, we)
    s = np.sqrt(1 - M * M)
    upx, upy, upz = s * np.cos(E), s * np.sin(E), M
    vx = c[0] + rho * (upx * zz + upz * zx)
    vy = rho * upy
    vz = c[1] + rho * (-upx * zx + upz * zz)
    r = np.sqrt(vx * vx + vy * vy + vz * vz)
    terms = W * eval_genlaguerre(k, l + 0.5, r * r) * r ** l * s_raw(l, m, vz / r, vx / np.hypot(vx, vy),
                                                                    vy / np.hypot(vx, vy))
    return terms.sum(), np.abs(terms).sum()


def closed_form(l, k, m, c, rho, n_nodes):
    t, w = np.polynomial.legendre.leggauss(n_nodes)
    cn = np.hypot(c[0], c[1])
    z = cn + rho * t
    s2 = z * z + rho * rho * (1 - t * t)
    Z = [np.ones_like(t), z]
    for j in range(1, l):
        Z.append(((2 * j + 1) * z * Z[j] - j * s2 * Z[j - 1]) / (j + 1))
    F = np.sum(w * eval_genlaguerre(k, l + 0.5, s2) * Z[l])
    return 2 * np.pi * s_raw(l, m, c[1] / cn, np.sign(c[0]) or 1.0, 0.0) * F


def test_sphere_sum_equals_the_one_dimensional_integral():
    rng = np.random.default_rng(3)
    L, NK = 6, 6
    d = L + NK - 1
    n_chi, n_eps = (d + 2) // 2, d + 1            # smallest rule exact for degree d
    n_nodes = (L + NK + 1) // 2
    worst = 0.0
    for _ in range(3):
        c = (abs(rng.normal()) * 1.2, rng.normal() * 1.2)    # U lies in the x-z plane with U_x >= 0
        rho = abs(rng.normal()) * 1.2
        ang = rng.uniform(0, 2 * np.pi)
        zx, zz = np.sin(ang), np.cos(ang)
        for l in range(L + 1):
            for k in range(NK):
                for m in range(-l, l + 1):
                    q, a = quadrature(l, k, m, c, rho, zx, zz, n_chi, n_eps)
                    worst = max(worst, abs(q - closed_form(l, k, m, c, rho, n_nodes)) / a)
    assert worst <= 1e-13


def test_one_eps_node_short_breaks_the_equality():
    rng = np.random.default_rng(4)
    L, NK = 6, 6
    d = L + NK - 1
    n_chi, n_eps = (d + 2) // 2, d + 1
    c, rho, zx, zz = (0.9, 0.4), 1.1, 0.6, 0.8
    l, k, m = L, NK - 1, L          # the integrand of largest degree
    q_ok, a = quadrature(l, k, m, c, rho, zx, zz, n_chi, n_eps)
    q_bad, _ = quadrature(l, k, m, c, rho, zx, zz, n_chi, n_eps - 1)
    ref = closed_form(l, k, m, c, rho, (L + NK + 1) // 2)
    assert abs(q_ok - ref) <= 1e-13 * a
    assert abs(q_bad - ref) > 1e-8 * a
    q_bad, _ = quadrature(l, k, m, c, rho, zx, zz, n_chi - 1, n_eps)
    assert abs(q_bad - ref) > 1e-8 * a
