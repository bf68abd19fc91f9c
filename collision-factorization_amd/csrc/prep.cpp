// Host-side preparation: real Gaunt COO, 1-D quadrature rules, sumfac tables.
// Native restatement of the reference's MATLAB host code (cited per function).
#include "prep.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <stdexcept>

namespace cfk {

namespace {

// log(n!) table built by sequential accumulation (src/SHL/gaunt_compute_values.m:3)
std::vector<double> log_factorials(int n) {
    std::vector<double> lf(n + 1, 0.0);
    double acc = 0.0;
    for (int i = 1; i <= n; ++i) {
        acc += std::log((double)i);
        lf[i] = acc;
    }
    return lf;
}

double sgn_pow(int e) { return (e % 2 == 0) ? 1.0 : -1.0; }

// Wigner 3j (l1 l2 l3; 0 0 0) by the log-space Racah closed form (gaunt_compute_values.m:49-57)
double w3j_zero(int l1, int l2, int l3, const std::vector<double> &lf) {
    int J = l1 + l2 + l3;
    if (J % 2) return 0.0;
    int g = J / 2;
    double lw = 0.5 * (lf[2 * g - 2 * l1] + lf[2 * g - 2 * l2] + lf[2 * g - 2 * l3] - lf[2 * g + 1]) + lf[g] -
                (lf[g - l1] + lf[g - l2] + lf[g - l3]);
    return sgn_pow(g) * std::exp(lw);
}

// Wigner 3j (l1 l2 l3; m1 m2 m3) by the Racah sum in log space (gaunt_compute_values.m:59-91)
double w3j(int l1, int l2, int l3, int m1, int m2, int m3, const std::vector<double> &lf) {
    if (m1 + m2 + m3 != 0) return 0.0;
    double log_delta = 0.5 * (lf[l1 + l2 - l3] + lf[l1 - l2 + l3] + lf[-l1 + l2 + l3] - lf[l1 + l2 + l3 + 1]);
    int tmin = std::max({0, l2 - l3 - m1, l1 - l3 + m2});
    int tmax = std::min({l1 - m1, l2 + m2, l1 + l2 - l3});
    double log_mag = 0.5 * (lf[l1 + m1] + lf[l1 - m1] + lf[l2 + m2] + lf[l2 - m2] + lf[l3 + m3] + lf[l3 - m3]);
    double s = 0.0;
    for (int t = tmin; t <= tmax; ++t) {
        double term = lf[t] + lf[l1 + l2 - l3 - t] + lf[l1 - m1 - t] + lf[l2 + m2 - t] + lf[l3 - l2 + m1 + t] +
                      lf[l3 - l1 - m2 + t];
        s += sgn_pow(t) * std::exp(-term);
    }
    return sgn_pow(l1 - l2 - m3) * std::exp(log_delta + log_mag) * s;
}

// Nonzeros of column (complex index) c of the complex->real matrix U (src/SHL/complex2real.m):
// U[p,p] = diag[p], U[p,d-1-p] = adiag[p] inside each l block.
int u_column(int L, int c, int rows[2], std::complex<double> vals[2]) {
    (void)L;
    int l = l_of_q0(c);
    int off = l * l, d = 2 * l + 1, p = c - off;
    const double s = 1.0 / std::sqrt(2.0);
    auto diag = [&](int r) -> std::complex<double> {
        int m = r - l;
        if (m < 0) return {0.0, s};
        if (m == 0) return {1.0, 0.0};
        return {sgn_pow(m) * s, 0.0};
    };
    auto adiag = [&](int r) -> std::complex<double> {
        int m = r - l;
        if (m < 0) return {0.0, -sgn_pow(-m) * s};
        if (m == 0) return {0.0, 0.0};
        return {s, 0.0};
    };
    int n = 0;
    rows[n] = off + p;
    vals[n++] = diag(p);
    int pr = d - 1 - p;  // row whose anti-diagonal entry sits in column p
    if (pr != p) {
        rows[n] = off + pr;
        vals[n++] = adiag(pr);
    }
    return n;
}

}  // namespace

RealGaunt real_gaunt(int L) {
    if (L < 0 || L > 40) throw std::invalid_argument("real_gaunt: L out of range [0,40]");
    const int nq = (L + 1) * (L + 1);
    auto lf = log_factorials(3 * L + 2);
    std::vector<double> G((size_t)nq * nq * nq, 0.0);  // G[a + nq*(b + nq*c)] (real part)
    std::map<std::array<int, 3>, double> w0cache;
    for (int l1 = 0; l1 <= L; ++l1)
        for (int l2 = 0; l2 <= L; ++l2) {
            int lo = std::abs(l1 - l2), hi = std::min(L, l1 + l2);
            int start = ((l1 + l2 + lo) % 2 == 0) ? lo : lo + 1;
            for (int l3 = start; l3 <= hi; l3 += 2) {
                double w0 = w3j_zero(l1, l2, l3, lf);
                if (w0 == 0.0) continue;
                double pref0 = std::sqrt((2.0 * l1 + 1) * (2.0 * l2 + 1) * (2.0 * l3 + 1) / (4.0 * M_PI)) * w0;
                for (int m1 = -l1; m1 <= l1; ++m1)
                    for (int m2 = -l2; m2 <= l2; ++m2) {
                        int m3 = m1 + m2;
                        if (std::abs(m3) > l3) continue;
                        double v = sgn_pow(std::abs(m3)) * pref0 * w3j(l1, l2, l3, m1, m2, -m3, lf);
                        if (v == 0.0) continue;
                        int ci = l1 * l1 + l1 + m1, cj = l2 * l2 + l2 + m2, ck = l3 * l3 + l3 + m3;
                        int ra[2], rb[2], rc[2];
                        std::complex<double> ua[2], ub[2], uc[2];
                        int na = u_column(L, ci, ra, ua), nb = u_column(L, cj, rb, ub), nc = u_column(L, ck, rc, uc);
                        for (int x = 0; x < na; ++x)
                            for (int y = 0; y < nb; ++y)
                                for (int z = 0; z < nc; ++z) {
                                    std::complex<double> w = ua[x] * ub[y] * std::conj(uc[z]) * v;
                                    G[(size_t)ra[x] + (size_t)nq * (rb[y] + (size_t)nq * rc[z])] += w.real();
                                }
                    }
            }
        }
    RealGaunt out;
    out.L = L;
    // MATLAB find order: column-major scan (q1 fastest, then q2, then q3)
    for (int c = 0; c < nq; ++c)
        for (int b = 0; b < nq; ++b)
            for (int a = 0; a < nq; ++a) {
                double v = G[(size_t)a + (size_t)nq * (b + (size_t)nq * c)];
                if (std::fabs(v) > 1e-12) {
                    out.q1.push_back(a + 1);
                    out.q2.push_back(b + 1);
                    out.q3.push_back(c + 1);
                    out.val.push_back(v);
                }
            }
    // ic_map = rank of the (l1,l2,l3) row in lexicographic order (unique(...,'rows'))
    std::map<std::array<int, 3>, int> rank;
    size_t n = out.val.size();
    for (size_t z = 0; z < n; ++z)
        rank[{l_of_q0(out.q1[z] - 1), l_of_q0(out.q2[z] - 1), l_of_q0(out.q3[z] - 1)}] = 0;
    int r = 0;
    for (auto &kv : rank) kv.second = ++r;
    out.n_l = r;
    out.t.resize(n);
    for (size_t z = 0; z < n; ++z)
        out.t[z] = rank[{l_of_q0(out.q1[z] - 1), l_of_q0(out.q2[z] - 1), l_of_q0(out.q3[z] - 1)}];
    return out;
}

// Gauss-Legendre nodes by Newton iteration on the Legendre recursion from the
// Chebyshev-plus-correction start (src/Gauss.m:38-105).
void gauss_legendre(int n, double a, double b, double *xo, double *wo) {
    if (n < 1) throw std::invalid_argument("gauss_legendre: n < 1");
    if (n == 1) {
        xo[0] = 0.5 * (a + b);
        wo[0] = b - a;
        return;
    }
    const int N = n - 1, N1 = N + 1, N2 = N + 2;
    std::vector<double> y(N1), y0(N1), Lp(N1);
    for (int i = 0; i < N1; ++i) {
        double xu = -1.0 + 2.0 * i / (double)(N1 - 1);
        y[i] = std::cos((2.0 * i + 1.0) * M_PI / (2.0 * N + 2.0)) + (0.27 / N1) * std::sin(M_PI * xu * N / N2);
    }
    std::fill(y0.begin(), y0.end(), 2.0);
    for (int it = 0; it < 100; ++it) {
        double md = 0.0;
        for (int i = 0; i < N1; ++i) md = std::max(md, std::fabs(y[i] - y0[i]));
        if (!(md > 2.220446049250313e-16)) break;
        for (int i = 0; i < N1; ++i) {
            double L0 = 1.0, L1 = y[i];
            double Lkm1 = L0, Lk = L1;  // L(:,k) and L(:,k+1) columns
            for (int k = 2; k <= N1; ++k) {
                double Ln = ((2 * k - 1) * y[i] * Lk - (k - 1) * Lkm1) / k;
                Lkm1 = Lk;
                Lk = Ln;
            }
            // Lkm1 = L(:,N1), Lk = L(:,N2)
            Lp[i] = N2 * (Lkm1 - y[i] * Lk) / (1.0 - y[i] * y[i]);
            y0[i] = y[i];
            y[i] = y0[i] - Lk / Lp[i];
        }
    }
    for (int i = 0; i < N1; ++i) {
        int j = N1 - 1 - i;  // flipud
        xo[i] = (a * (1.0 - y[j]) + b * (1.0 + y[j])) / 2.0;
        wo[i] = (b - a) / ((1.0 - y[j] * y[j]) * Lp[j] * Lp[j]) * ((double)N2 / N1) * ((double)N2 / N1);
    }
}

namespace {
// Eigenvalues of a symmetric tridiagonal matrix by Sturm-sequence bisection (ascending).
void tridiag_eigvals(const std::vector<double> &d, const std::vector<double> &e, std::vector<double> &ev) {
    int n = (int)d.size();
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r);
        hi = std::max(hi, d[i] + r);
    }
    auto count_below = [&](double x) {
        int c = 0;
        double q = d[0] - x;
        if (q < 0) ++c;
        for (int i = 1; i < n; ++i) {
            double qq = (q == 0.0) ? 1e-300 : q;
            q = d[i] - x - e[i - 1] * e[i - 1] / qq;
            if (q < 0) ++c;
        }
        return c;
    };
    ev.resize(n);
    for (int k = 0; k < n; ++k) {
        double a = lo, b = hi;
        for (int it = 0; it < 200 && b - a > 1e-15 * std::max(1.0, std::fabs(a) + std::fabs(b)); ++it) {
            double m = 0.5 * (a + b);
            if (count_below(m) > k)
                b = m;
            else
                a = m;
        }
        ev[k] = 0.5 * (a + b);
    }
}
}  // namespace

// Generalized Gauss-Laguerre: Jacobi-matrix eigenvalue seeds, Newton polishing on the
// three-term recursion, derivative weights (src/Gauss.m:128-173).
void gauss_gen_laguerre(int n, double alpha, double *xo, double *wo) {
    if (n < 1) throw std::invalid_argument("gauss_gen_laguerre: n < 1");
    std::vector<double> d(n), e(n > 1 ? n - 1 : 1), x;
    for (int i = 1; i <= n; ++i) d[i - 1] = 2.0 * i - 1.0 + alpha;
    for (int i = 1; i < n; ++i) e[i - 1] = std::sqrt(i * (i + alpha));
    tridiag_eigvals(d, e, x);
    std::vector<double> Lp1(n);
    for (int it = 0; it < 15; ++it) {
        double mdx = 0.0;
        for (int i = 0; i < n; ++i) {
            double L0 = 1.0, L1 = 1.0 + alpha - x[i], P0 = 0.0, P1 = -1.0;
            for (int k = 2; k <= n; ++k) {
                double L2 = ((2 * k - 1 + alpha - x[i]) * L1 - (k - 1 + alpha) * L0) / k;
                double P2 = ((2 * k - 1 + alpha - x[i]) * P1 - L1 - (k - 1 + alpha) * P0) / k;
                L0 = L1;
                L1 = L2;
                P0 = P1;
                P1 = P2;
            }
            if (n == 1) {
                L1 = 1.0 + alpha - x[i];
                P1 = -1.0;
            }
            double dx = L1 / P1;
            x[i] -= dx;
            Lp1[i] = P1;
            mdx = std::max(mdx, std::fabs(dx));
        }
        if (mdx < 1e-15) break;
    }
    double lc = std::lgamma(n + alpha + 1.0) - std::lgamma(n + 1.0);
    for (int i = 0; i < n; ++i) {
        xo[i] = x[i];
        wo[i] = std::exp(lc) / (x[i] * Lp1[i] * Lp1[i]);
    }
}

void sumfac_grid_sizes(int K_max, int L, int rp, int ap, int s[7]) {
    s[0] = rp + (int)std::ceil((3 * K_max + 1.5 * L + 3) / 2.0);
    s[1] = rp + 4 * K_max + 3 * L + 4;
    s[2] = ap + K_max + (int)std::ceil(1.5 * L) + 1;
    s[3] = ap + 4 * K_max + 3 * L + 4;
    s[4] = rp + 3 * K_max + (int)std::floor(1.5 * L) + 3;
    s[5] = K_max + (int)std::ceil(0.5 * L) + 1;
    s[6] = 2 * K_max + L + 1;
}

double legendre_plm(int l, int m, double x) {
    double pmm = 1.0;
    if (m > 0) {
        double s = std::sqrt(std::max((1.0 - x) * (1.0 + x), 0.0)), fact = 1.0;
        for (int i = 1; i <= m; ++i) {
            pmm *= -fact * s;
            fact += 2.0;
        }
    }
    if (l == m) return pmm;
    double p1 = x * (2 * m + 1) * pmm;
    if (l == m + 1) return p1;
    double pl = 0.0;
    for (int ll = m + 2; ll <= l; ++ll) {
        pl = (x * (2 * ll - 1) * p1 - (ll + m - 1) * pmm) / (ll - m);
        pmm = p1;
        p1 = pl;
    }
    return pl;
}

void sh_norm(int L, double *out) {
    for (int l = 0; l <= L; ++l)
        for (int m = -l; m <= l; ++m) {
            int am = std::abs(m);
            double base = std::sqrt(((2.0 * l + 1) / (4.0 * M_PI)) * std::exp(std::lgamma(l - am + 1.0) - std::lgamma(l + am + 1.0)));
            out[l * l + l + m] = (m == 0) ? base : std::sqrt(2.0) * base;
        }
}

void radial_norm(int N_K, int L, double *out) {
    for (int k = 0; k < N_K; ++k)
        for (int l = 0; l <= L; ++l) {
            double lm = -std::log(2.0) + std::lgamma(k + l + 1.5) - std::lgamma(k + 1.0);
            out[k + N_K * l] = std::exp(-0.5 * lm);
        }
}

SumfacGrids sumfac_grids(double alpha, const int s[7]) {
    SumfacGrids g;
    auto mk = [](std::vector<double> &x, std::vector<double> &w, int n) {
        x.resize(n);
        w.resize(n);
    };
    mk(g.x, g.W_x, s[0]);
    gauss_gen_laguerre(s[0], alpha / 2.0, g.x.data(), g.W_x.data());
    mk(g.u1, g.W_u1, s[1]);
    gauss_legendre(s[1], 0.0, 1.0, g.u1.data(), g.W_u1.data());
    mk(g.t1, g.W_t1, s[2]);
    gauss_legendre(s[2], 0.0, 1.0, g.t1.data(), g.W_t1.data());
    mk(g.y2, g.W_y2, s[3]);
    gauss_legendre(s[3], 0.0, 1.0, g.y2.data(), g.W_y2.data());
    mk(g.t2, g.W_t2, s[4]);
    gauss_legendre(s[4], 0.0, 1.0, g.t2.data(), g.W_t2.data());
    mk(g.mu_chi, g.W_chi, s[5]);
    gauss_legendre(s[5], -1.0, 1.0, g.mu_chi.data(), g.W_chi.data());
    mk(g.eps, g.W_eps, s[6]);
    for (int i = 0; i < s[6]; ++i) {
        g.eps[i] = 2.0 * M_PI * i / s[6];
        g.W_eps[i] = 2.0 * M_PI / s[6];
    }
    return g;
}

SumfacTablePlan sumfac_table_plan(const RealGaunt &g) {
    const int L = g.L, nq = (L + 1) * (L + 1), n_l = g.n_l;
    SumfacTablePlan P;
    SumfacTables &T = P.T;
    T.n_l = n_l;
    T.N_Q = nq;
    T.L = L;
    T.trip.assign((size_t)n_l * 3, 0);
    T.qi_off.assign(n_l + 1, 0);
    P.scale.assign(n_l, 0.0);
    P.loss_off.assign(n_l + 1, 0);
    P.ent_off.push_back(0);
    std::vector<std::vector<size_t>> members(n_l);
    for (size_t z = 0; z < g.val.size(); ++z) members[g.t[z] - 1].push_back(z);
    for (int t = 0; t < n_l; ++t) {
        const auto &mz = members[t];
        const size_t z0 = mz.front();
        const int li = l_of_q0(g.q1[z0] - 1), lj = l_of_q0(g.q2[z0] - 1), lk = l_of_q0(g.q3[z0] - 1);
        T.trip[3 * t] = li;
        T.trip[3 * t + 1] = lj;
        T.trip[3 * t + 2] = lk;
        const int qi0 = li * li + li, qj0 = lj * lj + lj;  // 0-based m=0 indices
        const double Yi0 = std::sqrt((2.0 * li + 1) / (4.0 * M_PI)), Yj0 = std::sqrt((2.0 * lj + 1) / (4.0 * M_PI));
        double D = 0.0;
        for (size_t z : mz) D += g.val[z] * g.val[z];
        P.scale[t] = (8.0 * M_PI * M_PI * Yj0) / D;
        // slots in order of first appearance of q_i among the entries with q_j = q_j0; each slot's
        // entries in member order (the order the reference accumulates them)
        std::vector<int> qlist;
        std::vector<std::vector<size_t>> slot_z;
        for (size_t z : mz) {
            const int qi = g.q1[z] - 1, qj = g.q2[z] - 1;
            if (qj != qj0) continue;
            int slot = -1;
            for (size_t s = 0; s < qlist.size(); ++s)
                if (qlist[s] == qi) slot = (int)s;
            if (slot < 0) {
                slot = (int)qlist.size();
                qlist.push_back(qi);
                slot_z.emplace_back();
            }
            slot_z[slot].push_back(z);
            if (qi == qi0) {
                P.loss_qk.push_back(g.q3[z] - 1);
                P.loss_g.push_back(g.val[z] * Yi0);
            }
        }
        P.loss_off[t + 1] = (int)P.loss_qk.size();
        for (size_t s = 0; s < qlist.size(); ++s) {
            T.qi.push_back(qlist[s]);
            P.slot_t.push_back(t);
            for (size_t z : slot_z[s]) {
                P.ent_qk.push_back(g.q3[z] - 1);
                P.ent_g.push_back(g.val[z]);
            }
            P.ent_off.push_back((int)P.ent_qk.size());
        }
        T.qi_off[t + 1] = (int)T.qi.size();
    }
    return P;
}

SumfacTables compact_mex_tables(int N_Q, int n_l, const double *L_triplets, const double *qi_valid_mat, int Nu,
                                int Nt, int Ny, const double *P_loss_p1, const double *P_gain_p1,
                                const double *P_loss_p2, const double *P_gain_p2, std::string &err) {
    SumfacTables T;
    T.n_l = n_l;
    T.N_Q = N_Q;
    int L = 0;
    while ((L + 1) * (L + 1) < N_Q) ++L;
    if ((L + 1) * (L + 1) != N_Q) {
        err = "N_Q is not a perfect square";
        return T;
    }
    T.L = L;
    T.trip.resize((size_t)n_l * 3);
    T.qi_off.assign(n_l + 1, 0);
    for (int t = 0; t < n_l; ++t) {
        for (int c = 0; c < 3; ++c) {
            double v = L_triplets[t + (size_t)n_l * c];
            if (!(v >= 0 && v <= L) || v != std::floor(v)) {
                err = "L_triplets entry out of range";
                return T;
            }
            T.trip[3 * t + c] = (int)v;
        }
        // qi_valid_mat[idx + N_Q*t], 1-based, -1 terminated (compute_rtensor_sumfac_mex.cpp:219-222)
        for (int idx = 0; idx < N_Q; ++idx) {
            double v = qi_valid_mat[idx + (size_t)N_Q * t];
            if ((int)v == -1) break;
            int q = (int)v - 1;
            if (q < 0 || q >= N_Q) {
                err = "qi_valid_mat entry out of range";
                return T;
            }
            T.qi.push_back(q);
        }
        T.qi_off[t + 1] = (int)T.qi.size();
    }
    const size_t ns = T.qi.size();
    T.P_gain_p1.resize((size_t)Nu * Nt * ns);
    T.P_gain_p2.resize((size_t)Ny * ns);
    T.P_loss_p1.resize((size_t)Nu * Nt * n_l);
    T.P_loss_p2.resize((size_t)Ny * n_l);
    for (int t = 0; t < n_l; ++t) {
        for (int s = T.qi_off[t]; s < T.qi_off[t + 1]; ++s) {
            const int q = T.qi[s];
            for (int u = 0; u < Nu; ++u)
                for (int tt = 0; tt < Nt; ++tt)
                    T.P_gain_p1[((size_t)u * Nt + tt) * ns + s] =
                        P_gain_p1[tt + (size_t)Nt * (u + (size_t)Nu * (q + (size_t)N_Q * t))];
            for (int i = 0; i < Ny; ++i) T.P_gain_p2[(size_t)i * ns + s] = P_gain_p2[i + (size_t)Ny * (q + (size_t)N_Q * t)];
        }
        for (int u = 0; u < Nu; ++u)
            for (int tt = 0; tt < Nt; ++tt)
                T.P_loss_p1[((size_t)u * Nt + tt) * n_l + t] = P_loss_p1[tt + (size_t)Nt * (u + (size_t)Nu * t)];
        for (int i = 0; i < Ny; ++i) T.P_loss_p2[(size_t)i * n_l + t] = P_loss_p2[i + (size_t)Ny * t];
    }
    return T;
}

}  // namespace cfk
