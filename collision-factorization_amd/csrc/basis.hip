// Basis evaluation at arbitrary velocities and moment diagnostics of coefficient vectors.
//
// phi_{klm}(v) = N_kl L_k^{(l+1/2)}(|v|^2) |v|^l Y_lm(v/|v|) (src/SpectralBasis.m:25-58): real spherical
// harmonics with the Condon-Shortley phase (m > 0: sqrt(2) cos(m phi), m < 0: sqrt(2) sin(|m| phi)),
// N_kl = exp(-(-log 2 + lnGamma(k+l+3/2) - lnGamma(k+1))/2), column k*N_Q + q with q = l^2 + l + m, and
// the reference's r = sqrt(|v|^2) + eps in the solid harmonic and the polar angle.
//
// Moments of f(v) = e^{-|v|^2} sum_i c_i phi_i(v) up to second order are exact linear functionals of
// the coefficients of degree l <= 2: with the radial integrals
//   int_0^inf e^{-r^2} L_k^{(a)}(r^2) r^{2l+2+2j} dr = Gamma(a+1)/2 (k = 0, j = 0),
//   Gamma(a+2)/2 (k = 0, j = 1), -Gamma(a+2)/2 (k = 1, j = 1), 0 otherwise (a = l + 1/2),
// and the monomials written in real spherical harmonics (1 = sqrt(4 pi) Y_00, v_z = r sqrt(4pi/3) Y_10,
// v_x = -r sqrt(4pi/3) Y_11, v_y = -r sqrt(4pi/3) Y_1-1, xy = r^2 sqrt(4pi/15) Y_2-2,
// xz = -r^2 sqrt(4pi/15) Y_21, yz = -r^2 sqrt(4pi/15) Y_2-1, x^2 - y^2 = r^2 sqrt(16pi/15) Y_22,
// 3z^2 - r^2 = r^2 sqrt(16pi/5) Y_20), each moment is a sum of at most four weighted coefficients.
#include <hip/hip_runtime.h>

#include <cmath>
#include <initializer_list>
#include <utility>
#include <vector>

#include "common.hpp"

namespace cfk {

namespace {

constexpr int BASIS_LMAX = 16;

// one thread per point: v column-major [3][n] (MATLAB's n x 3 array), Psi column-major [n][N_terms]
__global__ __launch_bounds__(128) void basis_eval_kernel(double *__restrict__ Psi, const double *__restrict__ v,
                                                         int64_t n, int K, int L, const double *__restrict__ norm_kl,
                                                         const double *__restrict__ nf_lm) {
    const int64_t p = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (p >= n) return;
    const int NQ = (L + 1) * (L + 1);
    const double vx = v[p], vy = v[n + p], vz = v[2 * n + p];
    const double r2 = vx * vx + vy * vy + vz * vz;
    const double r = sqrt(r2) + 2.220446049250313e-16;
    const double theta = acos(vz / r), ph = atan2(vy, vx);
    const double x = cos(theta);
    const double s = sqrt(fmax((1.0 - x) * (1.0 + x), 0.0));
    // P_{l-1}^m and P_{l-2}^m for every m (Condon-Shortley phase), advanced one degree per l
    double pm1[BASIS_LMAX + 1], pm2[BASIS_LMAX + 1], pmm[BASIS_LMAX + 1];
    {
        double pv = 1.0, fact = 1.0;
#pragma unroll
        for (int m = 0; m <= BASIS_LMAX; ++m) {  // P_m^m
            pmm[m] = pv;
            pv = -pv * fact * s;
            fact += 2.0;
            pm1[m] = 0.0;
            pm2[m] = 0.0;
        }
    }
    double rl = 1.0;  // r^l
    for (int l = 0; l <= L; ++l) {
        double P[BASIS_LMAX + 1];
#pragma unroll
        for (int m = 0; m <= BASIS_LMAX; ++m) {
            if (m > l) break;
            double val;
            if (m == l)
                val = pmm[m];
            else if (m == l - 1)
                val = x * (2 * m + 1) * pmm[m];
            else
                val = (x * (2 * l - 1) * pm1[m] - (l + m - 1) * pm2[m]) / (l - m);
            P[m] = val;
            pm2[m] = pm1[m];
            pm1[m] = val;
        }
        // Laguerre L_k^{(l+1/2)}(r^2) by the three-term recursion, one degree at a time
        const double a = l + 0.5;
        double lk2 = 0.0, lk1 = 0.0;
        for (int k = 0; k < K; ++k) {
            double lk;
            if (k == 0)
                lk = 1.0;
            else if (k == 1)
                lk = 1.0 + a - r2;
            else
                lk = ((2 * (k - 1) + 1 + a - r2) * lk1 - ((k - 1) + a) * lk2) / k;
            lk2 = lk1;
            lk1 = lk;
            const double rad = norm_kl[k * (L + 1) + l] * lk * rl;
            for (int m = -l; m <= l; ++m) {
                const int am = m < 0 ? -m : m;
                double y;
                if (m > 0)
                    y = 1.4142135623730951 * nf_lm[l * (BASIS_LMAX + 1) + am] * P[am] * cos(am * ph);
                else if (m == 0)
                    y = nf_lm[l * (BASIS_LMAX + 1)] * P[0];
                else
                    y = 1.4142135623730951 * nf_lm[l * (BASIS_LMAX + 1) + am] * P[am] * sin(am * ph);
                const int q = l * l + l + m;
                Psi[(int64_t)(k * NQ + q) * n + p] = rad * y;
            }
        }
        rl *= r;
    }
}

// Moment weights: m_j = sum_t w[j][t] c[idx[j][t]] over at most 4 terms.
struct MomentWeights {
    int idx[10][4];
    double w[10][4];
    int nt[10];
};

__global__ __launch_bounds__(256) void moments_kernel(double *__restrict__ out, const double *__restrict__ c, int64_t B,
                                                      int NK, int layout, MomentWeights mw) {
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    double m[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        double acc = 0.0;
        for (int t = 0; t < mw.nt[j]; ++t) {
            const int64_t i = mw.idx[j][t];
            const double ci = layout == CFK_LAYOUT_NATURAL ? c[b * NK + i] : c[i * B + b];
            acc += mw.w[j][t] * ci;
        }
        m[j] = acc;
    }
    // m: int f, int v_x f, int v_y f, int v_z f, int v_x^2 f, v_y^2, v_z^2, v_x v_y, v_x v_z, v_y v_z
    const double n = m[0];
    const double ux = m[1] / n, uy = m[2] / n, uz = m[3] / n;
    const double pxx = m[4] - n * ux * ux, pyy = m[5] - n * uy * uy, pzz = m[6] - n * uz * uz;
    const double pxy = m[7] - n * ux * uy, pxz = m[8] - n * ux * uz, pyz = m[9] - n * uy * uz;
    double *o = out + b * CFK_N_MOMENTS;
    o[0] = n;
    o[1] = ux;
    o[2] = uy;
    o[3] = uz;
    o[4] = 2.0 * (pxx + pyy + pzz) / (3.0 * n);  // f ~ n (pi T)^{-3/2} exp(-|v-u|^2 / T): int (v_d-u_d)^2 f = n T / 2
    o[5] = pxx;
    o[6] = pyy;
    o[7] = pzz;
    o[8] = pxy;
    o[9] = pxz;
    o[10] = pyz;
}

double n_kl(int k, int l) { return std::exp(-0.5 * (-std::log(2.0) + std::lgamma(k + l + 1.5) - std::lgamma(k + 1.0))); }

// sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) for m = 0..l
double nf_lm_host(int l, int m) {
    return std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * std::exp(std::lgamma(l - m + 1.0) - std::lgamma(l + m + 1.0)));
}

// The ten second-order moment functionals in (index into c, weight) form; index k*N_Q + q.
int moment_weights(int N_Q, int K, MomentWeights &mw) {
    if (N_Q < 9 || K < 2) return set_error(CFK_EUNSUP, "moments need L >= 2 and K >= 2 (N_Q=%d, K=%d)", N_Q, K);
    const double g32 = std::tgamma(1.5), g52 = std::tgamma(2.5), g72 = std::tgamma(3.5);
    const double s4p = std::sqrt(4.0 * M_PI), s4p3 = std::sqrt(4.0 * M_PI / 3.0), s4p15 = std::sqrt(4.0 * M_PI / 15.0);
    const double s16p15 = std::sqrt(16.0 * M_PI / 15.0), s16p5 = std::sqrt(16.0 * M_PI / 5.0);
    // radial parts: l = 0 (k = 0: int r^2 (j=0) and r^4 (j=1); k = 1: j = 1), l = 1 and l = 2 (k = 0, j = 0)
    const double m0 = n_kl(0, 0) * s4p * 0.5 * g32;                 // int f        per c_{0,00}
    const double r2_0 = n_kl(0, 0) * s4p * 0.5 * g52;               // int r^2 f    per c_{0,00}
    const double r2_1 = -n_kl(1, 0) * s4p * 0.5 * g52;              // int r^2 f    per c_{1,00}
    const double m1 = n_kl(0, 1) * s4p3 * 0.5 * g52;                // int v f      per c_{0,1m} (up to sign)
    const double m2 = n_kl(0, 2) * 0.5 * g72;                       // int r^2 Y_2m f per c_{0,2m}
    const int q00 = 0, q1m1 = 1, q10 = 2, q11 = 3, q2m2 = 4, q2m1 = 5, q20 = 6, q21 = 7, q22 = 8;
    const int k1 = N_Q;  // offset of k = 1
    auto set = [&](int j, std::initializer_list<std::pair<int, double>> terms) {
        mw.nt[j] = 0;
        for (auto &t : terms) {
            mw.idx[j][mw.nt[j]] = t.first;
            mw.w[j][mw.nt[j]] = t.second;
            ++mw.nt[j];
        }
    };
    set(0, {{q00, m0}});
    set(1, {{q11, -m1}});
    set(2, {{q1m1, -m1}});
    set(3, {{q10, m1}});
    // with R2 = int r^2 f, Z = int (3z^2 - r^2) f, D = int (x^2 - y^2) f:
    // int z^2 f = R2/3 + Z/3, int x^2 f = R2/3 - Z/6 + D/2, int y^2 f = R2/3 - Z/6 - D/2
    const double z = s16p5 * m2, d = s16p15 * m2;
    set(4, {{q00, r2_0 / 3.0}, {k1 + q00, r2_1 / 3.0}, {q20, -z / 6.0}, {q22, d / 2.0}});
    set(5, {{q00, r2_0 / 3.0}, {k1 + q00, r2_1 / 3.0}, {q20, -z / 6.0}, {q22, -d / 2.0}});
    set(6, {{q00, r2_0 / 3.0}, {k1 + q00, r2_1 / 3.0}, {q20, z / 3.0}});
    set(7, {{q2m2, s4p15 * m2}});
    set(8, {{q21, -s4p15 * m2}});
    set(9, {{q2m1, -s4p15 * m2}});
    return CFK_OK;
}

}  // namespace

}  // namespace cfk

using namespace cfk;

extern "C" int cfk_basis_evaluate(double *Psi_dev, const double *v_dev, int64_t n_pts, int32_t K, int32_t L,
                                  void *stream) {
    if (n_pts < 0) return set_error(CFK_EINVAL, "n_pts must be >= 0");
    if (K < 1 || L < 0 || L > BASIS_LMAX) return set_error(CFK_EUNSUP, "basis with K=%d, L=%d (K >= 1, L <= %d)", K, L, BASIS_LMAX);
    if (n_pts == 0) return CFK_OK;
    if (!Psi_dev || !v_dev) return set_error(CFK_EINVAL, "null pointer");
    int rc;
    if ((rc = check_device())) return rc;
    std::vector<double> nkl((size_t)K * (L + 1)), nf((size_t)(L + 1) * (BASIS_LMAX + 1), 0.0);
    for (int k = 0; k < K; ++k)
        for (int l = 0; l <= L; ++l) nkl[(size_t)k * (L + 1) + l] = n_kl(k, l);
    for (int l = 0; l <= L; ++l)
        for (int m = 0; m <= l; ++m) nf[(size_t)l * (BASIS_LMAX + 1) + m] = nf_lm_host(l, m);
    DevBuf<double> dn, dnf;
    if ((rc = dn.upload(nkl.data(), nkl.size())) || (rc = dnf.upload(nf.data(), nf.size()))) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(basis_eval_kernel, dim3((unsigned)((n_pts + 127) / 128)), dim3(128), 0, st, Psi_dev, v_dev, n_pts,
                       K, L, dn.p, dnf.p);
    CFK_HIP(hipGetLastError());
    CFK_HIP(hipStreamSynchronize(st));  // the coefficient tables are freed on return
    return CFK_OK;
}

extern "C" int cfk_moment_weights(int32_t N_Q, int32_t K, int32_t *idx, double *w) {
    if (!idx || !w) return set_error(CFK_EINVAL, "null pointer");
    MomentWeights mw;
    int rc;
    if ((rc = moment_weights(N_Q, K, mw))) return rc;
    for (int j = 0; j < 10; ++j)
        for (int t = 0; t < 4; ++t) {
            idx[4 * j + t] = t < mw.nt[j] ? mw.idx[j][t] : -1;
            w[4 * j + t] = t < mw.nt[j] ? mw.w[j][t] : 0.0;
        }
    return CFK_OK;
}

extern "C" int cfk_moments(double *out_dev, const double *c_dev, int64_t B, int32_t N_Q, int32_t K, int32_t layout,
                           void *stream) {
    if (B < 0) return set_error(CFK_EINVAL, "B must be >= 0");
    if (layout != CFK_LAYOUT_NATURAL && layout != CFK_LAYOUT_BATCH_FASTEST)
        return set_error(CFK_EINVAL, "unknown layout %d", layout);
    MomentWeights mw;
    int rc;
    if ((rc = moment_weights(N_Q, K, mw))) return rc;
    if (B == 0) return CFK_OK;
    if (!out_dev || !c_dev) return set_error(CFK_EINVAL, "null pointer");
    if ((rc = check_device())) return rc;
    hipLaunchKernelGGL(moments_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, (hipStream_t)stream, out_dev,
                       c_dev, B, N_Q * K, layout, mw);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}
