// Projection of Maxwellian mixtures onto the Laguerre x real-spherical-harmonic basis, in FP64:
//   c_i = sum_points W(v) Psi_i(v) f(v) / e^{-|v|^2}
// on the grid of tutorial_collisions_spectral_properties.m:100-158: generalized Gauss-Laguerre
// (alpha = 1/2) in x = |v|^2 with weights w/2, Gauss-Legendre in cos(theta), uniform phi.
//
// The basis factorizes as Psi_{k,l,m}(v) = R_{k,l}(|v|) N_{lm} P_l^{|m|}(cos theta) T_m(phi), so the
// sum runs as three small contractions per radial node instead of one dense N_pts x N_basis
// product: over phi against the trig table, over theta against the normalized Legendre table,
// then over the radial nodes against the radial table.  One workgroup per distribution.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "common.hpp"
#include "prep.hpp"

namespace cfk {
namespace {

constexpr int PT = 256;     // threads per workgroup
constexpr int PMAXK = 16;   // radial modes held per thread

struct ProjDev {
    const double *rad;   // [K][L+1][N_r]  R_{k,l}(r_a) * w_a / 2
    const double *leg;   // [N_Q][N_p]     N_{lm} P_l^{|m|}(mu_p) w_p
    const double *trig;  // [2L+1][N_phi]  (1, cos phi, sin phi, cos 2phi, sin 2phi, ...) * 2 pi / N_phi
    const double *r;     // [N_r]
    const double *mu;    // [N_p]
    const double *cphi, *sphi;  // [N_phi]
    int L, K, N_Q, N_r, N_p, N_phi, n_comp;
};

__device__ inline int trig_row(int q) {  // 0-based q = l^2 + l + m  ->  row of the trig table
    int l = 0;
    while ((l + 1) * (l + 1) <= q) ++l;
    const int m = q - l * l - l;
    return m == 0 ? 0 : (m > 0 ? 2 * m - 1 : -2 * m);
}

__device__ inline int l_of(int q) {
    int l = 0;
    while ((l + 1) * (l + 1) <= q) ++l;
    return l;
}

// params [B][n_comp][7] = (n, ux, uy, uz, Tx, Ty, Tz); component density
//   n pi^{-3/2} (Tx Ty Tz)^{-1/2} exp(-sum_d (v_d - u_d)^2 / T_d)
// c [K*N_Q][B] (batch fastest).
__global__ __launch_bounds__(PT) void project_kernel(ProjDev d, const double *__restrict__ params,
                                                     double *__restrict__ c, int64_t B) {
    extern __shared__ double sm[];
    const int M2 = 2 * d.L + 1, NA = d.N_p * d.N_phi;
    double *leg = sm;                         // N_Q * N_p
    double *trig = leg + d.N_Q * d.N_p;       // M2 * N_phi
    double *Fa = trig + M2 * d.N_phi;         // N_p * N_phi
    double *Bm = Fa + NA;                     // N_p * M2
    double *ux = Bm + d.N_p * M2;             // per point unit vector (sin th cos ph, sin th sin ph, cos th)
    double *uy = ux + NA;
    double *uz = uy + NA;
    const int tid = threadIdx.x;
    const int64_t b = blockIdx.x;
    for (int i = tid; i < d.N_Q * d.N_p; i += PT) leg[i] = d.leg[i];
    for (int i = tid; i < M2 * d.N_phi; i += PT) trig[i] = d.trig[i];
    for (int i = tid; i < NA; i += PT) {
        const int p = i / d.N_phi, j = i % d.N_phi;
        const double st = sqrt(fmax(0.0, (1.0 - d.mu[p]) * (1.0 + d.mu[p])));
        ux[i] = st * d.cphi[j];
        uy[i] = st * d.sphi[j];
        uz[i] = d.mu[p];
    }
    // mixture components: log of the prefactor and inverse temperatures
    double lc[3], cu[3][3], it[3][3];
    const double *pp = params + b * d.n_comp * 7;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        lc[i] = -INFINITY;
        if (i < d.n_comp && pp[7 * i] > 0.0) {
            const double *q = pp + 7 * i;
            lc[i] = log(q[0]) - 1.5 * log(M_PI) - 0.5 * log(q[4] * q[5] * q[6]);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                cu[i][a] = q[1 + a];
                it[i][a] = 1.0 / q[4 + a];
            }
        } else {
#pragma unroll
            for (int a = 0; a < 3; ++a) cu[i][a] = it[i][a] = 0.0;
        }
    }
    const int lq = tid < d.N_Q ? l_of(tid) : 0, tq = tid < d.N_Q ? trig_row(tid) : 0;
    double acc[PMAXK];
#pragma unroll
    for (int k = 0; k < PMAXK; ++k) acc[k] = 0.0;
    __syncthreads();
    for (int a = 0; a < d.N_r; ++a) {
        const double r = d.r[a], r2 = r * r;
        // f(v) e^{|v|^2} on the sphere of radius r (one exponential per component, no overflow)
        for (int i = tid; i < NA; i += PT) {
            const double vx = r * ux[i], vy = r * uy[i], vz = r * uz[i];
            double F = 0.0;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                if (lc[m] == -INFINITY) continue;
                const double dx = vx - cu[m][0], dy = vy - cu[m][1], dz = vz - cu[m][2];
                F += exp(lc[m] + r2 - dx * dx * it[m][0] - dy * dy * it[m][1] - dz * dz * it[m][2]);
            }
            Fa[i] = F;
        }
        __syncthreads();
        // phi: Bm[p][mi] = sum_j trig[mi][j] Fa[p][j]
        for (int o = tid; o < d.N_p * M2; o += PT) {
            const int p = o / M2, mi = o % M2;
            const double *t = trig + mi * d.N_phi, *f = Fa + p * d.N_phi;
            double s = 0.0;
            for (int j = 0; j < d.N_phi; ++j) s = fma(t[j], f[j], s);
            Bm[o] = s;
        }
        __syncthreads();
        // theta, then the radial node: c[k][q] += rad[k][l][a] * sum_p leg[q][p] Bm[p][mi(q)]
        if (tid < d.N_Q) {
            const double *lg = leg + tid * d.N_p;
            double A = 0.0;
            for (int p = 0; p < d.N_p; ++p) A = fma(lg[p], Bm[p * M2 + tq], A);
#pragma unroll
            for (int k = 0; k < PMAXK; ++k)
                if (k < d.K) acc[k] = fma(d.rad[((int64_t)k * (d.L + 1) + lq) * d.N_r + a], A, acc[k]);
        }
        __syncthreads();
    }
    if (tid < d.N_Q) {
#pragma unroll
        for (int k = 0; k < PMAXK; ++k)
            if (k < d.K) c[((int64_t)k * d.N_Q + tid) * B + b] = acc[k];
    }
}

// L_k^{(alpha)}(x) by the three-term recursion (src/SpectralBasis.m:143-162)
double gen_laguerre_k(int k, double alpha, double x) {
    if (k == 0) return 1.0;
    double p2 = 1.0, p1 = 1.0 + alpha - x;
    for (int i = 1; i < k; ++i) {
        const double v = ((2 * i + 1 + alpha - x) * p1 - (i + alpha) * p2) / (i + 1);
        p2 = p1;
        p1 = v;
    }
    return p1;
}

}  // namespace

void projection_radial_table(int L, int K, int n_rad, std::vector<double> &r, std::vector<double> &rad) {
    std::vector<double> x(n_rad), w(n_rad), rn((size_t)K * (L + 1));
    gauss_gen_laguerre(n_rad, 0.5, x.data(), w.data());
    radial_norm(K, L, rn.data());
    r.resize(n_rad);
    for (int a = 0; a < n_rad; ++a) r[a] = std::sqrt(x[a]);
    rad.assign((size_t)K * (L + 1) * n_rad, 0.0);
    for (int k = 0; k < K; ++k)
        for (int l = 0; l <= L; ++l)
            for (int a = 0; a < n_rad; ++a)
                rad[((size_t)k * (L + 1) + l) * n_rad + a] =
                    rn[k + (size_t)K * l] * gen_laguerre_k(k, l + 0.5, x[a]) * std::pow(r[a], l) * 0.5 * w[a];
}

}  // namespace cfk

using namespace cfk;

extern "C" int cfk_project_maxwellians(double *c_dev, const double *params_dev, int64_t B, int32_t n_comp, int32_t L,
                                       int32_t K, int32_t n_rad, int32_t n_pol, int32_t n_azi, void *stream) {
    if (!c_dev || !params_dev) return set_error(CFK_EINVAL, "null argument");
    if (B < 0 || B > 0x7fffffff) return set_error(CFK_EINVAL, "B=%lld out of range", (long long)B);
    if (n_comp < 1 || n_comp > 3) return set_error(CFK_EINVAL, "n_comp=%d outside 1..3", n_comp);
    if (L < 0 || L > 12 || K < 1 || K > PMAXK) return set_error(CFK_EINVAL, "L=%d (0..12) or K=%d (1..16) out of range", L, K);
    if (n_rad < 1 || n_rad > 256 || n_pol < 1 || n_pol > 64 || n_azi < 1 || n_azi > 64)
        return set_error(CFK_EINVAL, "grid %d x %d x %d out of range (256, 64, 64)", n_rad, n_pol, n_azi);
    const int N_Q = (L + 1) * (L + 1), M2 = 2 * L + 1, NA = n_pol * n_azi;
    const size_t lds = sizeof(double) * ((size_t)N_Q * n_pol + (size_t)M2 * n_azi + NA + (size_t)n_pol * M2 + 3 * NA);
    // one workgroup per distribution holds the grid's tables: up to the 160 KB of LDS a gfx950 workgroup
    // can address (the tutorials' 60 x 30 x 30 grid needs 81 KB at L = 12)
    if (lds > 160 * 1024) return set_error(CFK_EINVAL, "grid too large for one workgroup (%zu bytes of LDS)", lds);
    if (N_Q > PT) return set_error(CFK_EINVAL, "N_Q=%d exceeds %d", N_Q, PT);
    int rc;
    if ((rc = check_device())) return rc;
    if (B == 0) return CFK_OK;
    // host tables
    std::vector<double> r, rad, mu(n_pol), wp(n_pol), cphi(n_azi), sphi(n_azi);
    projection_radial_table(L, K, n_rad, r, rad);
    gauss_legendre(n_pol, -1.0, 1.0, mu.data(), wp.data());
    const double wphi = 2.0 * M_PI / n_azi;
    for (int j = 0; j < n_azi; ++j) {
        const double ph = 2.0 * M_PI * j / n_azi;
        cphi[j] = std::cos(ph);
        sphi[j] = std::sin(ph);
    }
    std::vector<double> shn(N_Q);
    sh_norm(L, shn.data());
    std::vector<double> leg((size_t)N_Q * n_pol), trig((size_t)M2 * n_azi);
    for (int q = 0; q < N_Q; ++q) {
        const int l = l_of_q0(q), m = q - l * l - l;
        for (int p = 0; p < n_pol; ++p) leg[(size_t)q * n_pol + p] = shn[q] * legendre_plm(l, std::abs(m), mu[p]) * wp[p];
    }
    for (int j = 0; j < n_azi; ++j) {
        const double ph = 2.0 * M_PI * j / n_azi;
        trig[j] = wphi;
        for (int m = 1; m <= L; ++m) {
            trig[(size_t)(2 * m - 1) * n_azi + j] = std::cos(m * ph) * wphi;
            trig[(size_t)(2 * m) * n_azi + j] = std::sin(m * ph) * wphi;
        }
    }
    std::vector<double> all;
    const size_t o_rad = 0, o_leg = o_rad + rad.size(), o_trig = o_leg + leg.size(), o_r = o_trig + trig.size(),
                 o_mu = o_r + r.size(), o_c = o_mu + mu.size(), o_s = o_c + cphi.size(), n_all = o_s + sphi.size();
    all.reserve(n_all);
    for (const auto *v : {&rad, &leg, &trig, &r, &mu, &cphi, &sphi}) all.insert(all.end(), v->begin(), v->end());
    DevBuf<double> dt;
    if ((rc = dt.upload(all.data(), all.size()))) return rc;
    ProjDev d{dt.p + o_rad, dt.p + o_leg, dt.p + o_trig, dt.p + o_r, dt.p + o_mu, dt.p + o_c, dt.p + o_s,
              L, K, N_Q, n_rad, n_pol, n_azi, n_comp};
    hipStream_t st = (hipStream_t)stream;
    if (lds > 64 * 1024)
        CFK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&project_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(project_kernel, dim3((unsigned)B), dim3(PT), lds, st, d, params_dev, c_dev, B);
    CFK_HIP(hipGetLastError());
    // the tables are freed on return: wait for the kernel that reads them
    CFK_HIP(hipStreamSynchronize(st));
    return CFK_OK;
}
