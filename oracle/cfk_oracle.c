/*
 * ORACLE (test infrastructure only) -- plain-C restatement of the reference's native kernels.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
 * and only as the checker or the CPU baseline.  The product library (libcfk.so) never links it.
 *
 * Each function follows the reference loop order and arithmetic so that, built with
 * -ffp-contract=off, it reproduces the reference MEX bit for bit on the same inputs
 * (single thread).  The reference itself cannot be built here: its sources include MATLAB's
 * "mex.h", which this image lacks.
 *
 *   oracle_naive          src/mex/naive_collision_kernel_mex.cpp:4-47
 *   oracle_angular_first  src/mex/angular_first_collision_kernel_mex.cpp:6-76
 *   oracle_radial_first   src/mex/radial_first_collision_kernel_mex.cpp:5-77
 *   oracle_dense          src/mex/dense_tensor_kernel_mex.cpp:4-31
 *   oracle_sumfac         src/mex/compute_rtensor_sumfac_mex.cpp:80-398
 *
 * Index arrays are 1-based doubles exactly as MATLAB passes them.  All Q outputs accumulate.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* naive_collision_kernel_mex.cpp:23-46: z -> k2 -> k3 -> k1, f2*f3 hoisted */
void oracle_naive(double *Q, const double *f, const double *labels, const double *vals,
                  const double *ic_map, const double *R, int64_t N_G, int N_Q, int K)
{
    int K2 = K * K, K3 = K2 * K;
    for (int64_t z = 0; z < N_G; ++z) {
        int q1 = (int)labels[z + 0 * N_G] - 1;
        int q2 = (int)labels[z + 1 * N_G] - 1;
        int q3 = (int)labels[z + 2 * N_G] - 1;
        int t = (int)ic_map[z] - 1;
        double g = vals[z];
        for (int k2 = 0; k2 < K; ++k2)
            for (int k3 = 0; k3 < K; ++k3) {
                double f2f3 = f[q2 + k2 * N_Q] * f[q3 + k3 * N_Q];
                for (int k1 = 0; k1 < K; ++k1) {
                    double r = R[k1 + k2 * K + k3 * K2 + (int64_t)t * K3];
                    Q[q1 + k1 * N_Q] += g * r * f2f3;
                }
            }
    }
}

/* angular_first_collision_kernel_mex.cpp:24-75: t-runs, q1-runs, Phi (K x K) then R.Phi */
void oracle_angular_first(double *Q, const double *f, const double *labels, const double *vals,
                          const double *ic_map, const double *R, int64_t N_G, int N_Q, int K)
{
    int K2 = K * K, K3 = K2 * K;
    double *Phi = (double *)malloc(sizeof(double) * (size_t)K2);
    int64_t z = 0;
    while (z < N_G) {
        int t = (int)ic_map[z] - 1;
        int64_t z_end_t = z;
        while (z_end_t < N_G && ((int)ic_map[z_end_t] - 1) == t) z_end_t++;
        int64_t zs = z;
        while (zs < z_end_t) {
            int q1 = (int)labels[zs] - 1;
            int64_t ze = zs;
            while (ze < z_end_t && ((int)labels[ze] - 1) == q1) ze++;
            for (int i = 0; i < K2; ++i) Phi[i] = 0.0;
            for (int64_t i = zs; i < ze; ++i) {
                int q2 = (int)labels[i + 1 * N_G] - 1;
                int q3 = (int)labels[i + 2 * N_G] - 1;
                double g = vals[i];
                for (int k3 = 0; k3 < K; ++k3) {
                    double f3 = f[q3 + k3 * N_Q];
                    for (int k2 = 0; k2 < K; ++k2) Phi[k2 + k3 * K] += g * f[q2 + k2 * N_Q] * f3;
                }
            }
            for (int k1 = 0; k1 < K; ++k1) {
                double qv = 0.0;
                for (int k3 = 0; k3 < K; ++k3)
                    for (int k2 = 0; k2 < K; ++k2)
                        qv += R[k1 + k2 * K + k3 * K2 + (int64_t)t * K3] * Phi[k2 + k3 * K];
                Q[q1 + k1 * N_Q] += qv;
            }
            zs = ze;
        }
        z = z_end_t;
    }
    free(Phi);
}

/* radial_first_collision_kernel_mex.cpp:35-76: Psi[z,k1] per t-run, then routed by g */
void oracle_radial_first(double *Q, const double *f, const double *labels, const double *vals,
                         const double *ic_map, const double *R, int64_t N_G, int N_Q, int K)
{
    int K2 = K * K, K3 = K2 * K;
    double *Psi = (double *)calloc((size_t)N_G * K, sizeof(double));
    int64_t z0 = 0;
    while (z0 < N_G) {
        int t = (int)ic_map[z0] - 1;
        int64_t z1 = z0;
        while (z1 < N_G && ((int)ic_map[z1] - 1) == t) z1++;
        int64_t nz = z1 - z0;
        for (int64_t z = 0; z < nz; ++z) {
            int64_t gz = z0 + z;
            int q2 = (int)labels[gz + 1 * N_G] - 1;
            int q3 = (int)labels[gz + 2 * N_G] - 1;
            for (int k1 = 0; k1 < K; ++k1) {
                double psi = 0.0;
                for (int k3 = 0; k3 < K; ++k3) {
                    double f3 = f[q3 + k3 * N_Q];
                    for (int k2 = 0; k2 < K; ++k2) {
                        double f2 = f[q2 + k2 * N_Q];
                        double r = R[k1 + k2 * K + k3 * K2 + (int64_t)t * K3];
                        psi += r * f2 * f3;
                    }
                }
                Psi[z * K + k1] = psi;
            }
        }
        for (int64_t z = 0; z < nz; ++z) {
            int64_t gz = z0 + z;
            int q1 = (int)labels[gz] - 1;
            double g = vals[gz];
            for (int k1 = 0; k1 < K; ++k1) Q[q1 + k1 * N_Q] += g * Psi[z * K + k1];
        }
        z0 = z1;
    }
    free(Psi);
}

/* dense_tensor_kernel_mex.cpp:14-29 */
void oracle_dense(double *Q, const double *f, const double *C, int64_t N)
{
    for (int64_t k = 0; k < N; ++k) {
        double fk = f[k];
        for (int64_t j = 0; j < N; ++j) {
            double fjfk = f[j] * fk;
            const double *c = C + (j + k * N) * N;
            for (int64_t i = 0; i < N; ++i) Q[i] += c[i] * fjfk;
        }
    }
}

/* Batched harness: B independent distributions, f/Q strided by N_Q*K.  Parallel over the batch
 * when nthreads > 1 (harness parallelism; the reference contraction is single-threaded). */
void oracle_angular_first_batch(double *Q, const double *f, const double *labels, const double *vals,
                                const double *ic_map, const double *R, int64_t N_G, int N_Q, int K,
                                int64_t B, int nthreads)
{
    int64_t stride = (int64_t)N_Q * K;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t b = 0; b < B; ++b)
        oracle_angular_first(Q + b * stride, f + b * stride, labels, vals, ic_map, R, N_G, N_Q, K);
}

/* ------------------------------------------------------------------------------------------ */
/* Sum-factorized quadrature (compute_rtensor_sumfac_mex.cpp)                                  */
/* ------------------------------------------------------------------------------------------ */

/* compute_rtensor_sumfac_mex.cpp:14-34 */
static double eval_radial(int n_idx, int l, double v, double norm_factor)
{
    int k = n_idx - 1;
    double v2 = v * v;
    double al = (double)l + 0.5;
    double Lk = 0.0;
    if (k == 0) {
        Lk = 1.0;
    } else if (k == 1) {
        Lk = 1.0 + al - v2;
    } else {
        double L0 = 1.0, L1 = 1.0 + al - v2;
        for (int i = 1; i < k; ++i) {
            Lk = ((2.0 * i + 1.0 + al - v2) * L1 - (i + al) * L0) / (i + 1.0);
            L0 = L1;
            L1 = Lk;
        }
    }
    return norm_factor * Lk * pow(v, l);
}

/* compute_rtensor_sumfac_mex.cpp:36-55 */
static double legendre_P(int l, int m, double x)
{
    double pmm = 1.0;
    if (m > 0) {
        double s = sqrt(fmax((1.0 - x) * (1.0 + x), 0.0));
        double fact = 1.0;
        for (int i = 1; i <= m; ++i) {
            pmm *= -fact * s;
            fact += 2.0;
        }
    }
    if (l == m) return pmm;
    double pmmp1 = x * (2 * m + 1) * pmm;
    if (l == m + 1) return pmmp1;
    double pll = 0.0;
    for (int ll = m + 2; ll <= l; ++ll) {
        pll = (x * (2 * ll - 1) * pmmp1 - (ll + m - 1) * pmm) / (ll - m);
        pmm = pmmp1;
        pmmp1 = pll;
    }
    return pll;
}

/* compute_rtensor_sumfac_mex.cpp:57-75 */
static void eval_SH(double theta, double phi, int L_max, const double *SH_Norm, double *Y)
{
    double x = cos(theta);
    for (int l = 0; l <= L_max; ++l)
        for (int m = 0; m <= l; ++m) {
            double P = legendre_P(l, m, x);
            if (m == 0) {
                int q = l * l + l;
                Y[q] = SH_Norm[q] * P;
            } else {
                int qp = l * l + l + m, qn = l * l + l - m;
                Y[qp] = SH_Norm[qp] * P * cos(m * phi);
                Y[qn] = SH_Norm[qn] * P * sin(m * phi);
            }
        }
}

/* rad[k + N_K*l] = eval_radial(k+1, l, v, RadialNorm[k + N_K*l]) for all k < N_K, l <= L_max.
 * A cache of the exact calls the reference makes repeatedly; results are bit-identical. */
static void radial_table(double v, int N_K, int L_max, const double *RadialNorm, double *rad)
{
    for (int l = 0; l <= L_max; ++l)
        for (int k = 0; k < N_K; ++k) rad[k + N_K * l] = eval_radial(k + 1, l, v, RadialNorm[k + N_K * l]);
}

typedef struct {
    int K_max, N_L, N_Q;
    double alpha;
    const double *x, *W_x; int N_x;
    const double *u1, *W_u1; int N_u1;
    const double *t1, *W_t1; int N_t1;
    const double *y2, *W_y2; int N_y2;
    const double *t2, *W_t2; int N_t2;
    const double *mu_chi, *W_chi; int N_chi;
    const double *eps, *W_eps; int N_eps;
    const double *RadialNorm, *SH_Norm, *L_triplets, *qi_valid_mat;
    const double *P_loss_p1, *P_gain_p1, *P_loss_p2, *P_gain_p2;
} sumfac_args;

/* Gain sweep over (chi, eps) for one (x, outer, t) point: compute_rtensor_sumfac_mex.cpp:177-238
 * (patch 1) and :302-363 (patch 2).  P_gain(q_i, t_chan) is addressed through pg_base/pg_stride. */
static void gain_sweep(const sumfac_args *a, int L_max, double v_large, double v_small, double c_beta,
                       double s_beta, double B_val, double Jac_t, const double *P_gain, int64_t pg_off,
                       int64_t pg_q_stride, int64_t pg_t_stride, double *Y1, double *Y2, double *rv1,
                       double *rv2, double *Phi_gain1, double *Phi_gain2)
{
    const double eps_safe = 1e-15;
    int N_K = a->K_max + 1, N_L = a->N_L, N_Q = a->N_Q;
    double U_x1 = v_small * s_beta, U_z1 = v_large + v_small * c_beta;
    double u_x1 = -v_small * s_beta, u_z1 = v_large - v_small * c_beta;
    double u_mag1 = sqrt(fmax(u_x1 * u_x1 + u_z1 * u_z1, eps_safe));
    double zx1 = u_x1 / u_mag1, zz1 = u_z1 / u_mag1;
    double U_x2 = v_large * s_beta, U_z2 = v_small + v_large * c_beta;
    double u_x2 = -v_large * s_beta, u_z2 = v_small - v_large * c_beta;
    double u_mag2 = sqrt(fmax(u_x2 * u_x2 + u_z2 * u_z2, eps_safe));
    double zx2 = u_x2 / u_mag2, zz2 = u_z2 / u_mag2;
    for (int n = 0; n < a->N_chi; ++n) {
        double c_chi = a->mu_chi[n];
        double s_chi = sqrt(fmax(1.0 - c_chi * c_chi, 0.0));
        double gain_base = a->W_chi[n] * B_val * Jac_t;
        for (int e = 0; e < a->N_eps; ++e) {
            double c_eps = cos(a->eps[e]), s_eps = sin(a->eps[e]);
            double u_px = s_chi * c_eps, u_py = s_chi * s_eps, u_pz = c_chi;
            double up_x1 = u_px * zz1 + u_pz * zx1;
            double up_z1 = -u_px * zx1 + u_pz * zz1;
            double vp_x1 = 0.5 * (U_x1 + u_mag1 * up_x1);
            double vp_y1 = 0.5 * (u_mag1 * u_py);
            double vp_z1 = 0.5 * (U_z1 + u_mag1 * up_z1);
            double vp_mag1 = sqrt(fmax(vp_x1 * vp_x1 + vp_y1 * vp_y1 + vp_z1 * vp_z1, eps_safe));
            eval_SH(acos(fmax(fmin(vp_z1 / vp_mag1, 1.0), -1.0)), atan2(vp_y1, vp_x1), L_max, a->SH_Norm, Y1);
            double up_x2 = u_px * zz2 + u_pz * zx2;
            double up_z2 = -u_px * zx2 + u_pz * zz2;
            double vp_x2 = 0.5 * (U_x2 + u_mag2 * up_x2);
            double vp_y2 = 0.5 * (u_mag2 * u_py);
            double vp_z2 = 0.5 * (U_z2 + u_mag2 * up_z2);
            double vp_mag2 = sqrt(fmax(vp_x2 * vp_x2 + vp_y2 * vp_y2 + vp_z2 * vp_z2, eps_safe));
            eval_SH(acos(fmax(fmin(vp_z2 / vp_mag2, 1.0), -1.0)), atan2(vp_y2, vp_x2), L_max, a->SH_Norm, Y2);
            radial_table(vp_mag1, N_K, L_max, a->RadialNorm, rv1);
            radial_table(vp_mag2, N_K, L_max, a->RadialNorm, rv2);
            for (int tc = 0; tc < N_L; ++tc) {
                int l1 = (int)a->L_triplets[tc];
                double g1 = 0.0, g2 = 0.0;
                for (int idx = 0; idx < N_Q; ++idx) {
                    int qi = (int)a->qi_valid_mat[idx + (int64_t)N_Q * tc];
                    if (qi == -1) break;
                    qi -= 1;
                    double w = P_gain[pg_off + pg_q_stride * qi + pg_t_stride * tc];
                    g1 += Y1[qi] * w;
                    g2 += Y2[qi] * w;
                }
                double wg1 = g1 * gain_base * a->W_eps[e];
                double wg2 = g2 * gain_base * a->W_eps[e];
                for (int k1 = 0; k1 < N_K; ++k1) {
                    Phi_gain1[k1 + N_K * tc] += wg1 * rv1[k1 + N_K * l1];
                    Phi_gain2[k1 + N_K * tc] += wg2 * rv2[k1 + N_K * l1];
                }
            }
        }
    }
}

/* Factorized K^3 contraction into R: compute_rtensor_sumfac_mex.cpp:241-261 / :365-385 */
static void factor_contract(const sumfac_args *a, double Jac, const double *rl, const double *rs,
                            const double *Pg1, const double *Pl1, const double *Pg2, const double *Pl2,
                            double *Rl)
{
    int N_K = a->K_max + 1, N_L = a->N_L;
    for (int tc = 0; tc < N_L; ++tc) {
        int l2 = (int)a->L_triplets[tc + N_L * 1];
        int l3 = (int)a->L_triplets[tc + N_L * 2];
        for (int k1 = 0; k1 < N_K; ++k1) {
            double net1 = Jac * (Pg1[k1 + N_K * tc] - Pl1[k1 + N_K * tc]);
            double net2 = Jac * (Pg2[k1 + N_K * tc] - Pl2[k1 + N_K * tc]);
            for (int k2 = 0; k2 < N_K; ++k2) {
                double r2l = rl[k2 + N_K * l2], r2s = rs[k2 + N_K * l2];
                for (int k3 = 0; k3 < N_K; ++k3) {
                    double r3l = rl[k3 + N_K * l3], r3s = rs[k3 + N_K * l3];
                    int64_t o = k1 + N_K * (k2 + N_K * (k3 + (int64_t)N_K * tc));
                    Rl[o] += net1 * r2l * r3s + net2 * r2s * r3l;
                }
            }
        }
    }
}

/* Loss accumulation: compute_rtensor_sumfac_mex.cpp:167-175 / :292-300 */
static void loss_accum(const sumfac_args *a, double loss_weight, const double *P_loss, int64_t pl_off,
                       int64_t pl_t_stride, const double *rl, const double *rs, double *Pl1, double *Pl2)
{
    int N_K = a->K_max + 1;
    for (int tc = 0; tc < a->N_L; ++tc) {
        int l1 = (int)a->L_triplets[tc];
        double li = loss_weight * P_loss[pl_off + pl_t_stride * tc];
        for (int k1 = 0; k1 < N_K; ++k1) {
            Pl1[k1 + N_K * tc] += rl[k1 + N_K * l1] * li;
            Pl2[k1 + N_K * tc] += rs[k1 + N_K * l1] * li;
        }
    }
}

/* Work buffers of one thread. */
typedef struct {
    double *Pl1, *Pl2, *Pg1, *Pg2, *Y1, *Y2, *rl, *rs, *rv1, *rv2;
    size_t nphi;
} sumfac_ws;

static void ws_alloc(sumfac_ws *w, const sumfac_args *a, int L_max)
{
    int N_K = a->K_max + 1;
    size_t nrad = (size_t)N_K * (L_max + 1);
    w->nphi = (size_t)N_K * a->N_L;
    w->Pl1 = calloc(w->nphi, sizeof(double)); w->Pl2 = calloc(w->nphi, sizeof(double));
    w->Pg1 = calloc(w->nphi, sizeof(double)); w->Pg2 = calloc(w->nphi, sizeof(double));
    w->Y1 = calloc(a->N_Q, sizeof(double)); w->Y2 = calloc(a->N_Q, sizeof(double));
    w->rl = calloc(nrad, sizeof(double)); w->rs = calloc(nrad, sizeof(double));
    w->rv1 = calloc(nrad, sizeof(double)); w->rv2 = calloc(nrad, sizeof(double));
}

static void ws_free(sumfac_ws *w)
{
    free(w->Pl1); free(w->Pl2); free(w->Pg1); free(w->Pg2); free(w->Y1); free(w->Y2);
    free(w->rl); free(w->rs); free(w->rv1); free(w->rv2);
}

/* Patch-1 unit (x node i, u node ui): compute_rtensor_sumfac_mex.cpp:141-261. */
static void patch1_unit(const sumfac_args *a, int L_max, int i, int ui, sumfac_ws *w, double *Rl)
{
    const double eps_safe = 1e-15;
    int N_K = a->K_max + 1, N_t1 = a->N_t1, N_u1 = a->N_u1, N_Q = a->N_Q;
    double x_i = a->x[i];
    double u = a->u1[ui];
    double v_large = sqrt(x_i / 2.0) * (1.0 + u);
    double v_small = sqrt(x_i / 2.0) * (1.0 - u);
    double Jac_u = 0.5 * a->W_x[i] * a->W_u1[ui] * u * exp(-x_i * u * u) * (v_large * v_large * v_small * v_small);
    memset(w->Pl1, 0, w->nphi * sizeof(double));
    memset(w->Pl2, 0, w->nphi * sizeof(double));
    memset(w->Pg1, 0, w->nphi * sizeof(double));
    memset(w->Pg2, 0, w->nphi * sizeof(double));
    radial_table(v_large, N_K, L_max, a->RadialNorm, w->rl);
    radial_table(v_small, N_K, L_max, a->RadialNorm, w->rs);
    for (int ti = 0; ti < N_t1; ++ti) {
        double t = a->t1[ti];
        double y = u * t;
        double Jac_t = a->W_t1[ti] * (4.0 * u * t);
        double c_beta = 1.0 - 2.0 * y * y;
        double s_beta = sqrt(fmax(1.0 - c_beta * c_beta, eps_safe));
        double u_mag = sqrt(2.0 * x_i) * u * sqrt(1.0 + (1.0 - u * u) * t * t);
        double B_val = pow(u_mag * sqrt(2.0), a->alpha) / pow(x_i, a->alpha / 2.0);
        double loss_weight = B_val * 2.0 * M_PI * 2.0 * Jac_t;
        loss_accum(a, loss_weight, a->P_loss_p1, ti + (int64_t)N_t1 * ui, (int64_t)N_t1 * N_u1, w->rl, w->rs, w->Pl1,
                   w->Pl2);
        gain_sweep(a, L_max, v_large, v_small, c_beta, s_beta, B_val, Jac_t, a->P_gain_p1, ti + (int64_t)N_t1 * ui,
                   (int64_t)N_t1 * N_u1, (int64_t)N_t1 * N_u1 * N_Q, w->Y1, w->Y2, w->rv1, w->rv2, w->Pg1, w->Pg2);
    }
    factor_contract(a, Jac_u, w->rl, w->rs, w->Pg1, w->Pl1, w->Pg2, w->Pl2, Rl);
}

/* Patch-2 row (x node i, y node yi, all inner t): compute_rtensor_sumfac_mex.cpp:267-386. */
static void patch2_row(const sumfac_args *a, int L_max, int i, int yi, sumfac_ws *w, double *Rl)
{
    const double eps_safe = 1e-15;
    int N_K = a->K_max + 1, N_y2 = a->N_y2, N_Q = a->N_Q;
    double x_i = a->x[i];
    double y = a->y2[yi];
    double Jac_y = 0.5 * a->W_x[i] * a->W_y2[yi] * y * (4.0 * y);
    double c_beta = 1.0 - 2.0 * y * y;
    double s_beta = sqrt(fmax(1.0 - c_beta * c_beta, eps_safe));
    for (int ti = 0; ti < a->N_t2; ++ti) {
        double t = a->t2[ti];
        double u = y * t;
        double v_large = sqrt(x_i / 2.0) * (1.0 + u);
        double v_small = sqrt(x_i / 2.0) * (1.0 - u);
        double Jac_t = a->W_t2[ti] * exp(-x_i * u * u) * (v_large * v_large * v_small * v_small);
        double u_mag = sqrt(2.0 * x_i) * y * sqrt(t * t + 1.0 - y * y * t * t);
        double B_val = pow(u_mag * sqrt(2.0), a->alpha) / pow(x_i, a->alpha / 2.0);
        double loss_weight = B_val * 2.0 * M_PI * 2.0 * Jac_t;
        memset(w->Pl1, 0, w->nphi * sizeof(double));
        memset(w->Pl2, 0, w->nphi * sizeof(double));
        memset(w->Pg1, 0, w->nphi * sizeof(double));
        memset(w->Pg2, 0, w->nphi * sizeof(double));
        radial_table(v_large, N_K, L_max, a->RadialNorm, w->rl);
        radial_table(v_small, N_K, L_max, a->RadialNorm, w->rs);
        loss_accum(a, loss_weight, a->P_loss_p2, yi, N_y2, w->rl, w->rs, w->Pl1, w->Pl2);
        gain_sweep(a, L_max, v_large, v_small, c_beta, s_beta, B_val, Jac_t, a->P_gain_p2, yi, N_y2,
                   (int64_t)N_y2 * N_Q, w->Y1, w->Y2, w->rv1, w->rv2, w->Pg1, w->Pg2);
        factor_contract(a, Jac_y, w->rl, w->rs, w->Pg1, w->Pl1, w->Pg2, w->Pl2, Rl);
    }
}

/* Contribution of work units [w0, w1) accumulated into R_local, where the units of x node i are, in
 * the reference's order, its N_u1 patch-1 units followed by its N_y2 patch-2 rows. */
static void sumfac_units(const sumfac_args *a, int x_begin, int64_t w0, int64_t w1, double *Rl)
{
    int L_max = (int)lround(sqrt((double)a->N_Q)) - 1;
    int per_x = a->N_u1 + a->N_y2;
    sumfac_ws w;
    ws_alloc(&w, a, L_max);
    for (int64_t wu = w0; wu < w1; ++wu) {
        int i = x_begin + (int)(wu / per_x);
        int r = (int)(wu % per_x);
        if (r < a->N_u1)
            patch1_unit(a, L_max, i, r, &w, Rl);
        else
            patch2_row(a, L_max, i, r - a->N_u1, &w, Rl);
    }
    ws_free(&w);
}

/* R_out[N_K,N_K,N_K,N_L] += contribution of x nodes [x_begin, x_end).  The work units (per x node:
 * patch-1 u units, then patch-2 y rows) are split into nthreads contiguous blocks; each thread sums
 * its block into a partial tensor in the reference's order and the partials are added in thread order
 * (deterministic for a given thread count; with one thread the order is the reference's single-thread
 * order).  The reference splits over x nodes only and merges under an OpenMP critical section in
 * arrival order (compute_rtensor_sumfac_mex.cpp:126-134, :391-396). */
void oracle_sumfac(double *R_out, int K_max, int N_L, int N_Q, double alpha,
                   const double *x, const double *W_x, int N_x,
                   const double *u1, const double *W_u1, int N_u1,
                   const double *t1, const double *W_t1, int N_t1,
                   const double *y2, const double *W_y2, int N_y2,
                   const double *t2, const double *W_t2, int N_t2,
                   const double *mu_chi, const double *W_chi, int N_chi,
                   const double *eps, const double *W_eps, int N_eps,
                   const double *RadialNorm, const double *SH_Norm, const double *L_triplets,
                   const double *qi_valid_mat, const double *P_loss_p1, const double *P_gain_p1,
                   const double *P_loss_p2, const double *P_gain_p2, int x_begin, int x_end, int nthreads)
{
    sumfac_args a = {K_max, N_L, N_Q, alpha, x, W_x, N_x, u1, W_u1, N_u1, t1, W_t1, N_t1,
                     y2, W_y2, N_y2, t2, W_t2, N_t2, mu_chi, W_chi, N_chi, eps, W_eps, N_eps,
                     RadialNorm, SH_Norm, L_triplets, qi_valid_mat, P_loss_p1, P_gain_p1, P_loss_p2, P_gain_p2};
    int N_K = K_max + 1;
    size_t nR = (size_t)N_K * N_K * N_K * N_L;
    if (x_begin < 0) x_begin = 0;
    if (x_end > N_x) x_end = N_x;
    if (nthreads < 1) nthreads = 1;
    int nx = x_end - x_begin;
    if (nx <= 0) return;
    int64_t nunits = (int64_t)nx * (N_u1 + N_y2);
    if (nthreads > nunits) nthreads = (int)nunits;
    double *parts = calloc(nR * (size_t)nthreads, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
#endif
    for (int th = 0; th < nthreads; ++th) {
        int64_t c0 = nunits * th / nthreads;
        int64_t c1 = nunits * (th + 1) / nthreads;
        sumfac_units(&a, x_begin, c0, c1, parts + nR * th);
    }
    for (int th = 0; th < nthreads; ++th)
        for (size_t i = 0; i < nR; ++i) R_out[i] += parts[nR * th + i];
    free(parts);
}
