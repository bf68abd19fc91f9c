// 5-D sum-factorized Duffy quadrature on the GPU (compute_rtensor_sumfac_mex.cpp:80-398).
//
// The reference sums, per x node, over patch-1 units (x,u) and patch-2 units (x,y,t); each unit
// builds Phi_gain/Phi_loss[k1,t_chan] from its gain points and then adds a factorized K^3 block to
// R (compute_rtensor_sumfac_mex.cpp:241-261, :365-385).  Here:
//
//   stage A (sumfac_units_kernel): one workgroup per unit, one lane per t_chan.  Gain points are
//     processed in batches of NPB; for each batch the workgroup evaluates the real SH at both
//     post-collision directions (one lane per order m) and the radial functions at both speeds
//     (one lane per degree l) into LDS, then every t_chan lane does its sparse dot over its valid
//     q_i and updates its Phi accumulators in registers, in the reference's gain-point order.
//     The unit's net vectors Jac*(Phi_gain - Phi_loss) and its radial tables at v_large/v_small are
//     written to HBM.
//   stage B (sumfac_fold_kernel): one thread per R element sums the units' rank-2 contributions
//     net1 r2(v_large) r3(v_small) + net2 r2(v_small) r3(v_large) in the reference's unit order
//     (x, then patch 1 over u, then patch 2 over (y,t)), so the result is deterministic.
//
// All arithmetic is FP64.  The P_gain tables are compacted to the valid q_i of each t_chan (the
// reference stores them dense over all q and reads only the valid ones,
// compute_rtensor_sumfac_mex.cpp:219-227).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "prep.hpp"

namespace cfk {

namespace {

constexpr int SF_THREADS = 256;
constexpr int NPB = 4;        // gain points per LDS batch
constexpr int MAXL = 12;      // L <= 12 (N_Q <= 169)
constexpr int MAXK = 16;      // N_K <= 16
constexpr int MAXSLOT = 2 * MAXL + 1;  // valid q_i per t_chan (2 l1 + 1)
constexpr int MAXG = 128;     // N_chi, N_eps <= 128
constexpr double EPS_SAFE = 1e-15;

struct SumfacDev {
    int NK, NL, NQ, L;
    double alpha;
    int Nx, Nu1, Nt1, Ny2, Nt2, Nchi, Neps;
    const double *x, *Wx, *u1, *Wu1, *t1, *Wt1, *y2, *Wy2, *t2, *Wt2, *mu_chi, *W_chi, *eps, *W_eps;
    const double *radnorm;  // [l][k] -> radnorm[k + NK*l]
    const double *shnorm;   // [NQ]
    const int *trip;        // [NL][3]
    const int *qi_off;      // [NL+1]
    const int *qi;          // [nslot] 0-based q
    const double *Pg1;      // [Nu1][Nt1][nslot]
    const double *Pl1;      // [Nu1][Nt1][NL]
    const double *Pg2;      // [Ny2][nslot]
    const double *Pl2;      // [Ny2][NL]
    int max_block_slots;    // max over t_chan blocks of their number of valid q_i
};

// eval_radial for all k at one (l, v): out[k] = norm[k,l] * L_k^{(l+1/2)}(v^2) * v^l
// (compute_rtensor_sumfac_mex.cpp:14-34, same recursion and operation order).
template <int NK>
__device__ inline void radial_all_k(int l, double v, const double *radnorm, double *out) {
    double vl = 1.0;  // v^l by repeated multiplication (the reference calls pow(v, l))
    for (int i = 0; i < l; ++i) vl *= v;
    const double v2 = v * v, al = (double)l + 0.5;
    double L0 = 1.0, L1 = 1.0 + al - v2;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        double Lk;
        if (k == 0)
            Lk = 1.0;
        else if (k == 1)
            Lk = L1;
        else {
            Lk = ((2.0 * (k - 1) + 1.0 + al - v2) * L1 - ((k - 1) + al) * L0) / ((k - 1) + 1.0);
            L0 = L1;
            L1 = Lk;
        }
        out[k] = radnorm[k + NK * l] * Lk * vl;
    }
}

// Real SH of one order m >= 0 for all degrees l = m..L (eval_SH / legendre_P,
// compute_rtensor_sumfac_mex.cpp:36-75): Y[l^2+l+m] = N P cos(m phi), Y[l^2+l-m] = N P sin(m phi).
// x = cos(theta); (cphi, sphi) = (cos phi, sin phi); cos/sin(m phi) come from the m-th complex power.
__device__ inline void sh_order_m(int L, int m, double x, double cphi, double sphi, const double *shnorm, double *Y) {
    double pmm = 1.0;
    if (m > 0) {
        const double s = sqrt(fmax((1.0 - x) * (1.0 + x), 0.0));
        double fact = 1.0;
        for (int i = 1; i <= m; ++i) {
            pmm *= -fact * s;
            fact += 2.0;
        }
    }
    double cm = 1.0, sm = 0.0;
    for (int i = 0; i < m; ++i) {
        const double c = cm * cphi - sm * sphi;
        sm = sm * cphi + cm * sphi;
        cm = c;
    }
    double p_lm2 = 0.0, p_lm1 = 0.0;
    for (int l = m; l <= L; ++l) {
        double P;
        if (l == m)
            P = pmm;
        else if (l == m + 1)
            P = x * (2 * m + 1) * pmm;
        else
            P = (x * (2 * l - 1) * p_lm1 - (l + m - 1) * p_lm2) / (l - m);
        p_lm2 = p_lm1;
        p_lm1 = P;
        if (m == 0) {
            const int q = l * l + l;
            Y[q] = shnorm[q] * P;
        } else {
            const int qp = l * l + l + m, qn = l * l + l - m;
            Y[qp] = shnorm[qp] * P * cm;
            Y[qn] = shnorm[qn] * P * sm;
        }
    }
}

// Per-(unit, inner t) geometry shared by all gain points of that t.
struct TGeom {
    double v_large, v_small, B_val, Jac_t, loss_weight;
    double U_x1, U_z1, u_mag1, zx1, zz1;
    double U_x2, U_z2, u_mag2, zx2, zz2;
};

// Post-collision speed and direction (compute_rtensor_sumfac_mex.cpp:199-213).  The reference
// forms theta = acos(clamp(vz/|v|)) and phi = atan2(vy, vx) and evaluates cos(theta), cos(m phi),
// sin(m phi); here cos(theta) = clamp(vz/|v|) and (cos phi, sin phi) = (vx, vy)/|(vx, vy)| directly
// (equal up to rounding; atan2(0,0) = 0 gives (1, 0)).
__device__ inline void post_collision(const TGeom &g, double c_chi, double s_chi, double c_eps, double s_eps, int side,
                                      double &vp_mag, double &ctheta, double &cphi, double &sphi) {
    const double u_px = s_chi * c_eps, u_py = s_chi * s_eps, u_pz = c_chi;
    const double zx = side ? g.zx2 : g.zx1, zz = side ? g.zz2 : g.zz1;
    const double Ux = side ? g.U_x2 : g.U_x1, Uz = side ? g.U_z2 : g.U_z1, um = side ? g.u_mag2 : g.u_mag1;
    const double up_x = u_px * zz + u_pz * zx;
    const double up_z = -u_px * zx + u_pz * zz;
    const double vx = 0.5 * (Ux + um * up_x);
    const double vy = 0.5 * (um * u_py);
    const double vz = 0.5 * (Uz + um * up_z);
    vp_mag = sqrt(fmax(vx * vx + vy * vy + vz * vz, EPS_SAFE));
    ctheta = fmax(fmin(vz / vp_mag, 1.0), -1.0);
    const double rho = sqrt(vx * vx + vy * vy);
    if (rho > 0.0) {
        cphi = vx / rho;
        sphi = vy / rho;
    } else {
        cphi = 1.0;
        sphi = 0.0;
    }
}

__device__ inline void make_tgeom(TGeom &g, double x_i, double v_large, double v_small, double c_beta, double s_beta,
                                  double u_mag, double Jac_t, double alpha) {
    g.v_large = v_large;
    g.v_small = v_small;
    g.B_val = pow(u_mag * sqrt(2.0), alpha) / pow(x_i, alpha / 2.0);
    g.Jac_t = Jac_t;
    g.loss_weight = g.B_val * 2.0 * M_PI * 2.0 * Jac_t;
    g.U_x1 = v_small * s_beta;
    g.U_z1 = v_large + v_small * c_beta;
    const double ux1 = -v_small * s_beta, uz1 = v_large - v_small * c_beta;
    g.u_mag1 = sqrt(fmax(ux1 * ux1 + uz1 * uz1, EPS_SAFE));
    g.zx1 = ux1 / g.u_mag1;
    g.zz1 = uz1 / g.u_mag1;
    g.U_x2 = v_large * s_beta;
    g.U_z2 = v_small + v_large * c_beta;
    const double ux2 = -v_large * s_beta, uz2 = v_small - v_large * c_beta;
    g.u_mag2 = sqrt(fmax(ux2 * ux2 + uz2 * uz2, EPS_SAFE));
    g.zx2 = ux2 / g.u_mag2;
    g.zz2 = uz2 / g.u_mag2;
}

// Stage A.  Unit id u in [0, n_units): x = x0 + u / Ux; r = u % Ux; r < Nu1 -> patch 1 (u_idx=r),
// else patch 2 (y_idx, t_idx) = divmod(r - Nu1, Nt2).  Grid: (n_units, ceil(NL / SF_THREADS)); the
// block's lanes are the t_chans [tc0, tc0 + SF_THREADS).
// Outputs: net[(unit*NL + tc)*2*NK + side*NK + k1], rad[unit][side][l][k] (side 0 = v_large).
//
// Per inner-t step the block stages its t_chans' P_gain slots in LDS (one coalesced load; they do not
// depend on the gain point).  Per batch of NPB gain points: (1) one thread per (point, side) computes
// the post-collision speed and direction; (2) SH tasks (one per (point, side, order m)) and radial
// tasks (one per (point, side, degree l)) fill LDS tables, each kind in its own half of the
// workgroup so no wave mixes the two; (3) every t_chan lane does its sparse dot over its slots and
// updates its Phi accumulators (registers) in the reference's gain-point order.
template <int NK>
__global__ __launch_bounds__(SF_THREADS) void sumfac_units_kernel(SumfacDev d, int x0, int64_t n_units,
                                                                  double *__restrict__ net, double *__restrict__ rad) {
    __shared__ double sY[NPB][2][(MAXL + 1) * (MAXL + 1)];
    __shared__ double sR[NPB][2][(MAXL + 1) * MAXK];
    __shared__ double sGeo[NPB][2][4];      // vp_mag, cos(theta), cos(phi), sin(phi)
    __shared__ double sW[NPB];              // gain_base per point
    __shared__ double sRadU[2][(MAXL + 1) * MAXK];  // radial tables at v_large / v_small
    __shared__ double sCE[MAXG], sSE[MAXG], sCC[MAXG], sSC[MAXG];
    __shared__ TGeom sG;
    extern __shared__ double sDyn[];        // [nsb] P_gain slots of this block, then [nsb] int q_i

    const int64_t unit = blockIdx.x;
    const int Ux = d.Nu1 + d.Ny2 * d.Nt2;
    const int xi = x0 + (int)(unit / Ux);
    const int r = (int)(unit % Ux);
    const bool patch1 = r < d.Nu1;
    const int tc0 = blockIdx.y * SF_THREADS;
    const int tc1 = min(d.NL, tc0 + SF_THREADS);
    const int tc = tc0 + threadIdx.x;
    const bool own = tc < d.NL;
    const int L = d.L;
    const int ns_all = d.qi_off[d.NL];
    const int sb0 = d.qi_off[tc0], nsb = d.qi_off[tc1] - sb0;
    double *sPv = sDyn;
    int *sQi = reinterpret_cast<int *>(sDyn + d.max_block_slots);
    const double x_i = d.x[xi];

    for (int i = threadIdx.x; i < d.Neps; i += SF_THREADS) {
        sCE[i] = cos(d.eps[i]);
        sSE[i] = sin(d.eps[i]);
    }
    for (int i = threadIdx.x; i < d.Nchi; i += SF_THREADS) {
        const double c = d.mu_chi[i];
        sCC[i] = c;
        sSC[i] = sqrt(fmax(1.0 - c * c, 0.0));
    }
    for (int i = threadIdx.x; i < nsb; i += SF_THREADS) sQi[i] = d.qi[sb0 + i];

    int l1 = 0, s_lo = 0, s_hi = 0;
    if (own) {
        l1 = d.trip[3 * tc];
        s_lo = d.qi_off[tc] - sb0;
        s_hi = d.qi_off[tc + 1] - sb0;
    }
    double Pg1[NK], Pg2[NK], Pl1[NK], Pl2[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) Pg1[k] = Pg2[k] = Pl1[k] = Pl2[k] = 0.0;

    // unit-level quantities
    double v_large, v_small, Jac_unit;
    int u_idx = 0, y_idx = 0, t2_idx = 0;
    if (patch1) {
        u_idx = r;
        const double u = d.u1[u_idx];
        v_large = sqrt(x_i / 2.0) * (1.0 + u);
        v_small = sqrt(x_i / 2.0) * (1.0 - u);
        Jac_unit = 0.5 * d.Wx[xi] * d.Wu1[u_idx] * u * exp(-x_i * u * u) * (v_large * v_large * v_small * v_small);
    } else {
        y_idx = (r - d.Nu1) / d.Nt2;
        t2_idx = (r - d.Nu1) % d.Nt2;
        const double y = d.y2[y_idx];
        Jac_unit = 0.5 * d.Wx[xi] * d.Wy2[y_idx] * y * (4.0 * y);
        const double u = y * d.t2[t2_idx];
        v_large = sqrt(x_i / 2.0) * (1.0 + u);
        v_small = sqrt(x_i / 2.0) * (1.0 - u);
    }
    // radial tables at v_large / v_small: one lane per (side, l)
    if (threadIdx.x < 2 * (L + 1)) {
        const int side = threadIdx.x / (L + 1), l = threadIdx.x % (L + 1);
        double tmp[NK];
        radial_all_k<NK>(l, side ? v_small : v_large, d.radnorm, tmp);
#pragma unroll
        for (int k = 0; k < NK; ++k) sRadU[side][l * MAXK + k] = tmp[k];
    }
    const int n_inner_t = patch1 ? d.Nt1 : 1;
    const int n_gain = d.Nchi * d.Neps;
    constexpr int HALF = SF_THREADS / 2;   // SH tasks in threads [0, HALF), radial tasks in [HALF, 2 HALF)
    for (int it = 0; it < n_inner_t; ++it) {
        __syncthreads();  // sG / sPv reuse
        if (threadIdx.x == 0) {
            if (patch1) {
                const double u = d.u1[u_idx], t = d.t1[it];
                const double y = u * t;
                const double Jac_t = d.Wt1[it] * (4.0 * u * t);
                const double c_beta = 1.0 - 2.0 * y * y;
                const double s_beta = sqrt(fmax(1.0 - c_beta * c_beta, EPS_SAFE));
                const double u_mag = sqrt(2.0 * x_i) * u * sqrt(1.0 + (1.0 - u * u) * t * t);
                make_tgeom(sG, x_i, v_large, v_small, c_beta, s_beta, u_mag, Jac_t, d.alpha);
            } else {
                const double y = d.y2[y_idx], t = d.t2[t2_idx];
                const double u = y * t;
                const double c_beta = 1.0 - 2.0 * y * y;
                const double s_beta = sqrt(fmax(1.0 - c_beta * c_beta, EPS_SAFE));
                const double Jac_t = d.Wt2[t2_idx] * exp(-x_i * u * u) * (v_large * v_large * v_small * v_small);
                const double u_mag = sqrt(2.0 * x_i) * y * sqrt(t * t + 1.0 - y * y * t * t);
                make_tgeom(sG, x_i, v_large, v_small, c_beta, s_beta, u_mag, Jac_t, d.alpha);
            }
        }
        {
            const double *src = patch1 ? d.Pg1 + ((int64_t)u_idx * d.Nt1 + it) * ns_all + sb0
                                       : d.Pg2 + (int64_t)y_idx * ns_all + sb0;
            for (int i = threadIdx.x; i < nsb; i += SF_THREADS) sPv[i] = src[i];
        }
        __syncthreads();
        // loss (compute_rtensor_sumfac_mex.cpp:167-175 / :292-300)
        if (own) {
            const double P = patch1 ? d.Pl1[((int64_t)u_idx * d.Nt1 + it) * d.NL + tc] : d.Pl2[(int64_t)y_idx * d.NL + tc];
            const double li = sG.loss_weight * P;
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                Pl1[k] += sRadU[0][l1 * MAXK + k] * li;
                Pl2[k] += sRadU[1][l1 * MAXK + k] * li;
            }
        }
        // gain points in batches
        for (int p0 = 0; p0 < n_gain; p0 += NPB) {
            const int np = min(NPB, n_gain - p0);
            __syncthreads();
            if (threadIdx.x < 2 * np) {
                const int j = threadIdx.x >> 1, side = threadIdx.x & 1;
                const int p = p0 + j;
                const int n = p / d.Neps, e = p % d.Neps;
                double vp_mag, ct, cp, sp;
                post_collision(sG, sCC[n], sSC[n], sCE[e], sSE[e], side, vp_mag, ct, cp, sp);
                sGeo[j][side][0] = vp_mag;
                sGeo[j][side][1] = ct;
                sGeo[j][side][2] = cp;
                sGeo[j][side][3] = sp;
                if (side == 0) sW[j] = d.W_chi[n] * sG.B_val * sG.Jac_t;
            }
            __syncthreads();
            {
                const int kind = threadIdx.x / HALF, idx = threadIdx.x % HALF;
                const int j = idx / (2 * (L + 1)), rem = idx % (2 * (L + 1));
                const int side = rem / (L + 1), ml = rem % (L + 1);
                if (j < np) {
                    if (kind == 0) {
                        sh_order_m(L, ml, sGeo[j][side][1], sGeo[j][side][2], sGeo[j][side][3], d.shnorm, sY[j][side]);
                    } else {
                        double tmp[NK];
                        radial_all_k<NK>(ml, sGeo[j][side][0], d.radnorm, tmp);
#pragma unroll
                        for (int k = 0; k < NK; ++k) sR[j][side][ml * MAXK + k] = tmp[k];
                    }
                }
            }
            __syncthreads();
            if (own) {
                for (int j = 0; j < np; ++j) {
                    const int p = p0 + j;
                    const int e = p % d.Neps;
                    double g1 = 0.0, g2 = 0.0;
                    for (int s = s_lo; s < s_hi; ++s) {
                        const int q = sQi[s];
                        const double w = sPv[s];
                        g1 += sY[j][0][q] * w;
                        g2 += sY[j][1][q] * w;
                    }
                    const double wg1 = g1 * sW[j] * d.W_eps[e];
                    const double wg2 = g2 * sW[j] * d.W_eps[e];
#pragma unroll
                    for (int k = 0; k < NK; ++k) {
                        Pg1[k] += wg1 * sR[j][0][l1 * MAXK + k];
                        Pg2[k] += wg2 * sR[j][1][l1 * MAXK + k];
                    }
                }
            }
        }
    }
    if (own) {
        double *o = net + (unit * d.NL + tc) * 2 * NK;
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            o[k] = Jac_unit * (Pg1[k] - Pl1[k]);
            o[NK + k] = Jac_unit * (Pg2[k] - Pl2[k]);
        }
    }
    if (blockIdx.y == 0) {
        double *ro = rad + unit * 2 * (L + 1) * NK;
        for (int i = threadIdx.x; i < 2 * (L + 1) * NK; i += SF_THREADS) {
            const int side = i / ((L + 1) * NK), rem = i % ((L + 1) * NK);
            ro[i] = sRadU[side][(rem / NK) * MAXK + rem % NK];
        }
    }
}

// Stage B: R[k1 + NK*(k2 + NK*(k3 + NK*tc))] += sum_u net1 r2(vl) r3(vs) + net2 r2(vs) r3(vl)
__global__ __launch_bounds__(256) void sumfac_fold_kernel(int NK, int NL, int L, const int *__restrict__ trip,
                                                          int64_t n_units, const double *__restrict__ net,
                                                          const double *__restrict__ rad, double *__restrict__ R) {
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = (int64_t)NK * NK * NK * NL;
    if (o >= total) return;
    const int k1 = (int)(o % NK);
    const int k2 = (int)((o / NK) % NK);
    const int k3 = (int)((o / ((int64_t)NK * NK)) % NK);
    const int tc = (int)(o / ((int64_t)NK * NK * NK));
    const int l2 = trip[3 * tc + 1], l3 = trip[3 * tc + 2];
    const int rstride = 2 * (L + 1) * NK;
    double acc = R[o];
    for (int64_t u = 0; u < n_units; ++u) {
        const double *nu = net + (u * NL + tc) * 2 * NK;
        const double *ru = rad + u * rstride;
        const double r2l = ru[l2 * NK + k2], r2s = ru[(L + 1) * NK + l2 * NK + k2];
        const double r3l = ru[l3 * NK + k3], r3s = ru[(L + 1) * NK + l3 * NK + k3];
        acc += nu[k1] * r2l * r3s + nu[NK + k1] * r2s * r3l;
    }
    R[o] = acc;
}

// ---------------------------------------------------------------------------------------------
// Matrix-core variant (the default).
//
// The valid q_i of a t_chan are the q1 labels of its Gaunt entries, all of degree l1 = trip[t][0]
// (src/GeneralCollisionTensor.m:130-160), so the reference's per-t_chan gain sum over the points
// p of one inner-t step (compute_rtensor_sumfac_mex.cpp:219-234)
//     Phi_gain[k][tc] += sum_p w(p) Rad(p, l1, k) sum_{s in slots(tc)} Y(p, q_s) P(s)
// reassociates into
//     M(l, k, q) = sum_p w(p) Rad(p, l, k) Y(p, q)          for every degree l and l(q) = l
//     Phi_gain[k][tc] += sum_{s in slots(tc)} P(s) M(l1, k, q_s).
// The p-sum is done once per degree instead of once per t_chan (about N_L / (L+1) = 47 times fewer
// operations at L = 12), and it is a dense product: for each side and degree l, the NK x (2l+1)
// block M_l = A_l B_l with A_l[k][p] = w(p) Rad(p, l, k), B_l[p][m] = Y(p, l^2 + m), inner length
// the N_chi N_eps gain points.  Each block is tiled 16 x 16 onto v_mfma_f64_16x16x4_f64 (rows k,
// columns m), FP64 operands and accumulation.  The slot contraction, the loss term (whose radial
// factor is fixed per unit) and the unit Jacobian follow per unit.  Sums run in a different order
// than the reference's; the results agree to rounding.
// ---------------------------------------------------------------------------------------------

constexpr int MF_THREADS = 512;         // 8 waves; two workgroups per CU give 4 waves per SIMD
constexpr int MF_WAVES = MF_THREADS / 64;
constexpr int MF_NPB = 12;              // gain points per batch (a multiple of 4, the MFMA inner length)
constexpr int MF_ITEMS = 2 * MF_NPB;    // (point, side) items per batch; item = side * MF_NPB + point
constexpr int MF_PAD = 32;              // operand reads run up to 31 doubles past the last item's row
                                        // (they land in unused rows/columns of the products)
constexpr int MF_SH_LANES = 192;        // evaluation lanes: SH tasks in [0, 192), radial tasks in
constexpr int MF_GEO_LANE = 384;        // [192, 384), the next batch's geometry from lane 384 on

// Operand row stride in doubles: n rounded up to odd.  The evaluation tasks store with one lane
// per item, so an odd stride puts 32 consecutive items on distinct LDS bank pairs.
__host__ __device__ inline int mf_stride(int n) { return n | 1; }
typedef double v4d __attribute__((ext_vector_type(4)));

// 1/i for the three-term recursions (exactly rounded constants; the products differ from the
// reference's divisions by at most one rounding per step)
__constant__ double c_rinv[32] = {0.0,        1.0,        1.0 / 2,  1.0 / 3,  1.0 / 4,  1.0 / 5,  1.0 / 6,
                                  1.0 / 7,    1.0 / 8,    1.0 / 9,  1.0 / 10, 1.0 / 11, 1.0 / 12, 1.0 / 13,
                                  1.0 / 14,   1.0 / 15,   1.0 / 16, 1.0 / 17, 1.0 / 18, 1.0 / 19, 1.0 / 20,
                                  1.0 / 21,   1.0 / 22,   1.0 / 23, 1.0 / 24, 1.0 / 25, 1.0 / 26, 1.0 / 27,
                                  1.0 / 28,   1.0 / 29,   1.0 / 30, 1.0 / 31};

// LDS layout of the matrix-core unit kernel, in doubles.  After a unit's point loop the M blocks
// [2][N_Q][NK] reuse [0, geo).
struct MfLayout {
    int y, a, geo, loss, trig, wts, norm, total;
};

// 16 x 16 tiles per wave for degrees up to L (2 sides, ceil((2l+1)/16) column blocks per degree,
// 8 waves): 3 up to L = 9, 5 up to L = 12.
inline int mf_tiles_per_wave(int L) {
    int n_side = 0;
    for (int l = 0; l <= L; ++l) n_side += (2 * l + 16) / 16;
    const int per_wave = (2 * n_side + MF_WAVES - 1) / MF_WAVES;
    return per_wave <= 3 ? 3 : 5;
}

__host__ __device__ inline MfLayout mf_layout(int L, int NK, int Nt1, int Nchi, int Neps) {
    const int NQ = (L + 1) * (L + 1);
    MfLayout m;
    m.y = 0;
    m.a = MF_ITEMS * mf_stride(NQ) + MF_PAD;
    m.geo = m.a + MF_ITEMS * mf_stride((L + 1) * NK) + MF_PAD;   // two batches of item geometry
    m.loss = m.geo + 2 * MF_ITEMS * 5;                // loss_weight per inner t
    m.trig = m.loss + Nt1;                            // cos/sin eps, mu/sin chi
    m.wts = m.trig + 2 * Neps + 2 * Nchi;             // W_chi, W_eps
    m.norm = m.wts + Nchi + Neps;                     // SH_Norm [N_Q], RadialNorm [(L+1) NK], 1/i [32],
    m.total = m.norm + NQ + (L + 1) * NK + 32 + 32;  // (2i-1)!! [32]
    return m;
}

// Unnormalized real SH of order m >= 0 for degrees m..L into Y: P_l^m(x) cos(m phi) at
// l^2+l+m and P_l^m(x) sin(m phi) at l^2+l-m (eval_SH / legendre_P,
// compute_rtensor_sumfac_mex.cpp:36-75, with SH_Norm applied later to the summed blocks and the
// recursion divisions taken as products with 1/(l-m) from rinv).
__device__ inline void sh_order_m_raw(int L, int m, double x, double cphi, double sphi, const double *rinv,
                                      double *Y) {
    double pmm = 1.0;
    if (m > 0) {
        const double s = sqrt(fmax((1.0 - x) * (1.0 + x), 0.0));
        double fact = 1.0;
        for (int i = 1; i <= m; ++i) {
            pmm *= -fact * s;
            fact += 2.0;
        }
    }
    double cm = 1.0, sm = 0.0;
    for (int i = 0; i < m; ++i) {
        const double c = cm * cphi - sm * sphi;
        sm = sm * cphi + cm * sphi;
        cm = c;
    }
    int qc = m * m + m;           // l^2 + l at l = m
    double p0 = pmm;              // P_{l-1}^m
    double p1 = x * (2 * m + 1) * pmm;
    Y[qc + m] = pmm * cm;
    if (m > 0) Y[qc - m] = pmm * sm;
    if (m == L) return;
    qc += 2 * m + 2;              // l = m + 1
    Y[qc + m] = p1 * cm;
    if (m > 0) Y[qc - m] = p1 * sm;
    double a = 2 * m + 3, bcoef = 2 * m + 1;   // 2l-1 and l+m-1 at l = m + 2
    for (int l = m + 2; l <= L; ++l) {
        const double P = (x * a * p1 - bcoef * p0) * rinv[l - m];
        p0 = p1;
        p1 = P;
        qc += 2 * l;              // l^2 + l from (l-1)^2 + (l-1)
        Y[qc + m] = P * cm;
        if (m > 0) Y[qc - m] = P * sm;
        a += 2.0;
        bcoef += 1.0;
    }
}

// w v^l L_k^{(l+1/2)}(v^2) for all k at one (l, v) into out[k] (eval_radial,
// compute_rtensor_sumfac_mex.cpp:14-34, with RadialNorm applied later to the summed blocks and the
// division by k taken as a product with the constant 1/k).
template <int NK>
__device__ inline void radial_raw(int l, double v, double w, double *out) {
    double vl = w;
    for (int i = 0; i < l; ++i) vl *= v;
    const double v2 = v * v, al = (double)l + 0.5;
    double L0 = 1.0, L1 = 1.0 + al - v2;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        double Lk;
        if (k == 0)
            Lk = 1.0;
        else if (k == 1)
            Lk = L1;
        else {
            Lk = ((2.0 * (k - 1) + 1.0 + al - v2) * L1 - ((k - 1) + al) * L0) * (1.0 / (double)k);
            L0 = L1;
            L1 = Lk;
        }
        out[k] = Lk * vl;
    }
}

// (c + i s)^m by binary powering: cos(m phi), sin(m phi) from cos(phi), sin(phi).
__device__ inline void cis_pow(int m, double c, double s, double &cm, double &sm) {
    double rc = 1.0, rs = 0.0, bc = c, bs = s;
    while (m) {
        if (m & 1) {
            const double t = rc * bc - rs * bs;
            rs = rc * bs + rs * bc;
            rc = t;
        }
        m >>= 1;
        if (m) {
            const double t = bc * bc - bs * bs;
            bs = 2.0 * bc * bs;
            bc = t;
        }
    }
    cm = rc;
    sm = rs;
}

// x^m by binary powering.
__device__ inline double ipow(double x, int m) {
    double r = 1.0;
    while (m) {
        if (m & 1) r *= x;
        m >>= 1;
        if (m) x *= x;
    }
    return r;
}

// Unnormalized real SH of the two orders m0 <= m1 = L - m0, degrees m..L, in one loop (two
// independent recursions per step).  P_m^m = (-1)^m (2m-1)!! s^m with s = sqrt(1 - x^2) and
// (2m-1)!! from dfact; other conventions as sh_order_m_raw.
__device__ inline void sh_pair_raw(int L, int m0, double x, double cphi, double sphi, const double *rinv,
                                   const double *dfact, double *Y) {
    const int m1 = L - m0;
    const bool two = m1 != m0;
    const double s = sqrt(fmax((1.0 - x) * (1.0 + x), 0.0));
    double c0, s0, c1, s1;
    cis_pow(m0, cphi, sphi, c0, s0);
    cis_pow(m1, cphi, sphi, c1, s1);
    const double pa = ((m0 & 1) ? -dfact[m0] : dfact[m0]) * ipow(s, m0);
    const double pb = ((m1 & 1) ? -dfact[m1] : dfact[m1]) * ipow(s, m1);
    // degree m and m + 1
    int qa = m0 * m0 + m0, qb = m1 * m1 + m1;
    double a0 = pa, a1 = x * (2 * m0 + 1) * pa;
    double b0 = pb, b1 = x * (2 * m1 + 1) * pb;
    Y[qa + m0] = a0 * c0;
    if (m0 > 0) Y[qa - m0] = a0 * s0;
    if (two) {
        Y[qb + m1] = b0 * c1;
        Y[qb - m1] = b0 * s1;
    }
    if (m0 == L) return;
    qa += 2 * m0 + 2;
    Y[qa + m0] = a1 * c0;
    if (m0 > 0) Y[qa - m0] = a1 * s0;
    if (two && m1 < L) {
        qb += 2 * m1 + 2;
        Y[qb + m1] = b1 * c1;
        Y[qb - m1] = b1 * s1;
    }
    double aa = 2 * m0 + 3, ab = 2 * m0 + 1;   // 2l-1, l+m-1 of chain a at l = m0 + 2
    double ba = 2 * m1 + 3, bb = 2 * m1 + 1;
    for (int i = 2; i <= L - m0; ++i) {
        const int la = m0 + i;
        const double P = (x * aa * a1 - ab * a0) * rinv[i];
        a0 = a1;
        a1 = P;
        qa += 2 * la;
        Y[qa + m0] = P * c0;
        if (m0 > 0) Y[qa - m0] = P * s0;
        aa += 2.0;
        ab += 1.0;
        if (two && i <= L - m1) {
            const int lb = m1 + i;
            const double Q = (x * ba * b1 - bb * b0) * rinv[i];
            b0 = b1;
            b1 = Q;
            qb += 2 * lb;
            Y[qb + m1] = Q * c1;
            Y[qb - m1] = Q * s1;
            ba += 2.0;
            bb += 1.0;
        }
    }
}

// w v^l L_k^{(l+1/2)}(v^2) for the two degrees l0 and l1 = L - l0 (if different), the two
// Laguerre recursions interleaved; outputs at A[l * NK + k] (conventions as radial_raw).
template <int NK>
__device__ inline void radial_pair_raw(int L, int l0, double v, double w, double *A) {
    const int l1 = L - l0;
    const bool two = l1 != l0;
    const double v2 = v * v;
    const double w0 = w * ipow(v, l0), w1 = w * ipow(v, l1);
    const double al0 = (double)l0 + 0.5, al1 = (double)l1 + 0.5;
    double P0 = 1.0, P1 = 1.0 + al0 - v2, Q0 = 1.0, Q1 = 1.0 + al1 - v2;
    double *o0 = A + l0 * NK, *o1 = A + l1 * NK;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        double Pk, Qk;
        if (k == 0) {
            Pk = 1.0;
            Qk = 1.0;
        } else if (k == 1) {
            Pk = P1;
            Qk = Q1;
        } else {
            Pk = ((2.0 * (k - 1) + 1.0 + al0 - v2) * P1 - ((k - 1) + al0) * P0) * (1.0 / (double)k);
            Qk = ((2.0 * (k - 1) + 1.0 + al1 - v2) * Q1 - ((k - 1) + al1) * Q0) * (1.0 / (double)k);
            P0 = P1;
            P1 = Pk;
            Q0 = Q1;
            Q1 = Qk;
        }
        o0[k] = Pk * w0;
        if (two) o1[k] = Qk * w1;
    }
}

// Tile t of a unit's M blocks: side, degree l and column block c (columns m = 16c .. 16c+15).
__device__ inline void mf_tile(int t, int L, int &side, int &l, int &c) {
    int n_side = 0;
    for (int ll = 0; ll <= L; ++ll) n_side += (2 * ll + 16) / 16;
    side = t >= n_side;
    int r = t - side * n_side;
    l = 0;
    while (r >= (2 * l + 16) / 16) {
        r -= (2 * l + 16) / 16;
        ++l;
    }
    c = r;
}

// Stage A, matrix-core form.  One workgroup per unit; block b maps to the patch-1 units of the
// chunk first (they carry N_t1 inner steps each, so they start early), then the patch-2 units.
// Storage index of a unit: (x - x0) * Ux + r, the reference's unit order.
// Outputs as the per-channel kernel: net[(unit*NL + tc)*2*NK + side*NK + k1], rad[unit][side][l][k].
template <int NK, int MT>
__global__ __launch_bounds__(MF_THREADS, 4) void sumfac_units_mf_kernel(SumfacDev d, int x0, int nx,
                                                                        double *__restrict__ net,
                                                                        double *__restrict__ rad) {
    extern __shared__ double lds[];
    __shared__ double sRadU[2][(MAXL + 1) * MAXK];
    __shared__ TGeom sG;

    const int L = d.L, NQ = d.NQ, NL = d.NL;
    const int AST = mf_stride((L + 1) * NK), YST = mf_stride(NQ);
    const MfLayout lay = mf_layout(L, NK, d.Nt1, d.Nchi, d.Neps);
    double *sY = lds + lay.y;
    double *sA = lds + lay.a;
    double *sGeo = lds + lay.geo;
    double *sLw = lds + lay.loss;
    double *sCE = lds + lay.trig, *sSE = sCE + d.Neps, *sCC = sSE + d.Neps, *sSC = sCC + d.Nchi;
    double *sWchi = lds + lay.wts, *sWeps = sWchi + d.Nchi;
    double *sShN = lds + lay.norm, *sRadN = sShN + NQ, *sRinv = sRadN + (L + 1) * NK, *sDfact = sRinv + 32;
    double *sM = lds;

    const int Ux = d.Nu1 + d.Ny2 * d.Nt2;
    const int b = blockIdx.x;
    int xl, r;
    if (b < nx * d.Nu1) {
        xl = b / d.Nu1;
        r = b % d.Nu1;
    } else {
        const int bb = b - nx * d.Nu1;
        xl = bb / (d.Ny2 * d.Nt2);
        r = d.Nu1 + bb % (d.Ny2 * d.Nt2);
    }
    const int64_t unit = (int64_t)xl * Ux + r;
    const int xi = x0 + xl;
    const bool patch1 = r < d.Nu1;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ns_all = d.qi_off[NL];
    const double x_i = d.x[xi];

    for (int i = tid; i < d.Neps; i += MF_THREADS) {
        sCE[i] = cos(d.eps[i]);
        sSE[i] = sin(d.eps[i]);
        sWeps[i] = d.W_eps[i];
    }
    for (int i = tid; i < d.Nchi; i += MF_THREADS) {
        const double c = d.mu_chi[i];
        sCC[i] = c;
        sSC[i] = sqrt(fmax(1.0 - c * c, 0.0));
        sWchi[i] = d.W_chi[i];
    }
    for (int i = tid; i < NQ; i += MF_THREADS) sShN[i] = d.shnorm[i];
    for (int i = tid; i < (L + 1) * NK; i += MF_THREADS) sRadN[i] = d.radnorm[i];
    for (int i = tid; i < 32; i += MF_THREADS) sRinv[i] = c_rinv[i];
    if (tid == 0) {
        double f = 1.0;
        sDfact[0] = 1.0;
        for (int i = 1; i < 32; ++i) {
            f *= (double)(2 * i - 1);
            sDfact[i] = f;
        }
    }
    // operand pads: read by products that land in unused rows/columns, kept finite
    for (int i = tid; i < lay.geo; i += MF_THREADS) lds[i] = 0.0;

    double v_large, v_small, Jac_unit;
    int u_idx = 0, y_idx = 0, t2_idx = 0;
    if (patch1) {
        u_idx = r;
        const double u = d.u1[u_idx];
        v_large = sqrt(x_i / 2.0) * (1.0 + u);
        v_small = sqrt(x_i / 2.0) * (1.0 - u);
        Jac_unit = 0.5 * d.Wx[xi] * d.Wu1[u_idx] * u * exp(-x_i * u * u) * (v_large * v_large * v_small * v_small);
    } else {
        y_idx = (r - d.Nu1) / d.Nt2;
        t2_idx = (r - d.Nu1) % d.Nt2;
        const double y = d.y2[y_idx];
        Jac_unit = 0.5 * d.Wx[xi] * d.Wy2[y_idx] * y * (4.0 * y);
        const double u = y * d.t2[t2_idx];
        v_large = sqrt(x_i / 2.0) * (1.0 + u);
        v_small = sqrt(x_i / 2.0) * (1.0 - u);
    }
    if (tid < 2 * (L + 1)) {
        const int side = tid / (L + 1), l = tid % (L + 1);
        double tmp[NK];
        radial_all_k<NK>(l, side ? v_small : v_large, d.radnorm, tmp);
#pragma unroll
        for (int k = 0; k < NK; ++k) sRadU[side][l * MAXK + k] = tmp[k];
    }

    // this wave's tiles
    int n_side = 0;
    for (int ll = 0; ll <= L; ++ll) n_side += (2 * ll + 16) / 16;
    const int n_tiles = 2 * n_side;
    // tiles t = wave + 4 i; past n_tiles a slot repeats tile 0 (its result unused), so the product
    // loop has no branches
    int t_aoff[MT], t_boff[MT], t_l[MT], t_c[MT], t_side[MT];
#pragma unroll
    for (int i = 0; i < MT; ++i) {
        const int t = wave + MF_WAVES * i;
        int side = 0, l = 0, c = 0;
        mf_tile(t < n_tiles ? t : 0, L, side, l, c);
        t_side[i] = side;
        t_l[i] = l;
        t_c[i] = c;
        t_aoff[i] = side * MF_NPB * AST + l * NK;
        t_boff[i] = side * MF_NPB * YST + l * l + 16 * c;
    }
    const int i16 = lane & 15, pq = lane >> 4;

    const int n_inner_t = patch1 ? d.Nt1 : 1;
    const int n_gain = d.Nchi * d.Neps;
    const int n_out = NL * 2 * NK;
    double *netu = net + unit * n_out;
    // evaluation tasks per batch: (item, j) covers SH orders (or radial degrees) j and L - j, so
    // every task runs L + 2 recursion steps; SH tasks on lanes [0, 128), radial tasks on
    // [128, MF_GEO_LANE), the next batch's geometry on the last MF_ITEMS lanes
    const int npair = (L + 2) / 2;
    const int n_ev = MF_ITEMS * npair;
    // geometry of the items of the batch at p0 into sGeo buffer gb (one thread per item)
    auto geometry = [&](int item, int p0, int gb) {
        const int side = item / MF_NPB, j = item % MF_NPB;
        const int p = p0 + j;
        const int pc = p < n_gain ? p : n_gain - 1;
        const int n = pc / d.Neps, e = pc % d.Neps;
        double vp_mag, ct, cp, sp;
        post_collision(sG, sCC[n], sSC[n], sCE[e], sSE[e], side, vp_mag, ct, cp, sp);
        double *g = sGeo + (gb * MF_ITEMS + item) * 5;
        g[0] = vp_mag;
        g[1] = ct;
        g[2] = cp;
        g[3] = sp;
        g[4] = p < n_gain ? (sWchi[n] * sG.B_val * sG.Jac_t) * sWeps[e] : 0.0;
    };
    for (int it = 0; it < n_inner_t; ++it) {
        __syncthreads();  // sG, sM of the previous step
        if (tid == 0) {
            if (patch1) {
                const double u = d.u1[u_idx], t = d.t1[it];
                const double y = u * t;
                const double Jac_t = d.Wt1[it] * (4.0 * u * t);
                const double c_beta = 1.0 - 2.0 * y * y;
                const double s_beta = sqrt(fmax(1.0 - c_beta * c_beta, EPS_SAFE));
                const double u_mag = sqrt(2.0 * x_i) * u * sqrt(1.0 + (1.0 - u * u) * t * t);
                make_tgeom(sG, x_i, v_large, v_small, c_beta, s_beta, u_mag, Jac_t, d.alpha);
            } else {
                const double y = d.y2[y_idx], t = d.t2[t2_idx];
                const double u = y * t;
                const double c_beta = 1.0 - 2.0 * y * y;
                const double s_beta = sqrt(fmax(1.0 - c_beta * c_beta, EPS_SAFE));
                const double Jac_t = d.Wt2[t2_idx] * exp(-x_i * u * u) * (v_large * v_large * v_small * v_small);
                const double u_mag = sqrt(2.0 * x_i) * y * sqrt(t * t + 1.0 - y * y * t * t);
                make_tgeom(sG, x_i, v_large, v_small, c_beta, s_beta, u_mag, Jac_t, d.alpha);
            }
        }
        __syncthreads();
        if (tid < MF_ITEMS) geometry(tid, 0, 0);
        v4d acc[MT];
#pragma unroll
        for (int i = 0; i < MT; ++i) acc[i] = v4d{0.0, 0.0, 0.0, 0.0};

        int gb = 0;
        for (int p0 = 0; p0 < n_gain; p0 += MF_NPB, gb ^= 1) {
            __syncthreads();  // this batch's geometry written; the previous batch's operands consumed
            const double *geo = sGeo + gb * MF_ITEMS * 5;
            if (tid < MF_SH_LANES) {
                for (int t = tid; t < n_ev; t += MF_SH_LANES) {
                    const int item = t % MF_ITEMS, j = t / MF_ITEMS;
                    const double *g = geo + 5 * item;
                    sh_pair_raw(L, j, g[1], g[2], g[3], sRinv, sDfact, sY + item * YST);
                }
            } else if (tid < MF_GEO_LANE) {
                for (int t = tid - MF_SH_LANES; t < n_ev; t += MF_GEO_LANE - MF_SH_LANES) {
                    const int item = t % MF_ITEMS, j = t / MF_ITEMS;
                    const double *g = geo + 5 * item;
                    radial_pair_raw<NK>(L, j, g[0], g[4], sA + item * AST);
                }
            } else if (tid < MF_GEO_LANE + MF_ITEMS && p0 + MF_NPB < n_gain) {
                geometry(tid - MF_GEO_LANE, p0 + MF_NPB, gb ^ 1);
            }
            __syncthreads();
#pragma unroll
            for (int g4 = 0; g4 < MF_NPB / 4; ++g4) {
                const int pt = 4 * g4 + pq;
                double a[MT], bb[MT];
#pragma unroll
                for (int i = 0; i < MT; ++i) {
                    a[i] = sA[t_aoff[i] + pt * AST + i16];
                    bb[i] = sY[t_boff[i] + pt * YST + i16];
                }
#pragma unroll
                for (int i = 0; i < MT; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bb[i], acc[i], 0, 0, 0);
            }
        }
        __syncthreads();  // operand reads done: [0, geo) becomes the M blocks
        // D layout of v_mfma_f64_16x16x4_f64: register r of a lane holds row 4 r + lane / 16, column
        // lane % 16 (the f32 16x16x4 form holds rows 4 (lane / 16) + r instead).  The blocks take
        // their SH_Norm and RadialNorm factors here.
#pragma unroll
        for (int i = 0; i < MT; ++i) {
            if (wave + MF_WAVES * i < n_tiles) {
                const int l = t_l[i], m = 16 * t_c[i] + i16;
                const int q = l * l + m;
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int k = 4 * rr + pq;
                    if (k < NK && m < 2 * l + 1)
                        sM[(t_side[i] * NQ + q) * NK + k] = acc[i][rr] * (sShN[q] * sRadN[k + NK * l]);
                }
            }
        }
        __syncthreads();
        // slot contraction: one (t_chan, side) per thread, all k
        const double *Pv = patch1 ? d.Pg1 + ((int64_t)u_idx * d.Nt1 + it) * ns_all : d.Pg2 + (int64_t)y_idx * ns_all;
        for (int pr = tid; pr < 2 * NL; pr += MF_THREADS) {
            const int tc = pr >> 1, side = pr & 1;
            const int s0 = d.qi_off[tc], s1 = d.qi_off[tc + 1];
            double gk[NK];
#pragma unroll
            for (int k = 0; k < NK; ++k) gk[k] = 0.0;
            for (int s = s0; s < s1; ++s) {
                const double P = Pv[s];
                const double *Mq = sM + (side * NQ + d.qi[s]) * NK;
#pragma unroll
                for (int k = 0; k < NK; ++k) gk[k] += P * Mq[k];
            }
            double *o = netu + pr * NK;
            if (it == 0) {
#pragma unroll
                for (int k = 0; k < NK; ++k) o[k] = gk[k];
            } else {
#pragma unroll
                for (int k = 0; k < NK; ++k) o[k] += gk[k];
            }
        }
        if (tid == 0) sLw[it] = sG.loss_weight;
    }
    __syncthreads();
    // loss: the radial factor is fixed per unit, so only loss_weight * P_loss is summed over t
    for (int pr = tid; pr < 2 * NL; pr += MF_THREADS) {
        const int tc = pr >> 1, side = pr & 1;
        const int l1 = d.trip[3 * tc];
        double ls = 0.0;
        if (patch1) {
            for (int it = 0; it < d.Nt1; ++it) ls += sLw[it] * d.Pl1[((int64_t)u_idx * d.Nt1 + it) * NL + tc];
        } else {
            ls = sLw[0] * d.Pl2[(int64_t)y_idx * NL + tc];
        }
        double *o = netu + pr * NK;
#pragma unroll
        for (int k = 0; k < NK; ++k) o[k] = Jac_unit * (o[k] - sRadU[side][l1 * MAXK + k] * ls);
    }
    double *ro = rad + unit * 2 * (L + 1) * NK;
    for (int i = tid; i < 2 * (L + 1) * NK; i += MF_THREADS) {
        const int side = i / ((L + 1) * NK), rem = i % ((L + 1) * NK);
        ro[i] = sRadU[side][(rem / NK) * MAXK + rem % NK];
    }
}

// ---------------------------------------------------------------------------------------------
// Exact-sphere variant (the default where it applies).
//
// For one (unit, inner t, side) the gain points are v'(omega) = c + rho R omega with c = U / 2,
// rho = |u| / 2, R the rotation taking z to the relative velocity's direction and omega on the unit
// sphere, sampled by the product rule (mu_chi, W_chi) x (eps, W_eps)
// (compute_rtensor_sumfac_mex.cpp:186-213).  A slot's integrand
//     Rad(|v'|, l, k) Y_q(v'/|v'|) = RadN SH_Norm L_k^{(l+1/2)}(|v'|^2) S_q(v'),
// S_q(v) = |v|^l P_l^|m|(v_z/|v|) {cos, sin}(|m| phi) a harmonic polynomial of degree l, agrees on
// the sphere with a polynomial of degree l + k in omega (|v'|^2 = |c|^2 + rho^2 + 2 rho c . R omega
// is linear there).  When the rule integrates every polynomial of degree L + N_K - 1 over the
// sphere exactly (checked on the host; the grid sizes of src/GeneralCollisionTensor.m:63-64 are
// larger), the gain sum is the surface integral.  On each
// circle omega . c_hat = t the factor L_k(|v'|^2) is constant and the mean of S_q over the circle is
// S_q(c_hat) Z_l(t), with Z_l = |v'|^l P_l(cos angle(v', c)) the zonal solid harmonic about c (the
// mean over rotations about c_hat of a degree-l harmonic is its zonal part), so
//     sum_p w_p Rad Y = 2 pi B Jac_t RadN SH_Norm S_q(c_hat) int_{-1}^{1} L_k(s(t)) Z_l(t) dt,
//     s(t) = |c + rho t c_hat|^2 + rho^2 (1 - t^2),
// a polynomial of degree k + l in t, integrated exactly by Gauss-Legendre with ceil((L + N_K) / 2)
// nodes.  A unit step then needs the SH at two directions and 2 (L+1) N_K one-dimensional integrals
// instead of 2 N_chi N_eps basis evaluations, and the t_chan sums factor over k:
//     Phi_gain[k][tc] += G[l1][k] * sum_{s in slots(tc)} P(s) SH_Norm S_{q_s}(c_hat).
// The result equals the reference's quadrature up to rounding.
//
// One workgroup per unit; thread tc owns t_chan tc (both sides, all k) in registers across the
// inner-t loop.  Per step: the integrals on 16-lane groups (one lane per node, a group per (side, l),
// summed across the group by lane shuffles), the SH at c_hat on one lane per (side, m), then the
// slot sums.  The geometry of all inner steps is formed first, one thread per step.
// ---------------------------------------------------------------------------------------------

constexpr int SP_NODES = 16;  // lanes per (side, l) group; nodes ceil((L + N_K) / 2) <= 14

struct SphereRule {
    double t[SP_NODES], w[SP_NODES];  // Gauss-Legendre on [-1, 1], padded with zero weights
};
constexpr int SP_MAX_THREADS = 640;  // sphere_threads at L <= 12

// Threads the kernel needs at degree L with NL t_chans: the integral groups, the SH lanes, the
// t_chan lanes.
inline int sphere_threads(int L, int NL) {
    const int ev = 2 * (L + 1) * SP_NODES + 2 * (L + 1);
    const int n = std::max(ev, NL);
    return (n + 63) / 64 * 64;
}

// Slot sum of t_chan tc for both sides: d_side = sum_{s in [s0, s1)} P(s) Y_side[q_s], loads issued
// 8 slots at a time.
__device__ inline void slot_sums(const double *__restrict__ Pv, const int *__restrict__ qi, int s0, int s1,
                                 const double *Y0, const double *Y1, double &d0, double &d1) {
    d0 = 0.0;
    d1 = 0.0;
    for (int j0 = s0; j0 < s1; j0 += 8) {
        double P[8];
        int q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const bool ok = j0 + u < s1;
            P[u] = ok ? Pv[j0 + u] : 0.0;
            q[u] = ok ? qi[j0 + u] : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            d0 += P[u] * Y0[q[u]];
            d1 += P[u] * Y1[q[u]];
        }
    }
}

// MULTI: the patch-1 units (N_t1 inner steps each, t_chan sums accumulated in registers), grid
// nx * N_u1; otherwise the patch-2 units (one inner step), grid nx * N_y2 * N_t2.  Storage index of
// a unit: (x - x0) * Ux + r, the reference's unit order.
template <int NK, bool MULTI>
__global__ __launch_bounds__(SP_MAX_THREADS) void sumfac_units_sphere_kernel(SumfacDev d,
                                                                             const SphereRule *__restrict__ rule,
                                                                             int x0, double *__restrict__ net,
                                                                             double *__restrict__ rad) {
    __shared__ double sY[2][(MAXL + 1) * (MAXL + 1)];
    __shared__ double sGk[2][(MAXL + 1) * NK];
    __shared__ double sRadU[2][(MAXL + 1) * MAXK];
    __shared__ TGeom sG[MULTI ? MAXG : 1];  // geometry of every inner step of the unit

    const int L = d.L, NL = d.NL;
    const int Ux = d.Nu1 + d.Ny2 * d.Nt2;
    const int b = blockIdx.x;
    int xl, r;
    if (MULTI) {
        xl = b / d.Nu1;
        r = b % d.Nu1;
    } else {
        xl = b / (d.Ny2 * d.Nt2);
        r = d.Nu1 + b % (d.Ny2 * d.Nt2);
    }
    const int64_t unit = (int64_t)xl * Ux + r;
    const int xi = x0 + xl;
    const int tid = threadIdx.x;
    const double x_i = d.x[xi];

    double v_large, v_small, Jac_unit;
    int u_idx = 0, y_idx = 0, t2_idx = 0;
    if (MULTI) {
        u_idx = r;
        const double u = d.u1[u_idx];
        v_large = sqrt(x_i / 2.0) * (1.0 + u);
        v_small = sqrt(x_i / 2.0) * (1.0 - u);
        Jac_unit = 0.5 * d.Wx[xi] * d.Wu1[u_idx] * u * exp(-x_i * u * u) * (v_large * v_large * v_small * v_small);
    } else {
        y_idx = (r - d.Nu1) / d.Nt2;
        t2_idx = (r - d.Nu1) % d.Nt2;
        const double y = d.y2[y_idx];
        Jac_unit = 0.5 * d.Wx[xi] * d.Wy2[y_idx] * y * (4.0 * y);
        const double u = y * d.t2[t2_idx];
        v_large = sqrt(x_i / 2.0) * (1.0 + u);
        v_small = sqrt(x_i / 2.0) * (1.0 - u);
    }
    // geometry of inner step it (as the matrix-core kernel)
    auto tgeom = [&](int it, TGeom &g) {
        if (MULTI) {
            const double u = d.u1[u_idx], t = d.t1[it];
            const double y = u * t;
            const double Jac_t = d.Wt1[it] * (4.0 * u * t);
            const double c_beta = 1.0 - 2.0 * y * y;
            const double s_beta = sqrt(fmax(1.0 - c_beta * c_beta, EPS_SAFE));
            const double u_mag = sqrt(2.0 * x_i) * u * sqrt(1.0 + (1.0 - u * u) * t * t);
            make_tgeom(g, x_i, v_large, v_small, c_beta, s_beta, u_mag, Jac_t, d.alpha);
        } else {
            const double y = d.y2[y_idx], t = d.t2[t2_idx];
            const double u = y * t;
            const double c_beta = 1.0 - 2.0 * y * y;
            const double s_beta = sqrt(fmax(1.0 - c_beta * c_beta, EPS_SAFE));
            const double Jac_t = d.Wt2[t2_idx] * exp(-x_i * u * u) * (v_large * v_large * v_small * v_small);
            const double u_mag = sqrt(2.0 * x_i) * y * sqrt(t * t + 1.0 - y * y * t * t);
            make_tgeom(g, x_i, v_large, v_small, c_beta, s_beta, u_mag, Jac_t, d.alpha);
        }
    };
    const int n_ev = 2 * (L + 1) * SP_NODES;
    const int n_inner_t = MULTI ? d.Nt1 : 1;
    // radial tables at the unit's two speeds, on lanes past the ones forming the geometry
    if (tid >= 64 && tid < 64 + 2 * (L + 1)) {
        const int side = (tid - 64) / (L + 1), l = (tid - 64) % (L + 1);
        double tmp[NK];
        radial_all_k<NK>(l, side ? v_small : v_large, d.radnorm, tmp);
#pragma unroll
        for (int k = 0; k < NK; ++k) sRadU[side][l * MAXK + k] = tmp[k];
    }
    for (int it = tid; it < n_inner_t; it += blockDim.x) tgeom(it, sG[it]);

    // this thread's t_chan
    const bool own = tid < NL;
    const int tc = own ? tid : 0;
    const int s0 = d.qi_off[tc], s1 = own ? d.qi_off[tc + 1] : s0;
    const int l1 = d.trip[3 * tc];
    double acc0[MULTI ? NK : 1], acc1[MULTI ? NK : 1];
#pragma unroll
    for (int k = 0; k < (MULTI ? NK : 1); ++k) {
        acc0[k] = 0.0;
        acc1[k] = 0.0;
    }
    double dl0 = 0.0, dl1 = 0.0;  // one-step units: the slot sums themselves
    const int ns_all = d.qi_off[NL];
    for (int it = 0; it < n_inner_t; ++it) {
        __syncthreads();  // geometry written; the previous step's sY / sGk consumed
        const TGeom &g = sG[it];
        if (tid < n_ev) {
            // integral task: group (side, l), lane i = node
            const int grp = tid / SP_NODES, i = tid % SP_NODES;
            const int side = grp / (L + 1), l = grp % (L + 1);
            const double Ux_ = side ? g.U_x2 : g.U_x1, Uz_ = side ? g.U_z2 : g.U_z1;
            const double cm = 0.5 * sqrt(Ux_ * Ux_ + Uz_ * Uz_);
            const double rho = 0.5 * (side ? g.u_mag2 : g.u_mag1);
            const double ti = rule->t[i], wi = rule->w[i];
            const double z = cm + rho * ti;
            const double s = z * z + rho * rho * ((1.0 - ti) * (1.0 + ti));
            double z0 = 1.0, z1 = z;  // Z_{j-1}, Z_j
            for (int j = 1; j < l; ++j) {
                const double zn = ((2 * j + 1) * z * z1 - j * s * z0) * c_rinv[j + 1];
                z0 = z1;
                z1 = zn;
            }
            const double Zl = l == 0 ? 1.0 : z1;
            const double al = (double)l + 0.5;
            const double wz = wi * Zl;
            double P0 = 1.0, P1 = 1.0 + al - s;
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                double Lk;
                if (k == 0)
                    Lk = 1.0;
                else if (k == 1)
                    Lk = P1;
                else {
                    Lk = ((2.0 * (k - 1) + 1.0 + al - s) * P1 - ((k - 1) + al) * P0) * (1.0 / (double)k);
                    P0 = P1;
                    P1 = Lk;
                }
                double v = wz * Lk;
                v += __shfl_xor(v, 8);
                v += __shfl_xor(v, 4);
                v += __shfl_xor(v, 2);
                v += __shfl_xor(v, 1);
                if (i == 0) sGk[side][l * NK + k] = v * (2.0 * M_PI * g.B_val * g.Jac_t) * d.radnorm[k + NK * l];
            }
        } else if (tid < n_ev + 2 * (L + 1)) {
            // SH task: (side, m) at c_hat
            const int j = tid - n_ev;
            const int side = j / (L + 1), m = j % (L + 1);
            const double Ux_ = side ? g.U_x2 : g.U_x1, Uz_ = side ? g.U_z2 : g.U_z1;
            const double cn = sqrt(Ux_ * Ux_ + Uz_ * Uz_);
            const double ct = cn > 0.0 ? fmax(fmin(Uz_ / cn, 1.0), -1.0) : 1.0;
            const double cphi = Ux_ < 0.0 ? -1.0 : 1.0;   // c lies in the x-z plane
            sh_order_m(L, m, ct, cphi, 0.0, d.shnorm, sY[side]);
        }
        __syncthreads();
        if (own) {
            const double *Pv =
                MULTI ? d.Pg1 + ((int64_t)u_idx * d.Nt1 + it) * ns_all : d.Pg2 + (int64_t)y_idx * ns_all;
            double d0, d1;
            slot_sums(Pv, d.qi, s0, s1, sY[0], sY[1], d0, d1);
            if (MULTI) {
                const double *G0 = sGk[0] + l1 * NK, *G1 = sGk[1] + l1 * NK;
#pragma unroll
                for (int k = 0; k < (MULTI ? NK : 1); ++k) {
                    acc0[k] += d0 * G0[k];
                    acc1[k] += d1 * G1[k];
                }
            } else {
                dl0 = d0;
                dl1 = d1;
            }
        }
    }
    if (MULTI) __syncthreads();  // every step's loss weight formed
    if (own) {
        // loss: the radial factor is fixed per unit, so only loss_weight * P_loss is summed over t
        double ls = 0.0;
        if (MULTI) {
            for (int it = 0; it < d.Nt1; ++it)
                ls += sG[it].loss_weight * d.Pl1[((int64_t)u_idx * d.Nt1 + it) * NL + tc];
        } else {
            ls = sG[0].loss_weight * d.Pl2[(int64_t)y_idx * NL + tc];
        }
        double *o = net + (unit * NL + tc) * 2 * NK;
        const double *G0 = sGk[0] + l1 * NK, *G1 = sGk[1] + l1 * NK;
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const double g0 = MULTI ? acc0[MULTI ? k : 0] : dl0 * G0[k];
            const double g1 = MULTI ? acc1[MULTI ? k : 0] : dl1 * G1[k];
            o[k] = Jac_unit * (g0 - sRadU[0][l1 * MAXK + k] * ls);
            o[NK + k] = Jac_unit * (g1 - sRadU[1][l1 * MAXK + k] * ls);
        }
    }
    double *ro = rad + unit * 2 * (L + 1) * NK;
    for (int i = tid; i < 2 * (L + 1) * NK; i += blockDim.x) {
        const int side = i / ((L + 1) * NK), rem = i % ((L + 1) * NK);
        ro[i] = sRadU[side][(rem / NK) * MAXK + rem % NK];
    }
}

// Stage B as one product per (l2, l3) class over the units:
//     R[(tc,k1), (k2,k3)] += sum_u net1[u][tc][k1] B1[u][(k2,k3)] + net2[u][tc][k1] B2[u][(k2,k3)],
//     B1 = r2l (x) r3s, B2 = r2s (x) r3l at the class's degrees,
// shared by every t_chan of the class.  Rows (tc, k1) of a class are cut into tiles of 16 (a row
// table built on the host; padding rows -1), columns are the N_K^2 pairs in blocks of 16, and one
// wave computes one row tile x all column blocks for one range of units, the accumulators of the
// blocks held in registers across the range: per 4 units one load of each net term per lane feeds
// 2 N_blocks v_mfma_f64_16x16x4_f64 (first all blocks' net1 products, then all net2 products, so
// no MFMA waits on the one before it), the next step's net operands already in flight.  The units'
// radial rows at l2 and l3 are staged in LDS 32 units at a time, (r2l, r2s, r3l, r3s) per k in one
// 32-byte entry.  Each unit range writes its partial R; sumfac_reduce_kernel adds the ranges in
// order (deterministic).  Workgroups go to the 8 XCDs round-robin by block index; the block ->
// (tile, range) table puts the tiles of one class and one range on one XCD, so they share that
// XCD's L2 copy of the radial rows and of the net rows of a t_chan split across two tiles.
constexpr int FOLD_SUB = 32;  // units staged per LDS round
constexpr int N_XCD = 8;
typedef double double4v __attribute__((ext_vector_type(4)));

template <int NK>
__global__ __launch_bounds__(64) void sumfac_fold_tiles_kernel(int NL, int L, const int *__restrict__ tile_rows,
                                                               const int *__restrict__ tile_l23,
                                                               const int2 *__restrict__ items, int64_t n_units,
                                                               int64_t units_per_range,
                                                               const double *__restrict__ net,
                                                               const double *__restrict__ rad,
                                                               double *__restrict__ partial) {
    constexpr int NPAIR = NK * NK, NBLK = (NPAIR + 15) / 16;
    constexpr int64_t NR1 = (int64_t)NK * NK * NK;
    __shared__ double4v sR[FOLD_SUB][NK + 1];  // per unit and k: (r2l, r2s, r3l, r3s); k = NK pads
    const int2 item = items[blockIdx.x];
    if (item.x < 0) return;  // padding of the per-XCD lists
    const int tile = item.x, range = item.y;
    const int lane = threadIdx.x, i16 = lane & 15, g = lane >> 4;
    const int l2 = tile_l23[2 * tile], l3 = tile_l23[2 * tile + 1];
    const int row = tile_rows[16 * tile + i16];
    const int64_t u_begin = (int64_t)range * units_per_range;
    const int64_t u_end = u_begin + units_per_range < n_units ? u_begin + units_per_range : n_units;
    const double arow = row >= 0 ? 1.0 : 0.0;
    const int64_t nstride = (int64_t)NL * 2 * NK;
    const double *nbase = net + (row >= 0 ? (int64_t)(row / NK) * 2 * NK + row % NK : 0);
    int k2o[NBLK], k3o[NBLK];
#pragma unroll
    for (int j = 0; j < NBLK; ++j) {
        const int p = 16 * j + i16;
        k2o[j] = p < NPAIR ? p / NK : NK;  // NK: padding (b = 0)
        k3o[j] = p < NPAIR ? p % NK : 0;
    }
    v4d acc[NBLK];
#pragma unroll
    for (int j = 0; j < NBLK; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
    const int rstride = 2 * (L + 1) * NK;
    // net operands of one 4-unit step (lane: unit g, row i16), fetched one step ahead
    auto fetch_net = [&](int64_t u, double &a1, double &a2) {
        const bool ok = u < u_end;
        const double *nu = nbase + (ok ? u : u_begin) * nstride;
        const double z = ok ? arow : 0.0;
        a1 = nu[0] * z;
        a2 = nu[NK] * z;
    };
    double a1n = 0.0, a2n = 0.0;
    if (u_begin < u_end) fetch_net(u_begin + g, a1n, a2n);
    if (lane < FOLD_SUB) sR[lane][NK] = double4v{0.0, 0.0, 0.0, 0.0};  // the padding pairs read these
    for (int64_t s0 = u_begin; s0 < u_end; s0 += FOLD_SUB) {
        const int ns = u_end - s0 < FOLD_SUB ? (int)(u_end - s0) : FOLD_SUB;
        __syncthreads();  // the previous round's rows consumed
        for (int i = lane; i < FOLD_SUB * NK; i += 64) {
            const int uu = i / NK, k = i % NK;
            const double *ru = rad + (s0 + (uu < ns ? uu : 0)) * rstride;
            const double z = uu < ns ? 1.0 : 0.0;
            sR[uu][k] = double4v{ru[l2 * NK + k] * z, ru[(L + 1) * NK + l2 * NK + k] * z, ru[l3 * NK + k] * z,
                                 ru[(L + 1) * NK + l3 * NK + k] * z};
        }
        __syncthreads();
        for (int q = 0; q < ns; q += 4) {
            const double a1 = a1n, a2 = a2n;
            fetch_net(s0 + q + 4 + g, a1n, a2n);
            const double4v *r = sR[q + g];
            double b1[NBLK], b2[NBLK];
#pragma unroll
            for (int j = 0; j < NBLK; ++j) {
                const double4v x2 = r[k2o[j]], x3 = r[k3o[j]];
                b1[j] = x2[0] * x3[3];  // r2l r3s
                b2[j] = x2[1] * x3[2];  // r2s r3l
            }
#pragma unroll
            for (int j = 0; j < NBLK; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1[j], acc[j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < NBLK; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b2[j], acc[j], 0, 0, 0);
        }
    }
    // D layout: register rr of a lane holds row 4 rr + lane / 16, column lane % 16
    double *pout = partial + (int64_t)range * NR1 * NL;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        const int orow = tile_rows[16 * tile + 4 * rr + g];
        if (orow < 0) continue;
        const int64_t obase = (int64_t)(orow / NK) * NR1 + orow % NK;
#pragma unroll
        for (int j = 0; j < NBLK; ++j) {
            const int p = 16 * j + i16;
            if (p < NPAIR) pout[obase + NK * (p / NK) + (int64_t)NK * NK * (p % NK)] = acc[j][rr];
        }
    }
}

// R[o] += sum over unit ranges r (in order) of partial[r][o].
__global__ __launch_bounds__(256) void sumfac_reduce_kernel(int64_t nR, int n_ranges, const double *__restrict__ partial,
                                                            double *__restrict__ R) {
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= nR) return;
    double s = R[o];
    for (int r = 0; r < n_ranges; ++r) s += partial[(int64_t)r * nR + o];
    R[o] = s;
}

// Row tiles of the class-level fold: t_chans grouped by (l2, l3) in first-appearance order, rows
// (tc, k1) of a class in tiles of 16 (tile_rows: tc * NK + k1, -1 for padding), tile_l23 per tile.
void fold_tiles(const SumfacTables &T, int NK, std::vector<int> &tile_rows, std::vector<int> &tile_l23,
                std::vector<int> &tile_class) {
    const int NL = T.n_l;
    std::vector<std::pair<int, int>> classes;
    std::vector<std::vector<int>> members;
    for (int tc = 0; tc < NL; ++tc) {
        const std::pair<int, int> key(T.trip[3 * tc + 1], T.trip[3 * tc + 2]);
        size_t c = 0;
        while (c < classes.size() && classes[c] != key) ++c;
        if (c == classes.size()) {
            classes.push_back(key);
            members.emplace_back();
        }
        members[c].push_back(tc);
    }
    tile_rows.clear();
    tile_l23.clear();
    tile_class.clear();
    for (size_t c = 0; c < classes.size(); ++c) {
        const int n_rows = (int)members[c].size() * NK;
        for (int r0 = 0; r0 < n_rows; r0 += 16) {
            for (int i = 0; i < 16; ++i) {
                const int r = r0 + i;
                tile_rows.push_back(r < n_rows ? members[c][r / NK] * NK + r % NK : -1);
            }
            tile_l23.push_back(classes[c].first);
            tile_l23.push_back(classes[c].second);
            tile_class.push_back((int)c);
        }
    }
}

// Wave slots of the fold kernel on this device (workgroups per CU at its register use x CUs).
template <int NK>
int fold_wave_slots() {
    int per_cu = 0, dev = 0, n_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sumfac_fold_tiles_kernel<NK>, 64, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * n_cu;
}

// Unit ranges of the fold: the fewest (>= 1, <= 32) for which n_tiles x ranges waves fill the last
// round of wave slots to at least 90 % (a partial last round leaves SIMDs idle at the end of the
// launch), else the best filled.
int fold_ranges(int NK, int n_tiles) {
    int slots = 0;
    switch (NK) {
#define X(KK)                                \
    case KK:                                 \
        slots = fold_wave_slots<KK>();       \
        break;
        X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
#undef X
    }
    if (slots <= 0 || n_tiles <= 0) return 1;
    int best = 1;
    double best_fill = 0.0;
    for (int r = 1; r <= 32; ++r) {
        const int64_t waves = (int64_t)n_tiles * r;
        const int64_t rounds = (waves + slots - 1) / slots;
        const double fill = (double)waves / (double)(rounds * slots);
        if (waves >= slots && fill >= 0.9) return r;
        if (fill > best_fill) {
            best_fill = fill;
            best = r;
        }
    }
    return best;
}

// Block -> (tile, range) for n_ranges unit ranges: the (range, class) groups are spread over the XCDs
// largest first onto the least loaded one; block b takes entry b / N_XCD of XCD b % N_XCD's list, the
// lists padded to one length with (-1, -1).
std::vector<int2> fold_items(const std::vector<int> &tile_class, int n_ranges) {
    const int n_tiles = (int)tile_class.size();
    const int n_classes = n_tiles ? tile_class.back() + 1 : 0;
    std::vector<std::vector<int>> tiles_of(n_classes);
    for (int t = 0; t < n_tiles; ++t) tiles_of[tile_class[t]].push_back(t);
    std::vector<int> order(n_classes);
    for (int c = 0; c < n_classes; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return tiles_of[a].size() > tiles_of[b].size(); });
    std::vector<std::vector<int2>> lists(N_XCD);
    for (int r = 0; r < n_ranges; ++r)
        for (int c : order) {
            size_t best = 0;
            for (size_t x = 1; x < lists.size(); ++x)
                if (lists[x].size() < lists[best].size()) best = x;
            for (int t : tiles_of[c]) lists[best].push_back(int2{t, r});
        }
    size_t len = 0;
    for (auto &l : lists) len = std::max(len, l.size());
    std::vector<int2> items(len * N_XCD, int2{-1, -1});
    for (int x = 0; x < N_XCD; ++x)
        for (size_t i = 0; i < lists[x].size(); ++i) items[i * N_XCD + x] = lists[x][i];
    return items;
}

// Compact P tables on the device (src/GeneralCollisionTensor.m:106-174), from a SumfacTablePlan.
// One workgroup per quadrature node: the patch-1 (u,t) nodes first ([N_u1][N_t1] order), then the
// patch-2 y nodes.  The real SH at phi = 0 of the node's polar angle (mu = 1 - 2 y^2, clamped) go to
// LDS (sin terms are 0 there), then one thread per slot sums its Gaunt entries in member order, as
// the MATLAB loop accumulates them, and applies the t_chan's SCALE; likewise the loss rows.
__global__ __launch_bounds__(256) void sumfac_ptable_kernel(
    int L, int n_l, int ns, int Nu, int Nt, int Ny, const double *__restrict__ u1, const double *__restrict__ t1,
    const double *__restrict__ y2, const double *__restrict__ shnorm, const int *__restrict__ slot_t,
    const int *__restrict__ ent_off, const int *__restrict__ ent_qk, const double *__restrict__ ent_g,
    const int *__restrict__ loss_off, const int *__restrict__ loss_qk, const double *__restrict__ loss_g,
    const double *__restrict__ scale, double *__restrict__ Pg1, double *__restrict__ Pl1, double *__restrict__ Pg2,
    double *__restrict__ Pl2) {
    __shared__ double sY[(MAXL + 1) * (MAXL + 1)];
    const int node = blockIdx.x;
    const bool p1 = node < Nu * Nt;
    const double y = p1 ? u1[node / Nt] * t1[node % Nt] : y2[node - Nu * Nt];
    const double mu = fmax(fmin(1.0 - 2.0 * y * y, 1.0), -1.0);
    const double x = cos(acos(mu));
    for (int m = threadIdx.x; m <= L; m += blockDim.x) {
        double pmm = 1.0;
        if (m > 0) {
            const double sq = sqrt(fmax((1.0 - x) * (1.0 + x), 0.0));
            double fact = 1.0;
            for (int i = 1; i <= m; ++i) {
                pmm *= -fact * sq;
                fact += 2.0;
            }
        }
        double p0 = pmm, pl = pmm;
        double p1v = x * (2 * m + 1) * pmm;
        for (int l = m; l <= L; ++l) {
            if (l == m + 1) {
                pl = p1v;
            } else if (l > m + 1) {
                pl = (x * (2 * l - 1) * p1v - (l + m - 1) * p0) / (l - m);
                p0 = p1v;
                p1v = pl;
            }
            const int q = l * l + l;
            sY[q + m] = shnorm[q + m] * pl;
            if (m > 0) sY[q - m] = 0.0;
        }
    }
    __syncthreads();
    double *Pg = p1 ? Pg1 + (size_t)node * ns : Pg2 + (size_t)(node - Nu * Nt) * ns;
    for (int sl = threadIdx.x; sl < ns; sl += blockDim.x) {
        double acc = 0.0;
        for (int e = ent_off[sl]; e < ent_off[sl + 1]; ++e) acc += ent_g[e] * sY[ent_qk[e]];
        Pg[sl] = acc * scale[slot_t[sl]];
    }
    double *Pl = p1 ? Pl1 + (size_t)node * n_l : Pl2 + (size_t)(node - Nu * Nt) * n_l;
    for (int t = threadIdx.x; t < n_l; t += blockDim.x) {
        double acc = 0.0;
        for (int e = loss_off[t]; e < loss_off[t + 1]; ++e) acc += loss_g[e] * sY[loss_qk[e]];
        Pl[t] = acc * scale[t];
    }
}

__global__ void conservation_zero_kernel(int NK, int NL, const int *__restrict__ trip, double *__restrict__ R) {
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = (int64_t)NK * NK * NK * NL;
    if (o >= total) return;
    const int k1 = (int)(o % NK);
    const int tc = (int)(o / ((int64_t)NK * NK * NK));
    const int l1 = trip[3 * tc];
    if ((l1 == 0 && k1 <= 1) || (l1 == 1 && k1 == 0)) R[o] = 0.0;
}

template <int NK, int MT>
int launch_units_mf_t(size_t dyn, const SumfacDev &d, int xc, int nx, int64_t n_units, double *net, double *rad) {
    CFK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&sumfac_units_mf_kernel<NK, MT>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    hipLaunchKernelGGL((sumfac_units_mf_kernel<NK, MT>), dim3((unsigned)n_units), dim3(MF_THREADS), dyn, nullptr, d,
                       xc, nx, net, rad);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

template <int NK>
int launch_units_mf(int mt, size_t dyn, const SumfacDev &d, int xc, int nx, int64_t n_units, double *net,
                    double *rad) {
    if (mt == 3) return launch_units_mf_t<NK, 3>(dyn, d, xc, nx, n_units, net, rad);
    return launch_units_mf_t<NK, 5>(dyn, d, xc, nx, n_units, net, rad);
}

int g_sumfac_kernel = CFK_SUMFAC_AUTO;

// True when the product rule (mu_chi, W_chi) x (eps, W_eps) integrates every polynomial of degree
// deg over the unit sphere exactly, to rounding: the chi rule integrates mu^j (j <= deg) over
// [-1, 1] and the eps rule integrates cos(m eps) and sin(m eps) (m <= deg) over [0, 2 pi).  The
// surface integrand of a monomial of degree deg in omega splits into such factors.
bool sphere_rule_exact(const SumfacGrids &G, int deg) {
    const size_t nc = G.mu_chi.size(), ne = G.eps.size();
    for (int j = 0; j <= deg; ++j) {
        double s = 0.0, a = 0.0;
        for (size_t n = 0; n < nc; ++n) {
            double p = 1.0;
            for (int e = 0; e < j; ++e) p *= G.mu_chi[n];
            s += G.W_chi[n] * p;
            a += std::fabs(G.W_chi[n] * p);
        }
        const double ex = (j & 1) ? 0.0 : 2.0 / (j + 1);
        if (!(std::fabs(s - ex) <= 1e-12 * std::max(a, ex))) return false;
    }
    double wsum = 0.0, wabs = 0.0;
    for (size_t e = 0; e < ne; ++e) {
        wsum += G.W_eps[e];
        wabs += std::fabs(G.W_eps[e]);
    }
    if (!(std::fabs(wsum - 2.0 * M_PI) <= 1e-12 * 2.0 * M_PI)) return false;
    for (int m = 1; m <= deg; ++m) {
        double c = 0.0, sn = 0.0;
        for (size_t e = 0; e < ne; ++e) {
            c += G.W_eps[e] * std::cos(m * G.eps[e]);
            sn += G.W_eps[e] * std::sin(m * G.eps[e]);
        }
        if (!(std::fabs(c) <= 1e-12 * wabs && std::fabs(sn) <= 1e-12 * wabs)) return false;
    }
    return true;
}

// Smallest |u|^2 (squared relative speed of make_tgeom, both sides) over the grid's units and inner
// steps of x nodes [x_begin, x_end).  Below EPS_SAFE the reference clamps it, and the rotation
// built from it is no longer orthogonal.
double min_relative_speed2(const SumfacGrids &G, int x_begin, int x_end) {
    double mn = INFINITY;
    auto side2 = [&](double vl, double vs, double c_beta, double s_beta) {
        const double ux1 = -vs * s_beta, uz1 = vl - vs * c_beta;
        const double ux2 = -vl * s_beta, uz2 = vs - vl * c_beta;
        mn = std::min(mn, std::min(ux1 * ux1 + uz1 * uz1, ux2 * ux2 + uz2 * uz2));
    };
    for (int xi = x_begin; xi < x_end; ++xi) {
        const double x = G.x[xi];
        for (size_t a = 0; a < G.u1.size(); ++a)
            for (size_t t = 0; t < G.t1.size(); ++t) {
                const double u = G.u1[a], y = u * G.t1[t];
                const double c_beta = 1.0 - 2.0 * y * y;
                const double s_beta = std::sqrt(std::max(1.0 - c_beta * c_beta, EPS_SAFE));
                side2(std::sqrt(x / 2.0) * (1.0 + u), std::sqrt(x / 2.0) * (1.0 - u), c_beta, s_beta);
            }
        for (size_t a = 0; a < G.y2.size(); ++a)
            for (size_t t = 0; t < G.t2.size(); ++t) {
                const double y = G.y2[a], u = y * G.t2[t];
                const double c_beta = 1.0 - 2.0 * y * y;
                const double s_beta = std::sqrt(std::max(1.0 - c_beta * c_beta, EPS_SAFE));
                side2(std::sqrt(x / 2.0) * (1.0 + u), std::sqrt(x / 2.0) * (1.0 - u), c_beta, s_beta);
            }
    }
    return mn;
}
int g_sumfac_chunk_nodes = 0;  // x nodes per chunk (0: sized by the buffer cap)

// Device time of the last quadrature call (HIP events around each launch on its stream).
struct SumfacTiming {
    int kind = 0;
    int64_t units = 0;
    int64_t chunks = 0;  // x-node chunks (unit-buffer launches) of the call
    double units_ms = 0.0, fold_ms = 0.0;
};
SumfacTiming g_last_timing;

struct SumfacBuffers {
    DevBuf<double> grid[14], radnorm, shnorm, Pg1, Pl1, Pg2, Pl2, net, rad, R;
    DevBuf<int> trip, qi_off, qi;
};

// T's P vectors are uploaded, or, with plan non-null, the P tables are built on the device.
int run_sumfac(const SumfacTables &T, const SumfacGrids &G, int NK, double alpha, int x_begin, int x_end,
               bool zero_conservation, double *R_host, const SumfacTablePlan *plan = nullptr) {
    int rc;
    if ((rc = check_device())) return rc;
    const int L = T.L, NL = T.n_l, NQ = T.N_Q;
    if (L > MAXL) return set_error(CFK_EUNSUP, "sumfac: L=%d > %d unsupported", L, MAXL);
    if (NK > MAXK || NK < 1) return set_error(CFK_EUNSUP, "sumfac: N_K=%d outside 1..%d", NK, MAXK);
    if ((int)G.mu_chi.size() > MAXG || (int)G.eps.size() > MAXG)
        return set_error(CFK_EUNSUP, "sumfac: N_chi and N_eps must be <= %d", MAXG);
    for (int t = 0; t < NL; ++t)
        if (T.qi_off[t + 1] - T.qi_off[t] > MAXSLOT)
            return set_error(CFK_EINVAL, "sumfac: t_chan %d has more than %d valid q_i", t + 1, MAXSLOT);
    SumfacBuffers b;
    const std::vector<double> *gv[14] = {&G.x, &G.W_x, &G.u1, &G.W_u1, &G.t1, &G.W_t1, &G.y2,
                                         &G.W_y2, &G.t2, &G.W_t2, &G.mu_chi, &G.W_chi, &G.eps, &G.W_eps};
    for (int i = 0; i < 14; ++i)
        if ((rc = b.grid[i].upload(gv[i]->data(), gv[i]->size()))) return rc;
    std::vector<double> rn((size_t)NK * (L + 1)), sn(NQ);
    radial_norm(NK, L, rn.data());
    sh_norm(L, sn.data());
    if ((rc = b.radnorm.upload(rn.data(), rn.size()))) return rc;
    if ((rc = b.shnorm.upload(sn.data(), sn.size()))) return rc;
    if ((rc = b.trip.upload(T.trip.data(), T.trip.size()))) return rc;
    if ((rc = b.qi_off.upload(T.qi_off.data(), T.qi_off.size()))) return rc;
    if ((rc = b.qi.upload(T.qi.data(), T.qi.size()))) return rc;
    if (plan) {
        const int Nu = (int)G.u1.size(), Nt = (int)G.t1.size(), Ny = (int)G.y2.size();
        const int ns = T.qi_off[NL];
        DevBuf<int> slot_t, ent_off, ent_qk, loss_off, loss_qk;
        DevBuf<double> ent_g, loss_g, scale;
        if ((rc = slot_t.upload(plan->slot_t.data(), plan->slot_t.size())) ||
            (rc = ent_off.upload(plan->ent_off.data(), plan->ent_off.size())) ||
            (rc = ent_qk.upload(plan->ent_qk.data(), plan->ent_qk.size())) ||
            (rc = ent_g.upload(plan->ent_g.data(), plan->ent_g.size())) ||
            (rc = loss_off.upload(plan->loss_off.data(), plan->loss_off.size())) ||
            (rc = loss_qk.upload(plan->loss_qk.data(), plan->loss_qk.size())) ||
            (rc = loss_g.upload(plan->loss_g.data(), plan->loss_g.size())) ||
            (rc = scale.upload(plan->scale.data(), plan->scale.size())))
            return rc;
        if ((rc = b.Pg1.alloc((size_t)Nu * Nt * ns)) || (rc = b.Pl1.alloc((size_t)Nu * Nt * NL)) ||
            (rc = b.Pg2.alloc((size_t)Ny * ns)) || (rc = b.Pl2.alloc((size_t)Ny * NL)))
            return rc;
        hipLaunchKernelGGL(sumfac_ptable_kernel, dim3((unsigned)(Nu * Nt + Ny)), dim3(256), 0, nullptr, L, NL, ns, Nu,
                           Nt, Ny, b.grid[2].p, b.grid[4].p, b.grid[6].p, b.shnorm.p, slot_t.p, ent_off.p, ent_qk.p,
                           ent_g.p, loss_off.p, loss_qk.p, loss_g.p, scale.p, b.Pg1.p, b.Pl1.p, b.Pg2.p, b.Pl2.p);
        CFK_HIP(hipGetLastError());
        CFK_HIP(hipDeviceSynchronize());  // the plan's buffers are freed on return from this block
    } else {
        if ((rc = b.Pg1.upload(T.P_gain_p1.data(), T.P_gain_p1.size()))) return rc;
        if ((rc = b.Pl1.upload(T.P_loss_p1.data(), T.P_loss_p1.size()))) return rc;
        if ((rc = b.Pg2.upload(T.P_gain_p2.data(), T.P_gain_p2.size()))) return rc;
        if ((rc = b.Pl2.upload(T.P_loss_p2.data(), T.P_loss_p2.size()))) return rc;
    }
    SumfacDev d;
    d.NK = NK;
    d.NL = NL;
    d.NQ = NQ;
    d.L = L;
    d.alpha = alpha;
    d.Nx = (int)G.x.size();
    d.Nu1 = (int)G.u1.size();
    d.Nt1 = (int)G.t1.size();
    d.Ny2 = (int)G.y2.size();
    d.Nt2 = (int)G.t2.size();
    d.Nchi = (int)G.mu_chi.size();
    d.Neps = (int)G.eps.size();
    d.x = b.grid[0].p;
    d.Wx = b.grid[1].p;
    d.u1 = b.grid[2].p;
    d.Wu1 = b.grid[3].p;
    d.t1 = b.grid[4].p;
    d.Wt1 = b.grid[5].p;
    d.y2 = b.grid[6].p;
    d.Wy2 = b.grid[7].p;
    d.t2 = b.grid[8].p;
    d.Wt2 = b.grid[9].p;
    d.mu_chi = b.grid[10].p;
    d.W_chi = b.grid[11].p;
    d.eps = b.grid[12].p;
    d.W_eps = b.grid[13].p;
    d.radnorm = b.radnorm.p;
    d.shnorm = b.shnorm.p;
    d.trip = b.trip.p;
    d.qi_off = b.qi_off.p;
    d.qi = b.qi.p;
    d.Pg1 = b.Pg1.p;
    d.Pl1 = b.Pl1.p;
    d.Pg2 = b.Pg2.p;
    d.Pl2 = b.Pl2.p;
    d.max_block_slots = 1;
    for (int t0 = 0; t0 < NL; t0 += SF_THREADS) {
        const int t1 = std::min(NL, t0 + SF_THREADS);
        d.max_block_slots = std::max(d.max_block_slots, T.qi_off[t1] - T.qi_off[t0]);
    }
    if ((size_t)d.max_block_slots * (sizeof(double) + sizeof(int)) > 64 * 1024)
        return set_error(CFK_EUNSUP, "sumfac: %d valid q_i per block exceed the LDS budget", d.max_block_slots);
    const int64_t nR = (int64_t)NK * NK * NK * NL;
    if ((rc = b.R.alloc((size_t)nR))) return rc;
    CFK_HIP(hipMemset(b.R.p, 0, (size_t)nR * sizeof(double)));
    x_begin = std::max(0, x_begin);
    x_end = std::min(d.Nx, x_end);
    const int64_t Ux = (int64_t)d.Nu1 + (int64_t)d.Ny2 * d.Nt2;
    // The matrix-core form takes each slot's radial factor at the degree of its q_i; it equals the
    // reference's l1 when every valid q_i of a t_chan has degree l1 (true of the tables built from
    // the Gaunt COO).  Other tables run on the per-channel kernel.
    bool slots_on_l1 = true;
    for (int t = 0; t < NL && slots_on_l1; ++t)
        for (int s = T.qi_off[t]; s < T.qi_off[t + 1]; ++s)
            if (l_of_q0(T.qi[s]) != T.trip[3 * t]) slots_on_l1 = false;
    // The exact-sphere form needs the gain rule to be exact for the integrand's degree and every
    // relative speed above the reference's clamp.
    const int kind_req = g_sumfac_kernel;
    bool sphere = false;
    if ((kind_req == CFK_SUMFAC_AUTO || kind_req == CFK_SUMFAC_EXACT_SPHERE) && slots_on_l1 &&
        (L + NK + 1) / 2 <= SP_NODES && d.Nt1 <= MAXG)
        sphere = sphere_rule_exact(G, L + NK - 1) && min_relative_speed2(G, x_begin, x_end) > EPS_SAFE;
    if (kind_req == CFK_SUMFAC_EXACT_SPHERE && !sphere)
        return set_error(CFK_EUNSUP,
                         "sumfac: the exact-sphere kernel needs a (chi, eps) rule exact to degree %d on the sphere, "
                         "every valid q_i of a t_chan of degree l1 and relative speeds above the clamp",
                         L + NK - 1);
    const bool mf = !sphere && (kind_req == CFK_SUMFAC_MATRIX_CORE || kind_req == CFK_SUMFAC_AUTO) && slots_on_l1;
    DevBuf<double> rule_dev;  // SphereRule on the device
    if (sphere) {
        SphereRule rule{};
        gauss_legendre((L + NK + 1) / 2, -1.0, 1.0, rule.t, rule.w);
        if ((rc = rule_dev.upload(rule.t, 2 * SP_NODES))) return rc;
        if (sphere_threads(L, NL) > SP_MAX_THREADS)
            return set_error(CFK_EUNSUP, "sumfac: %d t_chans exceed the exact-sphere kernel's workgroup", NL);
    }
    g_last_timing = SumfacTiming{};
    g_last_timing.kind = sphere ? CFK_SUMFAC_EXACT_SPHERE : mf ? CFK_SUMFAC_MATRIX_CORE : CFK_SUMFAC_PER_CHANNEL;
    hipEvent_t ev[3];
    for (auto &e : ev) CFK_HIP(hipEventCreate(&e));
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            for (int i = 0; i < 3; ++i) (void)hipEventDestroy(e[i]);
        }
    } ev_guard{ev};
    auto lap = [&](double &acc_ms, hipEvent_t a, hipEvent_t b2) -> int {
        CFK_HIP(hipEventSynchronize(b2));
        float ms = 0.f;
        CFK_HIP(hipEventElapsedTime(&ms, a, b2));
        acc_ms += ms;
        return CFK_OK;
    };
    // chunk of x nodes so the per-unit buffers stay below 1 GiB (per-channel kernel) or 4 GiB
    // (matrix-core kernel: more patch-1 units per launch balance the workgroups)
    const int64_t per_unit = (int64_t)NL * 2 * NK + 2 * (L + 1) * NK;
    const int64_t cap = (mf || sphere) ? ((int64_t)1 << 29) : ((int64_t)1 << 27);
    int64_t xs_per_chunk = std::max<int64_t>(1, cap / (per_unit * Ux));
    if (g_sumfac_chunk_nodes > 0) xs_per_chunk = std::min<int64_t>(xs_per_chunk, g_sumfac_chunk_nodes);
    if (x_end > x_begin) {
        const int64_t max_units = std::min<int64_t>(xs_per_chunk, x_end - x_begin) * Ux;
        if ((rc = b.net.alloc((size_t)(max_units * NL * 2 * NK)))) return rc;
        if ((rc = b.rad.alloc((size_t)(max_units * 2 * (L + 1) * NK)))) return rc;
    }
    // class-level fold (matrix-core and exact-sphere forms): row tiles, unit ranges for about 4 waves
    // per SIMD, one partial R per range
    DevBuf<int> fold_rows, fold_l23;
    DevBuf<int2> fold_blk;
    DevBuf<double> fold_part;
    int n_ranges = 1, n_fold_blocks = 0;
    if (mf || sphere) {
        std::vector<int> rows, l23, cls;
        fold_tiles(T, NK, rows, l23, cls);
        const int n_tiles = (int)cls.size();
        n_ranges = fold_ranges(NK, n_tiles);
        const std::vector<int2> items = fold_items(cls, n_ranges);
        n_fold_blocks = (int)items.size();
        if ((rc = fold_rows.upload(rows.data(), rows.size())) || (rc = fold_l23.upload(l23.data(), l23.size())) ||
            (rc = fold_blk.upload(items.data(), items.size())) ||
            (rc = fold_part.alloc((size_t)n_ranges * (size_t)nR)))
            return rc;
    }
    size_t mf_dyn = 0;
    if (mf) {
        const MfLayout lay = mf_layout(L, NK, d.Nt1, d.Nchi, d.Neps);
        if (2 * NQ * NK > lay.geo) return set_error(CFK_EUNSUP, "sumfac: M blocks exceed the operand area");
        if (MF_ITEMS * ((L + 2) / 2) > std::min(MF_SH_LANES, MF_GEO_LANE - MF_SH_LANES))
            return set_error(CFK_EUNSUP, "sumfac: L=%d needs more evaluation lanes", L);
        mf_dyn = (size_t)lay.total * sizeof(double);
        if (mf_dyn + 8 * 1024 > 160 * 1024)
            return set_error(CFK_EUNSUP, "sumfac: %zu bytes of LDS per workgroup exceed the budget", mf_dyn);
    }
    if (x_end > x_begin) {  // equal chunks
        const int64_t n_chunks = (x_end - x_begin + xs_per_chunk - 1) / xs_per_chunk;
        xs_per_chunk = (x_end - x_begin + n_chunks - 1) / n_chunks;
    }
    for (int xc = x_begin; xc < x_end; xc += (int)xs_per_chunk) {
        const int nx = (int)std::min<int64_t>(xs_per_chunk, x_end - xc);
        const int64_t n_units = (int64_t)nx * Ux;
        g_last_timing.chunks += 1;
        if (n_units > 0x7fffffffLL) return set_error(CFK_EINVAL, "sumfac: too many units per chunk");
        if (mf || sphere) {
            switch (NK) {
#define X(KK)                                                                                                    \
    case KK: {                                                                                                   \
        CFK_HIP(hipEventRecord(ev[0], nullptr));                                                                 \
        if (sphere) {                                                                                            \
            const SphereRule *rp = reinterpret_cast<const SphereRule *>(rule_dev.p);                             \
            const dim3 blk((unsigned)sphere_threads(L, NL));                                                     \
            if (nx * d.Nu1 > 0)                                                                                  \
                hipLaunchKernelGGL((sumfac_units_sphere_kernel<KK, true>), dim3((unsigned)(nx * d.Nu1)), blk, 0,  \
                                   nullptr, d, rp, xc, b.net.p, b.rad.p);                                        \
            CFK_HIP(hipGetLastError());                                                                          \
            if (nx * d.Ny2 * d.Nt2 > 0)                                                                          \
                hipLaunchKernelGGL((sumfac_units_sphere_kernel<KK, false>), dim3((unsigned)(nx * d.Ny2 * d.Nt2)), \
                                   blk, 0, nullptr, d, rp, xc, b.net.p, b.rad.p);                                \
            CFK_HIP(hipGetLastError());                                                                          \
        } else if ((rc = launch_units_mf<KK>(mf_tiles_per_wave(L), mf_dyn, d, xc, nx, n_units, b.net.p,         \
                                             b.rad.p)))                                                          \
            return rc;                                                                                           \
        CFK_HIP(hipEventRecord(ev[1], nullptr));                                                                 \
        const int64_t per_range = (n_units + n_ranges - 1) / n_ranges;                                          \
        hipLaunchKernelGGL(sumfac_fold_tiles_kernel<KK>, dim3((unsigned)n_fold_blocks), dim3(64), 0, nullptr, NL, \
                           L, fold_rows.p, fold_l23.p, fold_blk.p, n_units, per_range, b.net.p, b.rad.p,         \
                           fold_part.p);                                                                          \
        CFK_HIP(hipGetLastError());                                                                              \
        hipLaunchKernelGGL(sumfac_reduce_kernel, dim3((unsigned)((nR + 255) / 256)), dim3(256), 0, nullptr, nR,    \
                           n_ranges, fold_part.p, b.R.p);                                                        \
        CFK_HIP(hipGetLastError());                                                                              \
        CFK_HIP(hipEventRecord(ev[2], nullptr));                                                                 \
        break;                                                                                                   \
    }
                X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
#undef X
            }
            if ((rc = lap(g_last_timing.units_ms, ev[0], ev[1]))) return rc;
            if ((rc = lap(g_last_timing.fold_ms, ev[1], ev[2]))) return rc;
            g_last_timing.units += n_units;
            continue;
        }
        dim3 grid((unsigned)n_units, (unsigned)((NL + SF_THREADS - 1) / SF_THREADS));
        const size_t dyn = (size_t)d.max_block_slots * (sizeof(double) + sizeof(int));
        CFK_HIP(hipEventRecord(ev[0], nullptr));
        switch (NK) {
#define X(KK)                                                                                                   \
    case KK:                                                                                                    \
        hipLaunchKernelGGL(sumfac_units_kernel<KK>, grid, dim3(SF_THREADS), dyn, nullptr, d, xc, n_units, b.net.p, \
                           b.rad.p);                                                                            \
        break;
            X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
#undef X
        }
        CFK_HIP(hipGetLastError());
        CFK_HIP(hipEventRecord(ev[1], nullptr));
        hipLaunchKernelGGL(sumfac_fold_kernel, dim3((unsigned)((nR + 255) / 256)), dim3(256), 0, nullptr, NK, NL, L,
                           b.trip.p, n_units, b.net.p, b.rad.p, b.R.p);
        CFK_HIP(hipGetLastError());
        CFK_HIP(hipEventRecord(ev[2], nullptr));
        if ((rc = lap(g_last_timing.units_ms, ev[0], ev[1]))) return rc;
        if ((rc = lap(g_last_timing.fold_ms, ev[1], ev[2]))) return rc;
        g_last_timing.units += n_units;
    }
    if (zero_conservation) {
        hipLaunchKernelGGL(conservation_zero_kernel, dim3((unsigned)((nR + 255) / 256)), dim3(256), 0, nullptr, NK, NL,
                           b.trip.p, b.R.p);
        CFK_HIP(hipGetLastError());
    }
    CFK_HIP(hipMemcpy(R_host, b.R.p, (size_t)nR * sizeof(double), hipMemcpyDeviceToHost));
    return CFK_OK;
}

}  // namespace
}  // namespace cfk

using namespace cfk;

extern "C" int cfk_compute_rtensor_sumfac(
    double *R_out, int32_t K_max, int32_t N_L, int32_t N_Q, double alpha, const double *x_nodes, const double *W_x,
    int32_t N_x, const double *u1_nodes, const double *W_u1, int32_t N_u1, const double *t1_nodes, const double *W_t1,
    int32_t N_t1, const double *y2_nodes, const double *W_y2, int32_t N_y2, const double *t2_nodes, const double *W_t2,
    int32_t N_t2, const double *mu_chi, const double *W_chi, int32_t N_chi, const double *eps_vec, const double *W_eps,
    int32_t N_eps, const double *RadialNorm, const double *SH_Norm, const double *L_triplets,
    const double *qi_valid_mat, const double *P_loss_p1, const double *P_gain_p1, const double *P_loss_p2,
    const double *P_gain_p2, int32_t x_begin, int32_t x_end) {
    if (!R_out) return set_error(CFK_EINVAL, "null R_out");
    if (K_max < 0 || N_L < 1 || N_Q < 1) return set_error(CFK_EINVAL, "bad K_max/N_L/N_Q");
    const int32_t ns[7] = {N_x, N_u1, N_t1, N_y2, N_t2, N_chi, N_eps};
    for (int i = 0; i < 7; ++i)
        if (ns[i] < 1) return set_error(CFK_EINVAL, "empty quadrature grid %d", i);
    const double *ptrs[] = {x_nodes, W_x, u1_nodes, W_u1, t1_nodes, W_t1, y2_nodes, W_y2, t2_nodes, W_t2,
                            mu_chi, W_chi, eps_vec, W_eps, RadialNorm, SH_Norm, L_triplets, qi_valid_mat,
                            P_loss_p1, P_gain_p1, P_loss_p2, P_gain_p2};
    for (const double *p : ptrs)
        if (!p) return set_error(CFK_EINVAL, "null input array");
    std::string err;
    SumfacTables T = compact_mex_tables(N_Q, N_L, L_triplets, qi_valid_mat, N_u1, N_t1, N_y2, P_loss_p1, P_gain_p1,
                                        P_loss_p2, P_gain_p2, err);
    if (!err.empty()) return set_error(CFK_EINVAL, "%s", err.c_str());
    const int NK = K_max + 1;
    // The MEX takes RadialNorm/SH_Norm as inputs; the device code recomputes them from the same
    // closed forms, so check that the caller's tables agree.
    std::vector<double> rn((size_t)NK * (T.L + 1)), sn(N_Q);
    radial_norm(NK, T.L, rn.data());
    sh_norm(T.L, sn.data());
    for (size_t i = 0; i < rn.size(); ++i)
        if (std::fabs(rn[i] - RadialNorm[i]) > 1e-13 * std::fabs(rn[i]))
            return set_error(CFK_EINVAL, "RadialNorm differs from its closed form at %zu", i);
    for (int i = 0; i < N_Q; ++i)
        if (std::fabs(sn[i] - SH_Norm[i]) > 1e-13 * std::fabs(sn[i]))
            return set_error(CFK_EINVAL, "SH_Norm differs from its closed form at %d", i);
    SumfacGrids G;
    G.x.assign(x_nodes, x_nodes + N_x);
    G.W_x.assign(W_x, W_x + N_x);
    G.u1.assign(u1_nodes, u1_nodes + N_u1);
    G.W_u1.assign(W_u1, W_u1 + N_u1);
    G.t1.assign(t1_nodes, t1_nodes + N_t1);
    G.W_t1.assign(W_t1, W_t1 + N_t1);
    G.y2.assign(y2_nodes, y2_nodes + N_y2);
    G.W_y2.assign(W_y2, W_y2 + N_y2);
    G.t2.assign(t2_nodes, t2_nodes + N_t2);
    G.W_t2.assign(W_t2, W_t2 + N_t2);
    G.mu_chi.assign(mu_chi, mu_chi + N_chi);
    G.W_chi.assign(W_chi, W_chi + N_chi);
    G.eps.assign(eps_vec, eps_vec + N_eps);
    G.W_eps.assign(W_eps, W_eps + N_eps);
    return run_sumfac(T, G, NK, alpha, x_begin, x_end, false, R_out);
}

extern "C" int cfk_sumfac_set_kernel(int32_t kind) {
    if (kind != CFK_SUMFAC_MATRIX_CORE && kind != CFK_SUMFAC_PER_CHANNEL && kind != CFK_SUMFAC_EXACT_SPHERE &&
        kind != CFK_SUMFAC_AUTO)
        return set_error(CFK_EINVAL, "unknown sumfac kernel %d", kind);
    g_sumfac_kernel = kind;
    return CFK_OK;
}

extern "C" int cfk_sumfac_set_chunk_nodes(int32_t max_nodes) {
    if (max_nodes < 0) return set_error(CFK_EINVAL, "sumfac chunk: %d x nodes", max_nodes);
    g_sumfac_chunk_nodes = max_nodes;
    return CFK_OK;
}

extern "C" int cfk_sumfac_last_chunks(int64_t *n_chunks) {
    if (!n_chunks) return set_error(CFK_EINVAL, "null pointer");
    *n_chunks = g_last_timing.chunks;
    return CFK_OK;
}

extern "C" int cfk_sumfac_last_timing(int32_t *kind, int64_t *n_units, double *units_ms, double *fold_ms) {
    if (kind) *kind = g_last_timing.kind;
    if (n_units) *n_units = g_last_timing.units;
    if (units_ms) *units_ms = g_last_timing.units_ms;
    if (fold_ms) *fold_ms = g_last_timing.fold_ms;
    return CFK_OK;
}

extern "C" int cfk_build_rtensor(double *R_out, int32_t K_max, int32_t L, double alpha, const int32_t grid_sizes[7],
                                 const double *labels, const double *vals, const double *ic_map, int64_t N_G) {
    if (!R_out || !grid_sizes || !labels || !vals || !ic_map) return set_error(CFK_EINVAL, "null pointer");
    if (K_max < 0 || L < 0) return set_error(CFK_EINVAL, "bad K_max/L");
    for (int i = 0; i < 7; ++i)
        if (grid_sizes[i] < 1) return set_error(CFK_EINVAL, "grid size %d < 1", i);
    const int NQ = (L + 1) * (L + 1);
    RealGaunt g;
    g.L = L;
    int n_l = 0;
    for (int64_t z = 0; z < N_G; ++z) {
        for (int c = 0; c < 3; ++c) {
            const double v = labels[z + c * N_G];
            if (!(v >= 1 && v <= NQ) || v != std::floor(v)) return set_error(CFK_EINVAL, "label out of range");
        }
        if (!(ic_map[z] >= 1) || ic_map[z] != std::floor(ic_map[z])) return set_error(CFK_EINVAL, "bad ic_map");
        g.q1.push_back((int)labels[z]);
        g.q2.push_back((int)labels[z + N_G]);
        g.q3.push_back((int)labels[z + 2 * N_G]);
        g.val.push_back(vals[z]);
        g.t.push_back((int)ic_map[z]);
        n_l = std::max(n_l, (int)ic_map[z]);
    }
    g.n_l = n_l;
    std::vector<int> seen(n_l, 0);
    for (int t : g.t) seen[t - 1] = 1;
    for (int t = 0; t < n_l; ++t)
        if (!seen[t]) return set_error(CFK_EINVAL, "ic_map has no entry for t=%d", t + 1);
    SumfacGrids G;
    try {
        G = sumfac_grids(alpha, grid_sizes);
    } catch (const std::exception &e) {
        return set_error(CFK_EINVAL, "%s", e.what());
    }
    const SumfacTablePlan P = sumfac_table_plan(g);
    return run_sumfac(P.T, G, K_max + 1, alpha, 0, grid_sizes[0], true, R_out, &P);
}
