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

__global__ void conservation_zero_kernel(int NK, int NL, const int *__restrict__ trip, double *__restrict__ R) {
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = (int64_t)NK * NK * NK * NL;
    if (o >= total) return;
    const int k1 = (int)(o % NK);
    const int tc = (int)(o / ((int64_t)NK * NK * NK));
    const int l1 = trip[3 * tc];
    if ((l1 == 0 && k1 <= 1) || (l1 == 1 && k1 == 0)) R[o] = 0.0;
}

struct SumfacBuffers {
    DevBuf<double> grid[14], radnorm, shnorm, Pg1, Pl1, Pg2, Pl2, net, rad, R;
    DevBuf<int> trip, qi_off, qi;
};

int run_sumfac(const SumfacTables &T, const SumfacGrids &G, int NK, double alpha, int x_begin, int x_end,
               bool zero_conservation, double *R_host) {
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
    if ((rc = b.Pg1.upload(T.P_gain_p1.data(), T.P_gain_p1.size()))) return rc;
    if ((rc = b.Pl1.upload(T.P_loss_p1.data(), T.P_loss_p1.size()))) return rc;
    if ((rc = b.Pg2.upload(T.P_gain_p2.data(), T.P_gain_p2.size()))) return rc;
    if ((rc = b.Pl2.upload(T.P_loss_p2.data(), T.P_loss_p2.size()))) return rc;
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
    // chunk of x nodes so the per-unit buffers stay below ~1 GiB
    const int64_t per_unit = (int64_t)NL * 2 * NK + 2 * (L + 1) * NK;
    int64_t xs_per_chunk = std::max<int64_t>(1, (int64_t)(1 << 27) / (per_unit * Ux));
    if (x_end > x_begin) {
        const int64_t max_units = std::min<int64_t>(xs_per_chunk, x_end - x_begin) * Ux;
        if ((rc = b.net.alloc((size_t)(max_units * NL * 2 * NK)))) return rc;
        if ((rc = b.rad.alloc((size_t)(max_units * 2 * (L + 1) * NK)))) return rc;
    }
    for (int xc = x_begin; xc < x_end; xc += (int)xs_per_chunk) {
        const int nx = (int)std::min<int64_t>(xs_per_chunk, x_end - xc);
        const int64_t n_units = (int64_t)nx * Ux;
        if (n_units > 0x7fffffffLL) return set_error(CFK_EINVAL, "sumfac: too many units per chunk");
        dim3 grid((unsigned)n_units, (unsigned)((NL + SF_THREADS - 1) / SF_THREADS));
        const size_t dyn = (size_t)d.max_block_slots * (sizeof(double) + sizeof(int));
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
        hipLaunchKernelGGL(sumfac_fold_kernel, dim3((unsigned)((nR + 255) / 256)), dim3(256), 0, nullptr, NK, NL, L,
                           b.trip.p, n_units, b.net.p, b.rad.p, b.R.p);
        CFK_HIP(hipGetLastError());
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
    SumfacTables T = sumfac_tables(g, G);
    return run_sumfac(T, G, K_max + 1, alpha, 0, grid_sizes[0], true, R_out);
}
