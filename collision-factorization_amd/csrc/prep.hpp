// Host-side preparation of the collision tables (native replacement of the reference's MATLAB
// host code).  Declarations shared by geometry.cpp and the device code.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace cfk {

struct RealGaunt {
    int L = 0;
    // 1-based, MATLAB find order (sorted by q3, then q2, then q1)
    std::vector<int32_t> q1, q2, q3;
    std::vector<double> val;
    std::vector<int32_t> t;        // 1-based l-triplet rank (lexicographic (l1,l2,l3))
    int n_l = 0;
};

// src/GeneralCollisionTensor.m:251-281 (+ src/SHL/gaunt_generate_labels.m, gaunt_compute_values.m,
// complex2real.m)
RealGaunt real_gaunt(int L);

// src/Gauss.m:38-105
void gauss_legendre(int n, double a, double b, double *x, double *w);
// src/Gauss.m:128-173
void gauss_gen_laguerre(int n, double alpha, double *x, double *w);

// src/GeneralCollisionTensor.m:55-64
void sumfac_grid_sizes(int K_max, int L, int radial_pad, int angular_pad, int sizes[7]);

inline int l_of_q0(int q0) {  // 0-based q -> l
    int l = 0;
    while ((l + 1) * (l + 1) <= q0) ++l;
    return l;
}

// Associated Legendre P_l^m (Condon-Shortley phase), for 0 <= m <= l.
double legendre_plm(int l, int m, double x);
// Real spherical harmonic normalisation SH_Norm[q] (src/GeneralCollisionTensor.m:92-104)
void sh_norm(int L, double *out);
// RadialNorm[k + N_K*l] (src/GeneralCollisionTensor.m:82-90)
void radial_norm(int N_K, int L, double *out);
// Radial nodes r_a = sqrt(x_a) (n_rad generalized Gauss-Laguerre nodes, alpha = 1/2, in x = |v|^2) and
// rad[(k (L+1) + l) n_rad + a] = R_{k,l}(r_a) w_a / 2, the radial factor of the tutorials' projection.
void projection_radial_table(int L, int K, int n_rad, std::vector<double> &r, std::vector<double> &rad);

// The 1-D rules of the sumfac (src/GeneralCollisionTensor.m:66-79)
struct SumfacGrids {
    std::vector<double> x, W_x, u1, W_u1, t1, W_t1, y2, W_y2, t2, W_t2, mu_chi, W_chi, eps, W_eps;
};
SumfacGrids sumfac_grids(double alpha, const int sizes[7]);

// Compact sumfac tables.  The reference's P_gain tables are dense over all (q_i, t) columns;
// only the valid q_i of each t are ever read (compute_rtensor_sumfac_mex.cpp:219-227), so they
// are stored per (t, valid slot j) here.
struct SumfacTables {
    int n_l = 0, N_Q = 0, L = 0;
    std::vector<int32_t> trip;        // [n_l][3] (l1,l2,l3)
    std::vector<int32_t> qi_off;      // [n_l+1] offsets into the compact valid list
    std::vector<int32_t> qi;          // 0-based q_i, in the reference's qi_valid_mat order
    // Device layouts, fastest index last: patch 1 gain [N_u1][N_t1][slot], loss [N_u1][N_t1][t];
    // patch 2 gain [N_y2][slot], loss [N_y2][t].  For one quadrature step all slots (t_chans) are
    // contiguous, so a workgroup loads its share with coalesced reads.
    std::vector<double> P_gain_p1, P_loss_p1, P_gain_p2, P_loss_p2;
};

// Recipe for building the compact P tables on the device (src/GeneralCollisionTensor.m:106-174):
// the slot structure of SumfacTables plus, per slot, its Gaunt entries (q_k, G) in member order and,
// per t_chan, the loss entries (q_k, G Y_i0) and SCALE.  With Y = real SH at phi = 0 on a grid node,
//   P_gain[node][slot] = SCALE_t * sum_{e in slot} G_e Y[q_k(e)],
//   P_loss[node][t]    = SCALE_t * sum_{e in loss(t)} (G Y_i0)_e Y[q_k(e)].
struct SumfacTablePlan {
    SumfacTables T;                   // trip, qi_off, qi (P vectors left empty)
    std::vector<int32_t> slot_t;      // [ns] t_chan of each slot
    std::vector<int32_t> ent_off;     // [ns+1]
    std::vector<int32_t> ent_qk;      // 0-based q_k
    std::vector<double> ent_g;
    std::vector<int32_t> loss_off;    // [n_l+1]
    std::vector<int32_t> loss_qk;
    std::vector<double> loss_g;
    std::vector<double> scale;        // [n_l]
};
SumfacTablePlan sumfac_table_plan(const RealGaunt &g);

// Compact tables from the reference's dense MEX inputs (the cfk_compute_rtensor_sumfac path).
SumfacTables compact_mex_tables(int N_Q, int n_l, const double *L_triplets, const double *qi_valid_mat,
                                int N_u1, int N_t1, int N_y2, const double *P_loss_p1, const double *P_gain_p1,
                                const double *P_loss_p2, const double *P_gain_p2, std::string &err);

}  // namespace cfk
