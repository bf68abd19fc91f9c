/*
 * cfk.h -- C ABI of the MI355X collision-factorization library (libcfk.so, gfx950).
 *
 * The library implements the Wigner-Eckart-factorized spectral Boltzmann collision hot path:
 *   - the Gaunt-COO-routed quadratic contraction Q[q1,k1] += sum A f f, and
 *   - the 5-D Duffy sum-factorized quadrature that produces the radial tensor R.
 *
 * Conventions (identical to the reference MEX interface):
 *   - one distribution f / Q is an N_Q x K column-major array: f[q + k*N_Q]
 *     (q = l^2+l+m, real spherical harmonics; k = Laguerre degree);
 *   - R is K x K x K x N_L column-major: R[k1 + k2*K + k3*K^2 + t*K^3];
 *   - Gaunt COO: labels is N_G x 3 column-major (q1, q2, q3), vals N_G, ic_map N_G (t),
 *     all indices 1-based and stored as double, exactly as MATLAB hands them to a MEX;
 *   - the contraction entry points ACCUMULATE into Q (Q += Q(f)), as the reference MEX do.
 *
 * Every function returns CFK_OK (0) or a negative status; cfk_last_error() returns the
 * message of the last failure on the calling thread.  Nothing here ever falls back to a CPU
 * computation of the contraction or the quadrature: without a usable HIP device the compute
 * entry points fail with CFK_ENODEV.
 *
 * Batched device API: a distribution batch is B consecutive N_Q x K arrays ([B][K][N_Q],
 * CFK_LAYOUT_NATURAL) or the batch-fastest array [K][N_Q][B] (CFK_LAYOUT_BATCH_FASTEST).
 * Device pointers must come from hipMalloc (or torch) on the current device.
 */
#ifndef CFK_H
#define CFK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CFK_OK 0
#define CFK_EINVAL (-1)   /* bad argument: size, index out of range, null pointer, unsorted input */
#define CFK_EHIP (-2)     /* HIP runtime error */
#define CFK_ENOMEM (-3)   /* device or host allocation failed */
#define CFK_ENODEV (-4)   /* no usable HIP device */
#define CFK_EUNSUP (-5)   /* combination not supported (e.g. K too large for a variant) */

/* contraction variants (the four reference MEX kernels) */
#define CFK_VARIANT_ANGULAR_FIRST 0
#define CFK_VARIANT_RADIAL_FIRST 1
#define CFK_VARIANT_NAIVE 2

/* precision of a plan execution */
#define CFK_PREC_FP64 0   /* plain FP64 operator (parity with the reference at 1e-12) */
#define CFK_PREC_MIXED 1     /* FP32 symmetrized contraction, FP64 accumulation + conservation projection */
#define CFK_PREC_FP64_SYM 2  /* FP64 over the symmetrized tables (q2<=q3 half of the routing table) */

/* batch layouts */
#define CFK_LAYOUT_NATURAL 0        /* [B][K][N_Q]  (B MATLAB N_Q x K arrays back to back) */
#define CFK_LAYOUT_BATCH_FASTEST 1  /* [K][N_Q][B]  (transpose of natural: cells fastest, as a spatial solver stores them) */

const char *cfk_last_error(void);
int cfk_version(void);
int cfk_device_count(void);

/* ---------------------------------------------------------------------------------------
 * MEX-equivalent one-shot entry points: host pointers, one distribution, MATLAB layout.
 * A MEX gateway forwards its prhs[] here unchanged (mxGetPr of each argument, in order).
 * R_numel is mxGetNumberOfElements(prhs[5]) and bounds ic_map (the reference does not check).
 * ------------------------------------------------------------------------------------- */

/* replaces angular_first_collision_kernel_mex (src/mex/angular_first_collision_kernel_mex.cpp:6-76) */
int cfk_angular_first_collision_kernel(double *Q, const double *f, const double *labels, const double *vals,
                                       const double *ic_map, const double *R, int64_t N_G, int64_t R_numel,
                                       int32_t N_Q, int32_t K);
/* replaces radial_first_collision_kernel_mex (src/mex/radial_first_collision_kernel_mex.cpp:5-77) */
int cfk_radial_first_collision_kernel(double *Q, const double *f, const double *labels, const double *vals,
                                      const double *ic_map, const double *R, int64_t N_G, int64_t R_numel,
                                      int32_t N_Q, int32_t K);
/* replaces naive_collision_kernel_mex (src/mex/naive_collision_kernel_mex.cpp:4-47) */
int cfk_naive_collision_kernel(double *Q, const double *f, const double *labels, const double *vals,
                               const double *ic_map, const double *R, int64_t N_G, int64_t R_numel,
                               int32_t N_Q, int32_t K);
/* replaces dense_tensor_kernel_mex (src/mex/dense_tensor_kernel_mex.cpp:4-31): Q[i] += C[i,j,k] f[j] f[k] */
int cfk_dense_tensor_kernel(double *Q, const double *f, const double *C, int64_t N);

/* replaces compute_rtensor_sumfac_mex (src/mex/compute_rtensor_sumfac_mex.cpp:80-398).
 * The 26 MEX inputs in MEX order, each array followed by its element count where the MEX reads
 * one with mxGetNumberOfElements.  R_out (N_K^3*N_L doubles, N_K = K_max+1) is OVERWRITTEN with
 * the contribution of the x nodes [x_begin, x_end) (0, N_x for the full tensor), as the MEX
 * returns a fresh array.  No conservation zeroing (the MATLAB host applies it afterwards).
 * The quadrature runs on the kernel chosen by cfk_sumfac_set_kernel. */
int cfk_compute_rtensor_sumfac(double *R_out, int32_t K_max, int32_t N_L, int32_t N_Q, double alpha,
                               const double *x_nodes, const double *W_x, int32_t N_x,
                               const double *u1_nodes, const double *W_u1, int32_t N_u1,
                               const double *t1_nodes, const double *W_t1, int32_t N_t1,
                               const double *y2_nodes, const double *W_y2, int32_t N_y2,
                               const double *t2_nodes, const double *W_t2, int32_t N_t2,
                               const double *mu_chi, const double *W_chi, int32_t N_chi,
                               const double *eps_vec, const double *W_eps, int32_t N_eps,
                               const double *RadialNorm, const double *SH_Norm, const double *L_triplets,
                               const double *qi_valid_mat, const double *P_loss_p1, const double *P_gain_p1,
                               const double *P_loss_p2, const double *P_gain_p2, int32_t x_begin, int32_t x_end);

/* Kernel for the sumfac quadrature (process-wide; set it before the calls it should govern).
 * CFK_SUMFAC_AUTO (default): CFK_SUMFAC_EXACT_SPHERE where it applies, else CFK_SUMFAC_MATRIX_CORE
 *   (else CFK_SUMFAC_PER_CHANNEL for tables whose valid q_i are not all of degree l1).
 * CFK_SUMFAC_EXACT_SPHERE: when the (chi, eps) product rule integrates polynomials of degree
 *   L + K_max over the sphere exactly (the degree of the gain integrand there; the reference's own
 *   grid sizing, GeneralCollisionTensor.m:63-64, and larger Gauss-Legendre x uniform rules qualify),
 *   the gain sum of one unit step equals its surface
 *   integral, which reduces by rotational symmetry about the collision's centre of mass to the SH at
 *   one direction per side and one-dimensional Gauss-Legendre integrals, exact for the polynomial
 *   integrand; fails with CFK_EUNSUP where the rule is not exact.
 * CFK_SUMFAC_MATRIX_CORE: per inner-t step, the gain-point sum is taken once per SH degree on FP64
 *   matrix cores (M(l,k,q) = sum_p w Rad(p,l,k) Y(p,q)) and then contracted with each t_chan's
 *   P_gain slots.
 * CFK_SUMFAC_PER_CHANNEL: one lane per t_chan sums over the gain points in the reference's order
 *   (compute_rtensor_sumfac_mex.cpp:219-234), on the vector ALUs.
 * All are FP64 and agree with the reference to rounding; the unit fold of the first three is a
 * matrix-core product over units.  No counterpart in the reference (one code path there). */
#define CFK_SUMFAC_MATRIX_CORE 0
#define CFK_SUMFAC_PER_CHANNEL 1
#define CFK_SUMFAC_EXACT_SPHERE 2
#define CFK_SUMFAC_AUTO 3
int cfk_sumfac_set_kernel(int32_t kind);
/* Largest number of x nodes the sumfac processes per chunk (one stage-A and one fold launch over
 * per-unit buffers of that many nodes; process-wide).  0 (the default) sizes chunks by the unit
 * buffers' memory cap (at L=12, K=12, 48 nodes per axis: 4 chunks of 12 nodes).  Results do not
 * depend on the chunking beyond rounding (the fold adds each chunk's units into R).  No reference
 * counterpart (the reference's OpenMP threads each take x nodes, compute_rtensor_sumfac_mex.cpp:134). */
int cfk_sumfac_set_chunk_nodes(int32_t max_nodes);
/* Number of x-node chunks of the last sumfac call. */
int cfk_sumfac_last_chunks(int64_t *n_chunks);
/* Device time of the last sumfac call: the kernel that ran (tables whose valid q_i are not all of
 * degree l1 run per channel), the units processed, and the summed HIP-event durations of the
 * unit (stage A) and fold (stage B) launches.  Any pointer may be NULL.  No reference counterpart. */
int cfk_sumfac_last_timing(int32_t *kind, int64_t *n_units, double *units_ms, double *fold_ms);

/* ---------------------------------------------------------------------------------------
 * Host-side preparation (the MATLAB host code of the reference, native here).
 * ------------------------------------------------------------------------------------- */

/* Real Gaunt COO for degree L (src/GeneralCollisionTensor.m:251-281 with src/SHL/gaunt_*.m and
 * complex2real.m).  Two-call pattern: pass NULL arrays to get *N_G; then arrays of N_G*3, N_G, N_G
 * doubles (1-based, MATLAB find order: sorted by (q3,q2,q1)). */
int cfk_real_gaunt(int32_t L, int64_t *N_G, double *labels, double *vals, double *ic_map);

/* Gauss-Legendre on [a,b] (src/Gauss.m:38-105) and generalized Gauss-Laguerre for x^alpha e^-x
 * (src/Gauss.m:128-173); ascending nodes. */
int cfk_gauss_legendre(int32_t n, double a, double b, double *x, double *w);
int cfk_gauss_gen_laguerre(int32_t n, double alpha, double *x, double *w);

/* Grid sizes (N_x, N_u1, N_t1, N_y2, N_t2, N_chi, N_eps) for given paddings
 * (src/GeneralCollisionTensor.m:55-64). */
int cfk_sumfac_grid_sizes(int32_t K_max, int32_t L, int32_t radial_pad, int32_t angular_pad, int32_t sizes[7]);

/* Build R [N_K,N_K,N_K,N_L] on the GPU end to end: P tables (src/GeneralCollisionTensor.m:106-174,
 * compacted to the valid q_i; the slot structure on the host, the tables on the device), the
 * sumfac quadrature, and the conservation
 * zeroing (src/GeneralCollisionTensor.m:183-191).  grid_sizes as from cfk_sumfac_grid_sizes (or
 * chosen explicitly); labels/vals/ic_map as returned by cfk_real_gaunt. */
int cfk_build_rtensor(double *R_out, int32_t K_max, int32_t L, double alpha, const int32_t grid_sizes[7],
                      const double *labels, const double *vals, const double *ic_map, int64_t N_G);

/* Read one float64 dataset of an HDF5 file — a MATLAB v7.3 .mat file is one, such as the
 * reference's precalc tensors (the .mat files in src/precalc, written by precompute_collision_operator.m:61;
 * the TensorObj properties are datasets under "/#refs#").  `dataset` is a '/'-separated path.
 * dims[0..rank-1] are the HDF5 dimensions (C order), i.e. a MATLAB array's dimensions reversed, so
 * `out` holds the array in MATLAB's column-major order.  Two-call pattern: out = NULL returns
 * numel, rank and dims only.  Supported: the HDF5 subset MATLAB writes (superblock 0/1, symbol-table
 * groups, v1 object headers, compact/contiguous/chunked layouts, deflate and shuffle filters);
 * anything else returns CFK_EUNSUP.  Host-only; no reference counterpart (MATLAB's load). */
int cfk_h5_read_f64(const char *path, const char *dataset, double *out, int64_t capacity, int64_t *numel,
                    int32_t *rank, int64_t dims[8]);

/* Write n_vars float64 arrays as the top-level variables of a MATLAB v7.3 .mat file (an HDF5 file
 * behind the 512-byte MAT-file header), the format precompute_collision_operator.m:61 saves with
 * save(..., '-v7.3') and MATLAB's load reads: names[i] distinct MATLAB identifiers, ranks[i] the
 * number of MATLAB dimensions of variable i (1..8), dims those dimensions for all variables back to back,
 * data[i] the column-major values.  Stored uncompressed and contiguous with MATLAB_class "double";
 * at most 256 variables; cfk_h5_read_f64 reads them back.  Plain variables, not MATLAB class objects:
 * the reference saves its TensorObj, this writes the arrays a MATLAB host rebuilds it from.  Host only. */
int cfk_mat_write_f64(const char *path, int32_t n_vars, const char *const *names, const int32_t *ranks,
                      const int64_t *dims, const double *const *data);

/* ---------------------------------------------------------------------------------------
 * Device-resident plans: tables uploaded once, many batched executions.
 * ------------------------------------------------------------------------------------- */
typedef struct cfk_plan_s *cfk_plan;

/* labels/vals/ic_map: host COO (any order; it is stably sorted by (t,q1) as
 * benchmark_collision_contraction_scaling.m:82-85 does).  R: host R tensor (R_numel doubles). */
int cfk_plan_create(const double *labels, const double *vals, const double *ic_map, int64_t N_G,
                    const double *R, int64_t R_numel, int32_t N_Q, int32_t K, cfk_plan *out);
int cfk_plan_destroy(cfk_plan plan);

/* Q_dev += Q(f_dev) for B distributions on `stream` (hipStream_t, NULL = default).
 * Asynchronous.  variant: CFK_VARIANT_*; precision: CFK_PREC_*; layout: CFK_LAYOUT_*.
 * B = 0 and a plan with no Gaunt entries leave Q unchanged.  One call takes at most
 * min((2^32 - 1) / (8 N_Q K), (2^31 - 1) / (4 N_Q KP)) distributions, KP = K rounded up to a
 * multiple of 4 (f addressed through 32-bit buffer offsets; 552,336 at L=8, K=10); larger batches
 * return CFK_EINVAL before any launch and are split by the caller. */
int cfk_plan_execute(cfk_plan plan, int32_t variant, int32_t precision, int32_t layout, double *Q_dev,
                     const double *f_dev, int64_t B, void *stream);

/* Fixed device area, per GPU, of cfk_plan_execute_static. */
#define CFK_STATIC_SCRATCH_BYTES ((size_t)64 * 1024 * 1024)

/* cfk_plan_execute without any allocation: no hipMalloc, hipFree or allocating library call, so it may
 * be used where a caller cannot allocate -- inside HIP graph capture, in tight serving loops, under
 * harnesses that forbid allocation.  No counterpart in the reference (MATLAB allocates every MEX
 * output itself).  Same arguments, kernels and results as cfk_plan_execute; the staging it needs (the
 * mixed path's FP32 copy of f, and for CFK_LAYOUT_NATURAL the batch-fastest FP64 copies of f and Q)
 * lives in one fixed device array of CFK_STATIC_SCRATCH_BYTES per GPU, held until the process ends;
 * a call whose staging exceeds it returns CFK_EINVAL before launching anything (64 MiB covers L=8,
 * K=10 at 4096 distributions in every precision and layout).  The batch-fastest FP64 precisions need no
 * staging.  Calls that overlap on one GPU (several streams or threads) race on the shared array.
 * Kernel timing (cfk_plan_set_timing) is not recorded by this entry point. */
int cfk_plan_execute_static(cfk_plan plan, int32_t variant, int32_t precision, int32_t layout, double *Q_dev,
                            const double *f_dev, int64_t B, void *stream);

/* Autotuner: times every tile variant of the angular-first kernels -- k3 chunk width KC in
 * {K, ceil(K/2)} x distributions per lane NB in {1,2} x a third switch PF: for FP64 the prefetch of
 * the next nonzero's rows, for MIXED the radial stage on the matrix cores (v_mfma_f32_16x16x4_f32
 * with the operands moved between lanes by register transposes; its own chunk: KC = K up to K = 10,
 * ceil(K/2) above; NB = 1) -- for each precision on a synthetic batch of B distributions
 * (batch-fastest layout), each also at 1, 2 and 4 waves per workgroup (cfk_plan_set_waves) and, for
 * the symmetrized precisions, with the row split (cfk_plan_set_split), on
 * `stream` and keeps the fastest per precision in the plan.  When
 * tiles_out is not NULL it receives 9 ints: (KC, NB, PF) for FP64, MIXED, FP64_SYM in that order.
 * Allocates temporary device buffers; synchronizes `stream`.  The full variant set exists for K in
 * {6,8,10,12,14}; other K run the default tile. */
int cfk_plan_autotune(cfk_plan plan, int64_t B, void *stream, int32_t *tiles_out);

/* Linearized operator J = dQ/dc at f0 (host, N_Q x K MATLAB order; NULL = the absolute Maxwellian
 * c_000 = 1), written to J (host, (N_Q K) x (N_Q K) column-major, index k*N_Q + q): the reference's
 * J = C(:,:,1) + C(:,1,:) for f0 = e_000 (tests/TestWCUEigenvalues.m:37,
 * tutorial_stress_relaxation.m:51).  Computed exactly with the plan's FP64 operator on the GPU as
 * J e_j = (Q(f0 + e_j) - Q(f0 - e_j)) / 2.  Synchronizes `stream`. */
int cfk_plan_linearize(cfk_plan plan, const double *f0, double *J, void *stream);

/* Sets one precision's tile (KC, NB, PF as above; KC = 0 and NB = 0 select the defaults) without
 * timing anything.  For MIXED, PF = -1 (the default of a new plan) selects the kernel by K: the
 * matrix-core radial stage from K = 12 up (with the row split at K = 12), the VALU kernel below. */
int cfk_plan_set_tile(cfk_plan plan, int32_t precision, int32_t kc, int32_t nb, int32_t pf);

/* Waves per workgroup of one precision's angular-first kernel: wpb in {1, 2, 4} puts that many batch
 * tiles of the same output row in one workgroup (one CU), where they share the scalar-cache copies
 * of R_t and of the routing table; 0 selects the default (1).  The autotuner also picks it (per
 * precision, among 1, 2 and 4); cfk_plan_set_tile leaves it unchanged.  No reference counterpart
 * (the MEX kernels are single-threaded). */
int cfk_plan_set_waves(cfk_plan plan, int32_t precision, int32_t wpb);
/* The waves per workgroup the plan uses for a precision (1, 2 or 4). */
int cfk_plan_get_waves(cfk_plan plan, int32_t precision, int32_t *wpb_out);

/* Output mode of cfk_plan_execute / cfk_plan_execute_static for this plan: accumulate = 1 (the
 * default, as the MEX kernels: Q += Q(f)) or 0 (Q = Q(f), for a caller that wants the operator's value
 * itself, e.g. a time stepper; Q's previous contents are neither read nor needed).  Every variant,
 * precision and layout.  No reference counterpart. */
int cfk_plan_set_accumulate(cfk_plan plan, int32_t accumulate);

/* Row split for the symmetrized precisions (CFK_PREC_FP64_SYM, CFK_PREC_MIXED): with split = 1 each
 * batch tile runs as a workgroup of two waves that take the first and second half (by the cost model)
 * of every output row's slices; the second hands its partial sums to the first through LDS, which adds
 * them in that fixed order, so results stay deterministic.  Halves the work item for a better-balanced
 * end of the launch; a waves-per-workgroup setting of 2 or more then puts two batch tiles (four waves)
 * in a workgroup.  Ignored for CFK_PREC_FP64.  The autotuner also tries it.  No reference counterpart. */
int cfk_plan_set_split(cfk_plan plan, int32_t precision, int32_t split);
/* (split = -1: the default of the selected kernel -- on for the matrix-core MIXED kernel up to K = 12,
 * else off; cfk_plan_get_split reports that value.  A call of at most 1024 distributions also splits
 * rows for both symmetrized precisions (either mixed kernel) unless the plan fixes the split, and up to
 * 256 distributions the symmetrized FP64 kernel takes the whole k3 range in one chunk (K <= 10) unless
 * the plan fixes its tile: with few waves in flight a call lasts as long as its slowest wave.) */
int cfk_plan_get_split(cfk_plan plan, int32_t precision, int32_t *split_out);

/* Routing-table statistics of a plan, for work models: stats[0] N_G, [1] S (distinct (t,q1) slices),
 * and for the symmetrized tables (0 when unavailable) [2] nonzeros kept (q2 <= q3), [3] slices,
 * [4] q2 groups (distinct (t,q1,q2)), [5] (k2,k3)-symmetric slices (l2 == l3), [6] slices whose
 * partner block is l = 0 (FP64 radial stage on the mixed path); [7] N_L.  No reference counterpart. */
int cfk_plan_get_stats(cfk_plan plan, int64_t stats[8]);

/* Kernel timing: when enabled, every execute brackets its main contraction kernel (not the
 * layout/precision conversions around it) with a pair of HIP events on the execution stream.
 * cfk_plan_read_timing waits for the recorded events, returns the summed kernel time in ms and
 * the number of launches, and clears the record. */
int cfk_plan_set_timing(cfk_plan plan, int32_t enable);
int cfk_plan_read_timing(cfk_plan plan, double *total_ms, int64_t *launches);

/* ---------------------------------------------------------------------------------------
 * Batched uneven sizes: distributions of different (L, K) -- e.g. the cells of a spatial solver --
 * grouped by geometry, one plan per group, executed together in mixed precision.
 * ------------------------------------------------------------------------------------- */
typedef struct cfk_multi_s *cfk_multi;

/* Groups plans (K in {6,8,10,12,14}, symmetrized tables available, same device, at most 8). The
 * plans must outlive the multi object.  Each plan's autotuned mixed-precision k3 chunk (K or
 * ceil(K/2)) is used by the kernel of its K. */
int cfk_multi_create(const cfk_plan *plans, int32_t n_groups, cfk_multi *out);
int cfk_multi_destroy(cfk_multi multi);

/* Q_dev[g] += Q(f_dev[g]) for B[g] distributions of group g (batch-fastest layout [K][N_Q][B],
 * FP64 in and out, mixed precision inside, invariants projected); Q_dev[g] = Q(f_dev[g]) without
 * reading Q_dev[g] for a group whose plan is in overwrite mode (cfk_plan_set_accumulate(plan, 0)).  The work is split into
 * (group, row, batch tile) items, ordered heaviest first within each radial size K; one kernel per
 * K runs its items (one workgroup per item, dispatched in that order), the kernels of different K
 * on parallel graph branches.  The conversions and the kernels are captured into a hipGraph on the
 * first call; later calls with the same pointers, sizes and output modes replay it on `stream`.  The
 * graph is captured once per multi plan: after a call with another set, every call enqueues the same
 * kernels directly (the captured graph and the buffers it reads are kept until cfk_multi_destroy).
 * Each B[g] obeys cfk_plan_execute's per-call limit. */
int cfk_multi_execute(cfk_multi multi, double *const *Q_dev, const double *const *f_dev, const int64_t *B,
                      void *stream);

/* Replay the captured hipGraph (1, the default) or enqueue the same launch sequence directly on
 * `stream` at every execute (0). */
int cfk_multi_set_graph(cfk_multi multi, int32_t use_graph);

/* ---------------------------------------------------------------------------------------
 * Multilayer perceptron of the learned surrogate (no counterpart in the reference): n_in -> width
 * (GELU) x depth -> n_out, BF16 weights and activations in the products (MFMA), FP32 accumulation,
 * FP32 master weights and Adam moments.  Device arrays are feature-major, [features][batch].
 * Parameters are one flat FP32 array, layer by layer: W_l [out][in] row-major, then b_l [out].
 * ------------------------------------------------------------------------------------- */
typedef struct cfk_mlp_s *cfk_mlp;

/* Weights seeded He-uniform (hidden) / Glorot-uniform (output), zero biases. */
int cfk_mlp_create(int32_t n_in, int32_t width, int32_t depth, int32_t n_out, uint64_t seed, cfk_mlp *out);
int cfk_mlp_destroy(cfk_mlp mlp);
int cfk_mlp_num_params(cfk_mlp mlp, int64_t *n);
/* Host copies of the FP32 parameters / of the gradients of the last cfk_mlp_loss_grad call.
 * cfk_mlp_set_params also resets the Adam state. */
int cfk_mlp_get_params(cfk_mlp mlp, float *host);
int cfk_mlp_set_params(cfk_mlp mlp, const float *host);
int cfk_mlp_get_grads(cfk_mlp mlp, float *host);
/* Y_dev[n_out][B] = network(X_dev[n_in][B]) (FP32 in and out, BF16 inside). */
int cfk_mlp_forward(cfk_mlp mlp, float *Y_dev, const float *X_dev, int64_t B, void *stream);
/* Gradients of mean((network(X) - T)^2) over all n_out*B entries; *loss (if not null) receives
 * the loss (the call then waits for the stream). */
int cfk_mlp_loss_grad(cfk_mlp mlp, const float *X_dev, const float *T_dev, int64_t B, double *loss, void *stream);
/* One bias-corrected Adam step with the gradients of the last cfk_mlp_loss_grad call. */
int cfk_mlp_adam_step(cfk_mlp mlp, float lr, float beta1, float beta2, float eps, void *stream);

/* ---------------------------------------------------------------------------------------
 * Basis projection of Maxwellian mixtures (tutorial_collisions_spectral_properties.m:100-158):
 * c_i = sum over the grid of W(v) Psi_i(v) f(v) e^{|v|^2}, with n_rad generalized Gauss-Laguerre
 * (alpha = 1/2) nodes in |v|^2 (weights halved), n_pol Gauss-Legendre nodes in cos(theta) and n_azi
 * uniform phi.  params_dev [B][n_comp][7] = (n, ux, uy, uz, Tx, Ty, Tz) per component (n = 0 skips
 * it); component density n pi^{-3/2} (Tx Ty Tz)^{-1/2} exp(-sum_d (v_d - u_d)^2 / T_d).  Output
 * c_dev [K][N_Q][B] (batch fastest), FP64.  n_comp <= 3, L <= 12, K <= 16; the grid's tables must fit
 * one workgroup's LDS (160 KB: the 60 x 30 x 30 grid up to L = 12).  Waits for the stream.
 * ------------------------------------------------------------------------------------- */
int cfk_project_maxwellians(double *c_dev, const double *params_dev, int64_t B, int32_t n_comp, int32_t L, int32_t K,
                            int32_t n_rad, int32_t n_pol, int32_t n_azi, void *stream);

/* ---------------------------------------------------------------------------------------
 * Learned surrogate of the collision operator (no counterpart in the reference).  With
 * f = n (e1 + d), n = f[k=0,q=0]:  Q~(f) = n^2 (J d + N(d)), J the operator linearized about the
 * unit Maxwellian e1 (exact, FP64, one K x K block per l), N the MLP trained on the remainder
 * Q(f)/n^2 - J d; the five invariant coefficients of Q~ are then set to exactly zero.  All device
 * arrays are batch-fastest [K][N_Q][B] FP64.  Requires n > 0.
 * ------------------------------------------------------------------------------------- */
typedef struct cfk_surrogate_s *cfk_surrogate;

/* Builds J from the plan's FP64 operator (2 K N_Q evaluations) and an untrained network
 * K N_Q -> width x depth -> K N_Q.  The plan must outlive the surrogate. */
int cfk_surrogate_create(cfk_plan plan, int32_t width, int32_t depth, uint64_t seed, cfk_surrogate *out);
/* As cfk_surrogate_create; flags CFK_SURROGATE_ANCHORED selects the decomposition about each cell's own
 * Maxwellian c_M (same n, u, T; projected on the grid of cfk_project_maxwellians with 60 x 30 x 30
 * nodes): with g = (f - c_M) / n and d_M = c_M / n - e1,
 *   Q~(f) = Q(c_M) + n^2 (J g + |g| N(d_M, g / |g|, log |g|)),
 * Q(c_M) exact (FP64).  The network then has 2 K N_Q + 1 inputs.  Its error shrinks as a cell approaches
 * its Maxwellian, where the unit-Maxwellian form keeps a fixed fraction of the remainder Q(d).  Each
 * apply computes the anchors (c_M, Q(c_M)) of its cells; a relaxation computes them once (mass,
 * momentum and energy are invariants). */
#define CFK_SURROGATE_ANCHORED 1
int cfk_surrogate_create_ex(cfk_plan plan, int32_t width, int32_t depth, uint64_t seed, int32_t flags,
                            cfk_surrogate *out);
int cfk_surrogate_destroy(cfk_surrogate s);
/* The surrogate's network (owned by the surrogate) for parameter I/O. */
int cfk_surrogate_mlp(cfk_surrogate s, cfk_mlp *out);
/* Host copy of the linear blocks J_l[k1][k], (L+1) K K values. */
int cfk_surrogate_linear_blocks(cfk_surrogate s, double *JL_host);
/* Host copies of the per-feature input mean and inverse std (one per network input: K N_Q, or
 * 2 K N_Q + 1 anchored) and target mean and std (K N_Q each). */
int cfk_surrogate_get_normalization(cfk_surrogate s, float *mu_x, float *inv_s_x, float *mu_y, float *s_y);
/* Trains on n pairs (f_dev, Q_dev = Q(f)): per-feature standardization from the set, `epochs`
 * passes of shuffled minibatches of `batch` pairs (seeded), Adam with a cosine learning-rate decay
 * from lr to lr_final.  flags: CFK_FIT_GLOBAL_INPUT_SCALE / CFK_FIT_GLOBAL_TARGET_SCALE standardize
 * inputs / targets with one scale for all features (the RMS of the per-feature deviations) instead
 * of one per feature.  loss_hist (host, `epochs` values, may be null) receives the mean training
 * loss of each epoch. */
#define CFK_FIT_GLOBAL_INPUT_SCALE 1
#define CFK_FIT_GLOBAL_TARGET_SCALE 2
int cfk_surrogate_fit(cfk_surrogate s, const double *f_dev, const double *Q_dev, int64_t n, int32_t epochs,
                      int32_t batch, float lr, float lr_final, uint64_t seed, int32_t flags, double *loss_hist,
                      void *stream);
/* Each cell's anchor: cM_dev = the basis coefficients of the isotropic Maxwellian with the cell's n, u, T
 * (cfk_moments of f_dev), QM_dev = Q(cM) from the plan's symmetrized FP64 operator; [K N_Q][B] each,
 * batch-fastest.  The coefficients are those of cfk_project_maxwellians (60 x 30 x 30 grid) with the
 * azimuthal integral exact: M is axially symmetric about u, so by the Funk-Hecke theorem its degree-l
 * part is lambda_l(|v|) Y_lm(u / |u|) with lambda_l a polar integral of the z-drifted Maxwellian.  Any
 * surrogate (anchored or not) can compute them; synchronizes the stream. */
int cfk_surrogate_anchors(cfk_surrogate s, double *cM_dev, double *QM_dev, const double *f_dev, int64_t B, void *stream);
/* Q_dev = Q~(f_dev) (overwrites).  use_network = 0 evaluates n^2 J d only (anchored: Q(c_M) + n^2 J g). */
int cfk_surrogate_apply(cfk_surrogate s, double *Q_dev, const double *f_dev, int64_t B, int32_t use_network,
                        void *stream);

#define CFK_RELAX_EXACT 0      /* exact FP64 operator every step */
#define CFK_RELAX_SURROGATE 1  /* surrogate, sampled residual check, exact fallback */
#define CFK_RELAX_LINEAR 2     /* the linear part only (no network, no check) */

typedef struct cfk_relax_stats {
    int64_t steps_network;  /* steps advanced with the surrogate (or the linear part) */
    int64_t steps_exact;    /* steps advanced with the exact operator (mode EXACT, or fallbacks) */
    int64_t exact_evals;    /* exact distribution evaluations, residual samples included */
    double max_residual;    /* largest sampled residual (mode SURROGATE) */
    double last_residual;
} cfk_relax_stats;

/* n_steps forward-Euler steps c <- c + dt Q(c) of B cells (c_dev in place).  Mode SURROGATE:
 * a step evaluates Q~ for all cells, the exact operator for every sample_stride-th cell (offset
 * rotating with the step) and the sampled residual ||Q~ - Q|| / ||Q|| over those cells; a step
 * whose residual exceeds tol is redone with the exact operator for all cells, and the surrogate
 * then rests for `backoff` steps (exact steps, unchecked), backoff doubling with each consecutive
 * failed check up to 64 and returning to 1 after a passed one.  The conservation projection is
 * applied to every Q used, so mass, momentum and energy coefficients of c do not change.
 * residual_hist (host, n_steps values, may be null) receives each step's sampled residual (-1 when
 * not checked).  Waits for the stream. */
int cfk_surrogate_relax(cfk_surrogate s, double *c_dev, int64_t B, double dt, int32_t n_steps, int32_t mode, double tol,
                        int32_t sample_stride, double *residual_hist, cfk_relax_stats *stats, void *stream);

/* ---------------------------------------------------------------------------------------
 * Basis evaluation and moment diagnostics.
 * ------------------------------------------------------------------------------------- */

/* Psi = Basis.evaluate(v) (src/SpectralBasis.m:25-58): the basis functions
 * phi_{klm}(v) = N_kl L_k^{(l+1/2)}(|v|^2) |v|^l Y_lm(v/|v|) at n_pts velocities, K radial modes (K_max+1),
 * degrees 0..L (L <= 16).  v_dev is MATLAB's n_pts x 3 array (column-major: x block, y block, z block);
 * Psi_dev is n_pts x (K (L+1)^2) column-major, column k*N_Q + q.  Device pointers; waits for the
 * stream. */
int cfk_basis_evaluate(double *Psi_dev, const double *v_dev, int64_t n_pts, int32_t K, int32_t L, void *stream);

/* Moments of f(v) = e^{-|v|^2} sum c phi(v) for B coefficient vectors (N_Q x K each, layout
 * CFK_LAYOUT_*; N_Q >= 9, K >= 2): out_dev [B][CFK_N_MOMENTS] = (n, u_x, u_y, u_z, T, P_xx, P_yy, P_zz,
 * P_xy, P_xz, P_yz) with n = int f, n u = int v f, P_ij = int (v_i - u_i)(v_j - u_j) f and
 * T = 2 (P_xx + P_yy + P_zz) / (3 n) (a Maxwellian n (pi T)^{-3/2} exp(-|v-u|^2/T) has temperature T).
 * Mass, momentum and energy are the invariant coefficients the conservation projection keeps
 * (tutorial_collisions_spectral_properties.m:138-144); the traceless stress is the l = 2 part that
 * relaxes in tutorial_stress_relaxation.m.  Asynchronous. */
#define CFK_N_MOMENTS 11
int cfk_moments(double *out_dev, const double *c_dev, int64_t B, int32_t N_Q, int32_t K, int32_t layout, void *stream);
/* The ten linear functionals behind cfk_moments (int f, int v_d f, int v_i v_j f for xx, yy, zz, xy, xz,
 * yz): functional j is sum_t w[4j+t] c[idx[4j+t]] over the terms with idx >= 0 (index k*N_Q + q).
 * Host only. */
int cfk_moment_weights(int32_t N_Q, int32_t K, int32_t idx[40], double w[40]);

/* dense variant on device: Q_dev[b] += C f_b f_b for B distributions of length N <= 4096 (layout natural),
 * C streamed from HBM once per pass of up to 4 distributions (N <= 1024) or one (larger N).  Asynchronous
 * on `stream` (0: the default stream); the per-call workspace is stream-ordered. */
int cfk_dense_execute(double *Q_dev, const double *f_dev, const double *C_dev, int64_t N, int64_t B, void *stream);

/* Dense tensor on device from the COO and R, replacing assemble_full_tensor
 * (src/GeneralCollisionTensor.m:283-311; benchmark_collision_contraction_scaling.m:54-79 builds it inline):
 * C_dev[(k1 N_Q + q1) + (k2 N_Q + q2) N + (k3 N_Q + q3) N^2] = R[k1,k2,k3,t] g for every nonzero, 0
 * elsewhere, N = N_Q*K <= 4096 (C_dev holds N^3 doubles).  COO and R are host arrays as for the MEX
 * entry points (1-based double indices).  Returns after the tensor is written. */
int cfk_dense_assemble(double *C_dev, const double *labels, const double *vals, const double *ic_map, int64_t N_G,
                       const double *R, int64_t R_numel, int32_t N_Q, int32_t K, void *stream);

/* Conservation projection: set the five collision invariants (mass, three momenta, energy) of each
 * distribution to exactly zero: q=0,1,2,3 at k=0 and q=0 at k=1 (0-based), cf.
 * tutorial_collisions_spectral_properties.m:138-144. */
int cfk_conservation_project(double *Q_dev, int64_t B, int32_t N_Q, int32_t K, int32_t layout, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CFK_H */
