"""Benchmark: Q(f) evaluations per second, hard spheres, L=8, K=10, batch 4096, on one MI355X.

A step is one pass of the hot path over the whole batch resident in HBM: Q = Q(f) for 4096
distributions through the mixed-precision path (symmetrized FP32 contraction, FP64 accumulation
and conservation projection); --accumulate times Q = 0; Q += Q(f), the MEX semantics, instead.  Its error is measured against the plain FP64 GPU operator on the
same batch (must be <= 1e-6 normwise relative per distribution), and the FP64 path is timed too.

Prints ONE JSON line (see the "metric" key).  --gpus N > 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "collision-factorization_amd"))

METRIC = "Q(f) evals/sec at HS Lmax=8 batch 4096 vs FP64 GPU path @<=1e-6 rel err"
# MI355X published peaks: FP64 vector/matrix 78.6 TF/s; FP32 matrix and vector 157.3 TF/s.  Measured on the
# MI355X with register-only loops (wave64 FMAs per SIMD): FP32 FMAs whose operands are all VGPRs issue
# about every 2.7-3.6 cycles at 2-8 waves per SIMD (two waves' FMAs issue together; the ideal is 2 cycles,
# the published 157.3 TF/s), FP32 FMAs with an SGPR operand every 4.3-6.2 cycles (never together: at
# most 78.6 TF/s), FP64 FMAs every 4.3-5 cycles (78.6 TF/s).
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "fp32_vector_sgpr_operand": 78.6e12}
# HBM bytes per launch of each mixed kernel at this workload (4096 distributions, batch-fastest, Q = Q(f)),
# from two rocprofv3 PMC passes (FETCH_SIZE, WRITE_SIZE) scaled by a plain-copy calibration; the commands
# and raw figures are in the named profile.  A kernel without an entry here reports traffic as null.
PROFILED_TRAFFIC = {
    "valu": (53908 * 1024, "profiles/af_kernel_mixed_hs_l8_k10_b4096.txt"),
    "mfma": (55416 * 1024, "profiles/af_mx_kernel_mixed_hs_l8_k10_b4096.txt"),
}


def work_model(lab, ic, K):
    """Algorithmic FLOP per evaluation of the angular-first structure: 2 (N_G K^2 + S K^3)."""
    n_g = lab.shape[0]
    S = len(set(zip(np.asarray(ic).astype(int).tolist(), np.asarray(lab)[:, 0].astype(int).tolist())))
    return 2.0 * (n_g * K * K + S * K ** 3), n_g, S


def executed_flops(st, K, matrix_cores):
    """FLOP per evaluation the mixed kernel executes on the symmetrized tables (plan statistics st), by the
    issue rate that bounds them:
      dual  -- FP32 FMAs with VGPR operands only: the outer products f2 (x) h of stage 1 (one per q2 group);
      single -- FMAs with a wave-uniform SGPR operand: the h sums g f3 (Gaunt value), the radial stage on
                the VALU (R from the scalar path: full blocks K^3, symmetric blocks K^2 (K+1)/2) and the
                FP64 radial stage of the precise slices;
      mfma  -- the radial stage on the matrix cores (16x16x4 steps of 4 values of (k2,k3), rows k1 padded
               to 16; executed, not useful, FLOPs)."""
    plain = st["sym_S"] - st["sym_symmetric"] - st["sym_precise"]
    dual = 2.0 * st["sym_groups"] * K * K
    single = 2.0 * (st["sym_N_G"] * K + st["sym_precise"] * K ** 3)
    if not matrix_cores:
        single += 2.0 * (plain * K ** 3 + st["sym_symmetric"] * K * K * (K + 1) / 2)
        return dual, single, 0.0
    kc = K if K <= 10 else (K + 1) // 2
    steps_full = (K * kc + 3) // 4 * ((K + kc - 1) // kc)
    steps_sym = (K * (K + 1) // 2 + 3) // 4 if kc == K else steps_full
    mfma = (plain * steps_full + st["sym_symmetric"] * steps_sym) * 4 * 16 * 2.0
    return dual, single, mfma


def cpu_baseline(lab, val, ic, R, nq, K, f, budget_s):
    """The oracle (plain-C restatement of angular_first_collision_kernel_mex) on one core."""
    sys.path.insert(0, ROOT)
    import oracle
    n = 4
    t0 = time.perf_counter()
    oracle.angular_first_batch(np.zeros(n * nq * K), f[:n], lab, val, ic, R, nq, K, nthreads=1)
    per = (time.perf_counter() - t0) / n
    m = int(max(8, min(f.shape[0], budget_s / max(per, 1e-9))))
    t0 = time.perf_counter()
    oracle.angular_first_batch(np.zeros(m * nq * K), f[:m], lab, val, ic, R, nq, K, nthreads=1)
    dt = time.perf_counter() - t0
    out = {"value": m / dt, "unit": "evals/s", "cores": 1, "kind": "port",
           "sample": "%d of the %d distributions, oracle angular-first (restated reference MEX), 1 thread, %.1f s"
                     % (m, f.shape[0], dt)}
    # the reference contraction is single-threaded; an all-core figure needs parallelism over the
    # batch added by the harness (OpenMP in the oracle's batch driver), reported separately
    cores = min(os.cpu_count() or 1, 16)
    m2 = int(max(8, min(f.shape[0], cores * budget_s / max(per, 1e-9) / 2)))
    t0 = time.perf_counter()
    oracle.angular_first_batch(np.zeros(m2 * nq * K), f[:m2], lab, val, ic, R, nq, K, nthreads=cores)
    dt2 = time.perf_counter() - t0
    out["all_cores"] = {"value": m2 / dt2, "cores": cores,
                        "note": "harness-level parallelism over the batch, %d distributions, %.1f s" % (m2, dt2)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed-kernel", default="auto", choices=["auto", "valu", "mfma"],
                    help="hs-l8-k10: the mixed-precision kernel -- auto: the autotuner's pick; valu / mfma: the VALU or "
                         "the matrix-core kernel with the row split and two batch tiles per workgroup (no autotuning "
                         "of the mixed path), e.g. to profile one kernel")
    ap.add_argument("--no-compare", action="store_true",
                    help="rtensor: skip the per-channel kernel run kept beside the matrix-core one")
    ap.add_argument("--accumulate", action="store_true",
                    help="default workload: time Q = 0; Q += Q(f) (the MEX semantics) instead of Q = Q(f)")
    ap.add_argument("--workload", default="hs-l8-k10", choices=["hs-l8-k10", "uneven", "rtensor", "surrogate", "copy", "dense"],
                    help="hs-l8-k10: the metric; uneven: 4096 distributions with L in {4,6,8,12}, K=L+2 "
                         "(batched item kernels in a hipGraph); rtensor: GPU tensor generation, L=12, K=12, 48^7 "
                         "grids; surrogate: training on 2e5 mixture pairs and 16384-cell relaxation; copy: "
                         "a plain 512 MiB device copy, three times (calibration of HBM traffic counters); dense: the "
                         "dense-tensor contraction (dense_tensor_kernel_mex) at L=8, K=10, N=810, C = 4.25 GB on the "
                         "device, one distribution per call and four per call")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus %d" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    import cfk
    from cfk.inputs import perturbed_maxwellians

    if args.workload == "uneven":
        return bench_uneven(args, torch, cfk, perturbed_maxwellians)
    if args.workload == "rtensor":
        return bench_rtensor(args, cfk)
    if args.workload == "surrogate":
        return bench_surrogate(args, torch, cfk)
    if args.workload == "copy":
        return bench_copy(torch)
    if args.workload == "dense":
        return bench_dense(args, torch, cfk, perturbed_maxwellians)

    L, K_max, gamma, B = 8, 9, 1.0, args.batch
    K = K_max + 1
    nq = (L + 1) ** 2
    grid = (48,) * 7   # Gauss-Legendre 32 points per dimension, padded 1.5x, on every axis
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    t1 = time.perf_counter()
    R = cfk.build_rtensor(K_max, L, gamma, grid, geometry=(lab, val, ic))
    t2 = time.perf_counter()
    plan = cfk.CollisionPlan(ls, vs, ics, R, nq, K)
    tiles = plan.autotune(B)
    if args.mixed_kernel != "auto":
        plan.set_tile("mixed", 0, 1, 1 if args.mixed_kernel == "mfma" else 0)
        plan.set_split("mixed", 1)
        plan.set_waves("mixed", 2)
        tiles["mixed"] = (K, 1, 1 if args.mixed_kernel == "mfma" else 0)
    F, n_g, S = work_model(ls, ics, K)

    f_host = perturbed_maxwellians(L, K, B, seed=0)
    f = torch.from_numpy(f_host).to(dev)
    # batch-fastest layout [N_Q*K][B] is what a spatial solver holds; convert once, outside timing
    fT = f.reshape(B, K * nq).t().contiguous()
    Q = torch.zeros(K * nq, B, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    # one step = Q(f) for the batch: overwrite mode (Q = Q(f), the operator value a time stepper
    # wants), or with --accumulate the MEX semantics on a zeroed Q (Q = 0; Q += Q(f))
    plan.set_accumulate(args.accumulate)

    def step(precision):
        if args.accumulate:
            Q.zero_()
        plan.execute(Q, fT, B, precision=precision, layout="batch_fastest", stream=stream)

    # FP64 reference result on the same batch
    step("fp64")
    torch.cuda.synchronize()
    Q64 = Q.clone()

    def timed(precision, steps, warmup):
        """(seconds per step, kernel seconds per launch): the steps timed with nothing else on the
        stream, then the same number of steps again with the library's HIP-event kernel timer on (its
        event records add launch overhead, so they stay out of the step time)."""
        for _ in range(warmup):
            step(precision)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(steps):
            step(precision)
        torch.cuda.synchronize()
        te = time.perf_counter()
        plan.read_timing()
        plan.set_timing(True)
        for _ in range(steps):
            step(precision)
        torch.cuda.synchronize()
        plan.set_timing(False)
        kms, nl = plan.read_timing()
        return (te - ts) / steps, kms / max(nl, 1) / 1e3

    t64, k64 = timed("fp64", max(3, args.steps // 4), 1)
    tsym, ksym = timed("fp64_sym", max(3, args.steps // 4), 1)
    tmix, kmix = timed("mixed", args.steps, args.warmup)
    Qm = Q.clone()
    err = (Qm - Q64).abs().amax(dim=0) / Q64.abs().amax(dim=0)
    max_rel = float(err.max())
    inv_rows = [0, 1, 2, 3, nq]   # (q=0..3, k=0) and (q=0, k=1) in the [K*N_Q] row index k*N_Q + q
    inv_abs = float(Qm[inv_rows].abs().max())

    value = B / tmix
    achieved = F * B / kmix
    matrix_cores = tiles["mixed"][2] == 1
    kind = "mfma" if matrix_cores else "valu"
    st = plan.stats()
    e_dual, e_single, e_mfma = executed_flops(st, K, matrix_cores)
    # bound on the executed work: dual-issue FMAs at 157.3 TF/s, SGPR-operand FMAs at 78.6 TF/s and MFMA
    # FLOPs at 157.3 TF/s, added (on gfx950 FP32 MFMAs and FP32 VALU work of other waves on the same SIMD
    # do not overlap: measured, waves of both kinds together take as long as the two alone added)
    t_exec = (e_dual / PEAK["fp32"] + e_single / PEAK["fp32_vector_sgpr_operand"] + e_mfma / PEAK["fp32"]) * B
    traffic, traffic_src = PROFILED_TRAFFIC.get(kind, (None, None)) if B == 4096 else (None, None)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tmix * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "hard spheres (gamma=1), L=8, K=10 radial modes, batch 4096 perturbed Maxwellians, "
                               "mixed-precision angular-first contraction", "L": L, "K": K, "batch": B,
                   "gamma": gamma, "N_G": n_g, "slices": S, "sumfac_grid": list(grid),
                   "step": "Q = 0; Q += Q(f)" if args.accumulate else "Q = Q(f)"},
        "max_rel_err_vs_fp64": max_rel,
        "invariants_max_abs": inv_abs,
        "fp64_evals_per_s": B / t64,
        "fp64_sym_evals_per_s": B / tsym,
        "speedup_vs_fp64": t64 / tmix,
        "speedup_vs_fp64_sym": tsym / tmix,
        "kernel_ms": {"mixed": kmix * 1e3, "fp64": k64 * 1e3, "fp64_sym": ksym * 1e3},
        "flop_per_eval": F,
        "tiles_kc": tiles,
        "waves_per_workgroup": {prec: plan.waves(prec) for prec in ("fp64", "mixed", "fp64_sym")},
        "row_split": {prec: plan.split(prec) for prec in ("fp64", "mixed", "fp64_sym")},
        "setup_s": {"geometry": t1 - t0, "rtensor_sumfac_gpu": t2 - t1},
        "roofline": {"bound": "valu" if kind == "valu" else "mfma", "achieved": achieved / 1e12,
                     "peak": PEAK["fp32"] / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK["fp32"],
                     "achieved_basis": "work-model FLOPs 2(N_G K^2 + S K^3) per eval over the kernel's HIP-event "
                                       "time; the kernel executes fewer (executed)",
                     "kernel": "matrix-core radial stage" if matrix_cores else "VALU",
                     "traffic": traffic, "traffic_source": traffic_src,
                     "executed": {"valu_dual_issue_flop_per_eval": e_dual, "valu_sgpr_operand_flop_per_eval": e_single,
                                  "mfma_flop_per_eval": e_mfma,
                                  "tflops": (e_dual + e_single + e_mfma) * B / kmix / 1e12,
                                  "frac_of_fp32_peak": (e_dual + e_single + e_mfma) * B / kmix / PEAK["fp32"],
                                  "bound_ms": t_exec * 1e3, "frac_of_bound": t_exec / kmix,
                                  "bound_basis": "dual-issue VGPR FMAs at 157.3 TF/s + SGPR-operand FMAs at 78.6 TF/s "
                                                 "+ MFMA at 157.3 TF/s, added"},
                     "fp64_path_frac": F * B / k64 / PEAK["fp64"]},
        "cpu_baseline": None,
    }
    if max_rel > 1e-6:
        out["error"] = "mixed path exceeds 1e-6 relative error"
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ls, vs, ics, R, nq, K, f_host, args.cpu_seconds)
    plan.close()
    print(json.dumps(out))


def bench_uneven(args, torch, cfk, perturbed_maxwellians):
    """4096 hard-sphere distributions with L drawn uniformly from {4,6,8,12}, K = L+2, one multi plan."""
    B = args.batch
    rng = np.random.Generator(np.random.PCG64(7))
    Ls = [4, 6, 8, 12]
    counts = np.bincount(rng.integers(0, 4, size=B), minlength=4)
    plans, fs, Qs, refs, F = [], [], [], [], 0.0
    setup0 = time.perf_counter()
    for i, L in enumerate(Ls):
        K = L + 2
        nq = (L + 1) ** 2
        lab, val, ic = cfk.real_gaunt(L)
        ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
        R = cfk.build_rtensor(K - 1, L, 1.0, cfk.sumfac_grid_sizes(K - 1, L, 0, 0), geometry=(lab, val, ic))
        plan = cfk.CollisionPlan(ls, vs, ics, R, nq, K)
        n = int(counts[i])
        f = torch.from_numpy(perturbed_maxwellians(L, K, n, seed=10 + i)).cuda().reshape(n, -1).t().contiguous()
        ref = torch.zeros_like(f)
        plan.execute(ref, f, n, precision="fp64", layout="batch_fastest")
        plans.append(plan)
        fs.append(f)
        Qs.append(torch.zeros_like(f))
        refs.append(ref)
        F += n * work_model(ls, ics, K)[0]
    setup = time.perf_counter() - setup0
    # one step writes Q = Q(f) for every group (each plan in overwrite mode: Q is not read)
    for plan in plans:
        plan.set_accumulate(False)
    mp = cfk.MultiPlan(plans)
    stream = torch.cuda.current_stream()

    def step():
        mp.execute(Qs, fs, list(counts), stream=stream)

    def separate():
        for p, Q, f, n in zip(plans, Qs, fs, counts):
            p.execute(Q, f, int(n), precision="mixed", layout="batch_fastest", stream=stream)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - ts) / args.steps

    t_sep = timed(separate)
    mp.set_graph(False)
    t_eager = timed(step)
    mp.set_graph(True)
    t = timed(step)
    errs = {}
    for L, Q, ref in zip(Ls, Qs, refs):
        errs["L%d" % L] = float(((Q - ref).abs().amax(dim=0) / ref.abs().amax(dim=0)).max())
    out = {"metric": "Q(f) evals/sec, batched uneven sizes (L in {4,6,8,12}, K=L+2), batch 4096",
           "value": B / t, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "uneven L in {4,6,8,12}, K=L+2, hard spheres, batch %d" % B,
                      "counts": [int(c) for c in counts]},
           "max_rel_err_vs_fp64_per_L": errs, "separate_launches_evals_per_s": B / t_sep,
           "no_graph_evals_per_s": B / t_eager,
           "algorithmic_tflops": F / t / 1e12, "setup_s": setup}
    print(json.dumps(out))


def sumfac_flops(L, NK, sizes, S):
    """Algorithmic FP64 flops of the matrix-core quadrature: per inner-t step and side the degree
    blocks M (2 NK N_Q per gain point) and the slot contraction (2 S NK); per unit the fold
    (two products per R element: 4 NK^3 N_L)."""
    Nx, Nu1, Nt1, Ny2, Nt2, Nchi, Neps = sizes
    NQ = (L + 1) ** 2
    n_l = {4: 42, 6: 106, 8: 215, 10: 381, 12: 616}.get(L)
    n_tau = Nx * (Nu1 * Nt1 + Ny2 * Nt2)
    n_units = Nx * (Nu1 + Ny2 * Nt2)
    f_units = n_tau * 2 * (2 * NK * NQ * Nchi * Neps + 2 * S * NK)
    f_fold = n_units * 4 * NK ** 3 * n_l
    return f_units, f_fold


def bench_rtensor(args, cfk):
    """GPU tensor generation at L=12, K=12 radial modes, 48 nodes per axis, gamma = 0 and 1, on the
    default kernel (exact-sphere: this rule is exact for the gain integrand); the matrix-core
    quadrature and (unless --no-compare) the per-channel kernel (reference summation order) timed
    beside it on gamma = 1."""
    L, K_max = 12, 11
    NK = K_max + 1
    sizes = (48,) * 7
    geo = cfk.real_gaunt(L)
    S = len({(int(t), int(q)) for t, q in zip(geo[2], geo[0][:, 0])})
    res = {}
    timing = {}
    Rs = {}
    runs = [("auto", (0.0, 1.0)), ("matrix_core", (1.0,))] + ([] if args.no_compare else [("per_channel", (1.0,))])
    try:
        for kind, gammas in runs:
            cfk.set_sumfac_kernel(kind)
            for gamma in gammas:
                t0 = time.perf_counter()
                R = cfk.build_rtensor(K_max, L, gamma, sizes, geometry=geo)
                key = "%s_gamma%d" % (kind, int(gamma))
                res[key + "_s"] = time.perf_counter() - t0
                res[key + "_max_abs_R"] = float(np.max(np.abs(R)))
                timing[key] = cfk.sumfac_last_timing()
                if gamma == 1.0:
                    Rs[kind] = R
                print("%s: %.3f s, kernels %s" % (key, res[key + "_s"], timing[key]), file=sys.stderr, flush=True)
    finally:
        cfk.set_sumfac_kernel("auto")
    scale = float(np.max(np.abs(Rs["matrix_core"])))
    diffs = {k: float(np.max(np.abs(v - Rs["matrix_core"])) / scale) for k, v in Rs.items() if k != "matrix_core"}
    cpu = None
    if not args.no_cpu:
        # the oracle's sumfac (plain-C restatement of compute_rtensor_sumfac_mex, OpenMP over the
        # quadrature units like the reference) on one of the 48 x nodes, all host cores
        sys.path.insert(0, ROOT)
        import oracle
        from oracle import geometry as G
        prep = G.sumfac_prep(K_max, L, 1.0, *[np.asarray(a) for a in geo], G.sumfac_grids(1.0, sizes))
        cores = min(os.cpu_count() or 1, 16)
        t0 = time.perf_counter()
        oracle.sumfac(prep, 20, 21, nthreads=cores)
        dt = time.perf_counter() - t0
        cpu = {"value": dt * sizes[0], "unit": "s", "cores": cores, "kind": "port",
               "sample": "x node 20 of 48 (gamma 1) in %.1f s, scaled by 48; oracle sumfac (restated reference "
                         "MEX, OpenMP over quadrature units as the reference)" % dt}
    f_units, f_fold = sumfac_flops(L, NK, sizes, S)
    tm = timing["matrix_core_gamma1"]
    ts = timing["auto_gamma1"]
    gain_points = 48 * (48 * 48 + 48 * 48) * 48 * 48
    out = {"metric": "tensor generation seconds, L=12, K=12, 48 nodes per axis", "value": res["auto_gamma1_s"],
           "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "compute_rtensor_sumfac, L=12, K=12, 48^7, gamma 1, default kernel (%s)"
                                  % ts["kernel"], "gain_points": gain_points, "slots": S},
           "kernel_ms": timing,
           "max_rel_diff_vs_matrix_core": diffs,
           "speedup_vs_matrix_core": res["matrix_core_gamma1_s"] / res["auto_gamma1_s"],
           "speedup_vs_per_channel": (res["per_channel_gamma1_s"] / res["auto_gamma1_s"]
                                      if "per_channel_gamma1_s" in res else None),
           "cpu_baseline": cpu,
           "matrix_core_roofline": {"bound": "mfma", "kernel": "sumfac_units_mf_kernel",
                                    "achieved": f_units / (tm["units_ms"] * 1e-3) / 1e12, "peak": 78.6,
                                    "unit": "TFLOP/s", "frac": f_units / (tm["units_ms"] * 1e-3) / 1e12 / 78.6,
                                    "flop_units": f_units},
           "fold": {"kernel": "sumfac_fold_tiles_kernel", "flop": f_fold,
                    "tflops": f_fold / (ts["fold_ms"] * 1e-3) / 1e12,
                    "frac": f_fold / (ts["fold_ms"] * 1e-3) / 1e12 / 78.6},
           **res}
    print(json.dumps(out))


def bench_copy(torch):
    """Three plain copies of 512 MiB on the device: the reference reading for FETCH_SIZE and
    WRITE_SIZE counters (bytes read and written are known exactly)."""
    n = 64 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 3
    print(json.dumps({"metric": "device copy bandwidth", "value": 2 * n * 8 / dt / 1e9, "unit": "GB/s",
                      "n_gpus": 1, "higher_is_better": True, "config": {"workload": "copy 512 MiB x 3"}}))


def bench_dense(args, torch, cfk, perturbed_maxwellians):
    """Dense-tensor contraction Q += C f f (dense_tensor_kernel_mex) for hard spheres, L=8, K=10: N = 810,
    C (N^3 doubles, 4.25 GB) assembled on the device from the COO and the sumfac R.  One call streams C
    once for up to 4 distributions, so the bound is HBM bandwidth: bytes = N^3 * 8 per call."""
    L, K_max, gamma = 8, 9, 1.0
    K = K_max + 1
    nq = (L + 1) ** 2
    N = nq * K
    lab, val, ic = cfk.real_gaunt(L)
    R = cfk.build_rtensor(K_max, L, gamma, (48,) * 7, geometry=(lab, val, ic))
    C = torch.empty(N ** 3, dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    cfk.dense_assemble(C, lab, val, ic, R, nq, K)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    stream = torch.cuda.current_stream()
    res = {}
    for B in (1, 4):
        f = torch.from_numpy(perturbed_maxwellians(L, K, B, seed=0).reshape(B, N)).cuda()
        Q = torch.zeros(B, N, dtype=torch.float64, device="cuda")
        for _ in range(max(1, args.warmup)):
            cfk.dense_execute(Q, f, C, N, B, stream=stream)
        steps = max(3, args.steps // 2)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(steps):
            cfk.dense_execute(Q, f, C, N, B, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        res[B] = ms
    bytes_per_call = N ** 3 * 8
    gbs = bytes_per_call / (res[1] * 1e-3) / 1e9
    print(json.dumps({
        "metric": "dense-tensor Q(f) evals/sec (dense_tensor_kernel_mex), HS L=8, K=10, N=810",
        "value": 1e3 / res[1], "unit": "evals/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
        "data": "synthetic", "config": {"workload": "dense contraction, hard spheres, L=8, K=10, N=810, C=4.25 GB",
                                         "N": N},
        "ms_per_call": {"B=1": res[1], "B=4": res[4]}, "evals_per_s_B4": 4e3 / res[4],
        "assemble_s": t_asm,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0,
                     "achieved_basis": "N^3 * 8 bytes of C per call (B=1) over the call's HIP-event time "
                                       "(stream kernel plus the partial-sum reduction)",
                     "achieved_B4": bytes_per_call / (res[4] * 1e-3) / 1e9},
    }))


def bench_surrogate(args, torch, cfk):
    """Learned surrogate, hard spheres L=8, K=10: training (2e5 mixture pairs, 40 epochs) and
    forward-Euler relaxation of 16384 cells, 200 steps of 0.005 / mu_stress, exact stepping against
    the surrogate with the 1%-sampled residual check at 5%.  Then the decomposition about each
    cell's own Maxwellian (anchored), trained on 1e5 mixtures plus states every 15 steps along 300
    exact steps of 12000 other mixtures, relaxing the same cells with the check at 5% and 10% (and
    the unit-Maxwellian surrogate at 10% beside it)."""
    L, K, B, steps, tol = 8, 10, 16384, 200, 5e-2
    nq = (L + 1) ** 2
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    R = cfk.build_rtensor(K - 1, L, 1.0, (48,) * 7, geometry=(lab, val, ic))
    plan = cfk.CollisionPlan(ls, vs, ics, R, nq, K)

    def pairs(n, seed):
        f = cfk.project_maxwellians(cfk.maxwellian_mixture_params(n, seed=seed), L, K)
        Q = torch.zeros_like(f)
        for s0 in range(0, n, 16384):
            e = min(n, s0 + 16384)
            fs = f[:, s0:e].contiguous()
            Qs = torch.zeros_like(fs)
            plan.execute(Qs, fs, e - s0, precision="fp64_sym", layout="batch_fastest")
            Q[:, s0:e] = Qs
        cfk.conservation_project(Q, n, nq, K, layout="batch_fastest")
        return f, Q

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f, Q = pairs(200000, 1)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    fv, Qv = pairs(4096, 2)
    sur = cfk.Surrogate(plan, width=1024, depth=3, seed=0)
    t0 = time.perf_counter()
    hist = sur.fit(f, Q, epochs=40, batch=1024, lr=1e-3, lr_final=1e-5, seed=3)
    t_fit = time.perf_counter() - t0
    del f, Q
    rel = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    err_net, err_lin = rel(sur.apply(fv), Qv), rel(sur.apply(fv, use_network=False), Qv)
    c0 = cfk.project_maxwellians(cfk.maxwellian_mixture_params(B, seed=5), L, K)
    dt = 0.005 / abs(sur.linear_blocks()[2][0][0])
    out_modes = {}
    finals = {}
    for mode in ("exact", "surrogate", "linear"):
        c = c0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, _ = sur.relax(c, dt, steps, mode=mode, tol=tol, sample_stride=100)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        finals[mode] = c
        out_modes[mode] = {"s": el, "ms_per_step": el / steps * 1e3, **st}
    chg = float(torch.linalg.norm(finals["exact"] - c0))
    for mode in ("surrogate", "linear"):
        out_modes[mode]["final_err_rel_change"] = float(torch.linalg.norm(finals[mode] - finals["exact"])) / chg

    # anchored decomposition: training set with states along exact relaxations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fa, Qa = pairs(100000, 1)
    c = cfk.project_maxwellians(cfk.maxwellian_mixture_params(12000, seed=7), L, K)
    states = [c.clone()]
    Qc = torch.zeros_like(c)
    for s0 in range(300):
        Qc.zero_()
        plan.execute(Qc, c, c.shape[1], precision="fp64_sym", layout="batch_fastest")
        cfk.conservation_project(Qc, c.shape[1], nq, K, layout="batch_fastest")
        c = c + dt * Qc
        if (s0 + 1) % 15 == 0:
            states.append(c.clone())
    ftr = torch.cat(states, dim=1).contiguous()
    Qtr = torch.zeros_like(ftr)
    for s0 in range(0, ftr.shape[1], 65536):
        e = min(ftr.shape[1], s0 + 65536)
        fs = ftr[:, s0:e].contiguous()
        Qs = torch.zeros_like(fs)
        plan.execute(Qs, fs, e - s0, precision="fp64_sym", layout="batch_fastest")
        Qtr[:, s0:e] = Qs
    cfk.conservation_project(Qtr, ftr.shape[1], nq, K, layout="batch_fastest")
    fa = torch.cat([fa, ftr], dim=1).contiguous()
    Qa = torch.cat([Qa, Qtr], dim=1).contiguous()
    del ftr, Qtr, states
    torch.cuda.synchronize()
    t_data_a = time.perf_counter() - t0
    sur_a = cfk.Surrogate(plan, width=1024, depth=3, seed=0, anchored=True)
    t0 = time.perf_counter()
    hist_a = sur_a.fit(fa, Qa, epochs=40, batch=1024, lr=1e-3, lr_final=1e-5, seed=3)
    t_fit_a = time.perf_counter() - t0
    n_pairs_a = fa.shape[1]
    del fa, Qa
    err_net_a = rel(sur_a.apply(fv), Qv)
    err_lin_a = rel(sur_a.apply(fv, use_network=False), Qv)
    anch = {}
    for name, s_, mode, tl in (("anchored@0.05", sur_a, "surrogate", 0.05), ("anchored@0.1", sur_a, "surrogate", 0.1),
                               ("unit@0.1", sur, "surrogate", 0.1), ("anchored_linear", sur_a, "linear", 0.1)):
        c = c0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, _ = s_.relax(c, dt, steps, mode=mode, tol=tl, sample_stride=100)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        anch[name] = {"s": el, "ms_per_step": el / steps * 1e3,
                      "final_err_rel_change": float(torch.linalg.norm(c - finals["exact"])) / chg, **st}
    # one evaluation of all cells: surrogate vs exact FP64 vs mixed
    Qb = torch.zeros_like(c0)

    def timed(fn, n=20):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n * 1e3

    t_sur = timed(lambda: sur.apply(c0, out=Qb))
    t_sur_a = timed(lambda: sur_a.apply(c0, out=Qb))   # anchors (moments, c_M, exact Q(c_M)) included
    t_f64 = timed(lambda: plan.execute(Qb, c0, B, precision="fp64_sym", layout="batch_fastest"))
    t_mix = timed(lambda: plan.execute(Qb, c0, B, precision="mixed", layout="batch_fastest"))
    gemm_flop = 2.0 * B * (810 * 1024 + 2 * 1024 * 1024 + 1024 * 810)
    out = {"metric": "surrogate relaxation, 16384 cells x 200 forward-Euler steps, HS L=8 K=10",
           "value": B * steps / out_modes["surrogate"]["s"], "unit": "cell-steps/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "bf16", "data": "synthetic",
           "config": {"workload": "surrogate 810 -> 3 x 1024 GELU -> 810, 2e5 training pairs, 40 epochs, "
                                  "tolerance %g on 1%% sampled cells" % tol},
           "exact_cell_steps_per_s": B * steps / out_modes["exact"]["s"],
           "train_data_s": t_data, "fit_s": t_fit, "loss_first_last": [float(hist[0]), float(hist[-1])],
           "heldout_rel_err": {"network": err_net, "linear_part_only": err_lin},
           "relax": out_modes,
           "eval_ms_16384": {"surrogate": t_sur, "anchored_surrogate_with_anchors": t_sur_a, "exact_fp64": t_f64,
                             "mixed": t_mix},
           "network_tflops": gemm_flop / (t_sur * 1e-3) / 1e12,
           "anchored": {"train_pairs": n_pairs_a, "train_data_s": t_data_a, "fit_s": t_fit_a,
                        "loss_first_last": [float(hist_a[0]), float(hist_a[-1])],
                        "heldout_rel_err": {"network": err_net_a, "linear_part_only": err_lin_a},
                        "relax": anch}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
