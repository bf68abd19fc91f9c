"""GPU parity of the contraction path (libcfk, HIP) against the CPU oracle and the golden vectors.

Bar: FP64 variants within 1e-12 normwise relative per distribution (max_b |Q-Q_ref|_inf /
|Q_ref|_inf); the mixed path within 1e-6 with the five invariants exactly zero.
"""
import os

import numpy as np
import pytest

import cfk
import oracle
from cfk.inputs import perturbed_maxwellians
from conftest import GOLDEN, load_precalc
from oracle import geometry as G

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL_MIXED = 1e-6


def rel_err(Q, Qref, nq, K):
    Q = np.asarray(Q).reshape(-1, nq * K)
    Qref = np.asarray(Qref).reshape(-1, nq * K)
    return float(np.max(np.max(np.abs(Q - Qref), axis=1) / np.max(np.abs(Qref), axis=1)))


def test_config0_mex_entry_points_match_golden():
    d = np.load(os.path.join(GOLDEN, "config0_maxwell_l4_k6.npz"))
    L, K = 4, 6
    nq = (L + 1) ** 2
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    f = d["f"].reshape(nq, K, order="F")
    for fn, args in ((cfk.angular_first_collision_kernel_mex, (ls, vs, ics)),
                     (cfk.radial_first_collision_kernel_mex, (ls, vs, ics)),
                     (cfk.naive_collision_kernel_mex, (lab, val, ic))):
        print("config0", fn.__name__, flush=True)
        Q = np.zeros((nq, K), order="F")
        fn(Q, f, *args, d["R"], nq, K)
        assert rel_err(Q.ravel(order="F"), d["Q_naive"], nq, K) <= TOL64
        assert rel_err(Q.ravel(order="F"), d["Q_af"], nq, K) <= TOL64
        fn(Q, f, *args, d["R"], nq, K)          # accumulates in place, as the MEX
        assert rel_err(Q.ravel(order="F"), 2 * d["Q_naive"], nq, K) <= TOL64
    N = nq * K
    C = G.assemble_dense(lab.astype(np.int64), val, ic.astype(np.int64), d["R"], nq, K)
    Qd = np.zeros(N, order="F")
    cfk.dense_tensor_kernel_mex(Qd, d["f"], C, N)
    assert rel_err(Qd, d["Q_naive"], nq, K) <= TOL64


def test_config0_sumfac_on_gpu_matches_golden():
    d = np.load(os.path.join(GOLDEN, "config0_maxwell_l4_k6.npz"))
    R = cfk.build_rtensor(5, 4, 0.0, tuple(int(v) for v in d["sizes"]), geometry=cfk.real_gaunt(4))
    assert np.max(np.abs(R - d["R"])) <= TOL64 * np.max(np.abs(d["R"]))


@pytest.mark.parametrize("L,gamma", [(2, 0.0), (4, 1.0)])
def test_sumfac_gpu_reproduces_reference_fixtures(L, gamma):
    # the reference's own R (K_max=4, pad 20) through the native prep + GPU quadrature
    R = cfk.build_rtensor(4, L, gamma, cfk.sumfac_grid_sizes(4, L, 20, 20), geometry=cfk.real_gaunt(L))
    Rf = load_precalc(L, gamma)[3]
    assert np.max(np.abs(R - Rf)) <= 1e-13 * np.max(np.abs(Rf))


def test_sumfac_mex_entry_point_matches_oracle_on_x_ranges():
    K_max, L, gamma = 3, 4, 1.0
    lab, val, ic = G.real_gaunt_geometry(L)
    grids = G.sumfac_grids(gamma, G.sumfac_grid_sizes(K_max, L, 6, 6))
    args = G.sumfac_prep(K_max, L, gamma, lab, val, ic, grids)
    prhs = [args[k] for k in oracle.SUMFAC_ORDER]
    n_x = args["x"].size
    full = cfk.compute_rtensor_sumfac_mex(*prhs)
    ref = oracle.sumfac(args, nthreads=1)
    assert full.shape == (K_max + 1,) * 3 + (int(args["N_L"]),)
    assert np.max(np.abs(full.ravel(order="F") - ref)) <= TOL64 * np.max(np.abs(ref))
    part = cfk.compute_rtensor_sumfac_mex(*prhs, x_range=(2, 5))
    refp = oracle.sumfac(args, 2, 5, nthreads=1)
    assert np.max(np.abs(part.ravel(order="F") - refp)) <= TOL64 * np.max(np.abs(refp))
    assert n_x > 5


def config1_R():
    path = os.path.join(GOLDEN, "sumfac_hs_l8_k10_g48.npz")
    if os.path.exists(path):
        return np.load(path)["R"]
    return cfk.build_rtensor(9, 8, 1.0, (48,) * 7, geometry=cfk.real_gaunt(8))


@pytest.fixture(scope="module")
def config1(torch_cuda):
    L, K = 8, 10
    nq = (L + 1) ** 2
    lab, val, ic = cfk.real_gaunt(L)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    R = config1_R()
    plan = cfk.CollisionPlan(ls, vs, ics, R, nq, K)
    return dict(L=L, K=K, nq=nq, lab=ls, val=vs, ic=ics, R=R, plan=plan)


@pytest.mark.parametrize("B", [1, 256])
def test_config1_batched_parity(config1, torch_cuda, B):
    torch = torch_cuda
    c = config1
    nq, K = c["nq"], c["K"]
    f = perturbed_maxwellians(c["L"], K, B, seed=11 + B)
    Qref = oracle.angular_first_batch(np.zeros(B * nq * K), f, c["lab"], c["val"], c["ic"], c["R"], nq, K,
                                      nthreads=8)
    fd = torch.from_numpy(f).cuda()
    for variant in ("angular_first", "radial_first", "naive"):
        print("config1 B=%d %s" % (B, variant), flush=True)
        Qd = torch.zeros(B * nq * K, dtype=torch.float64, device="cuda")
        c["plan"].execute(Qd, fd, B, variant=variant)
        torch.cuda.synchronize()
        assert rel_err(Qd.cpu().numpy(), Qref, nq, K) <= TOL64, variant
    # symmetrized FP64 (half the routing table) is still the FP64 operator
    Qd = torch.zeros(B * nq * K, dtype=torch.float64, device="cuda")
    c["plan"].execute(Qd, fd, B, precision="fp64_sym")
    torch.cuda.synchronize()
    assert rel_err(Qd.cpu().numpy(), Qref, nq, K) <= TOL64
    # mixed precision: error bound and exact conservation
    Qd = torch.zeros(B * nq * K, dtype=torch.float64, device="cuda")
    c["plan"].execute(Qd, fd, B, precision="mixed")
    torch.cuda.synchronize()
    Qm = Qd.cpu().numpy().reshape(B, K, nq)
    assert rel_err(Qm, Qref, nq, K) <= TOL_MIXED
    assert np.all(Qm[:, 0, 0:4] == 0.0) and np.all(Qm[:, 1, 0] == 0.0)


def test_config1_layouts_and_accumulation(config1, torch_cuda):
    torch = torch_cuda
    c = config1
    nq, K, B = c["nq"], c["K"], 100
    f = perturbed_maxwellians(c["L"], K, B, seed=5)
    fd = torch.from_numpy(f).cuda()
    base = torch.from_numpy(np.random.default_rng(0).standard_normal(B * K * nq)).cuda()
    Qn = base.clone()
    c["plan"].execute(Qn, fd, B)
    # batch-fastest layout: [N_Q*K][B]
    fT = fd.reshape(B, K * nq).t().contiguous()
    QT = base.reshape(B, K * nq).t().contiguous()
    c["plan"].execute(QT, fT, B, layout="batch_fastest")
    torch.cuda.synchronize()
    a = Qn.cpu().numpy()
    b = QT.t().contiguous().cpu().numpy().ravel()
    assert np.array_equal(a, b)
    Qref = oracle.angular_first_batch(base.cpu().numpy().copy(), f, c["lab"], c["val"], c["ic"], c["R"], nq, K,
                                      nthreads=8)
    assert rel_err(a, Qref, nq, K) <= TOL64


def test_overwrite_mode_writes_q_without_reading_it(config1, torch_cuda):
    """cfk_plan_set_accumulate(0): Q = Q(f) for every variant, precision, mixed kernel and layout, with
    Q's previous contents (NaN here) never read; equal to accumulating into zeros."""
    torch = torch_cuda
    c = config1
    nq, K, B = c["nq"], c["K"], 70
    f = perturbed_maxwellians(c["L"], K, B, seed=8)
    fd = torch.from_numpy(f).cuda()
    fT = fd.reshape(B, K * nq).t().contiguous()
    plan = c["plan"]
    runs = [("fp64", v, None) for v in ("angular_first", "radial_first", "naive")]
    runs += [("fp64_sym", "angular_first", None), ("mixed", "angular_first", 0), ("mixed", "angular_first", 1)]
    try:
        for precision, variant, pf in runs:
            if pf is not None:
                plan.set_tile("mixed", 0, 0, pf)
            for layout, fin in (("natural", fd), ("batch_fastest", fT)):
                print("overwrite: %s %s pf=%s %s" % (precision, variant, pf, layout), flush=True)
                plan.set_accumulate(True)
                Qa = torch.zeros(B * K * nq, dtype=torch.float64, device="cuda")
                plan.execute(Qa, fin, B, variant=variant, precision=precision, layout=layout)
                plan.set_accumulate(False)
                Qo = torch.full((B * K * nq,), float("nan"), dtype=torch.float64, device="cuda")
                plan.execute(Qo, fin, B, variant=variant, precision=precision, layout=layout)
                torch.cuda.synchronize()
                assert torch.equal(Qa, Qo), (precision, variant, pf, layout)
    finally:
        plan.set_accumulate(True)
        plan.set_tile("mixed", 0, 0, None)


def test_config1_sumfac_on_gpu_matches_golden():
    path = os.path.join(GOLDEN, "sumfac_hs_l8_k10_g48.npz")
    if not os.path.exists(path):
        pytest.skip("golden tensor not generated yet (tests/golden/make_sumfac_golden.py)")
    d = np.load(path)
    R = cfk.build_rtensor(9, 8, 1.0, tuple(int(v) for v in d["sizes"]), geometry=cfk.real_gaunt(8))
    assert np.max(np.abs(R - d["R"])) <= TOL64 * np.max(np.abs(d["R"]))


def test_autotuned_tiles_keep_parity(config1, torch_cuda):
    torch = torch_cuda
    c = config1
    nq, K, B = c["nq"], c["K"], 192
    tiles = c["plan"].autotune(B)
    assert set(tiles) == {"fp64", "mixed", "fp64_sym"}
    assert all(kc in (K, (K + 1) // 2) and nb in (1, 2) and pf in (0, 1) for kc, nb, pf in tiles.values()), tiles
    assert c["plan"].split("mixed") in (0, 1)
    f = perturbed_maxwellians(c["L"], K, B, seed=99)
    Qref = oracle.angular_first_batch(np.zeros(B * nq * K), f, c["lab"], c["val"], c["ic"], c["R"], nq, K,
                                      nthreads=8)
    fd = torch.from_numpy(f).cuda()
    for prec, tol in (("fp64", TOL64), ("fp64_sym", TOL64), ("mixed", TOL_MIXED)):
        Qd = torch.zeros(B * nq * K, dtype=torch.float64, device="cuda")
        c["plan"].execute(Qd, fd, B, precision=prec)
        torch.cuda.synchronize()
        assert rel_err(Qd.cpu().numpy(), Qref, nq, K) <= tol, prec


def test_every_mixed_tile_keeps_parity(config1, torch_cuda):
    """Each mixed-precision tile variant (k3 chunk, distributions per lane, radial stage on the VALU
    or on the matrix cores) against the oracle, on a ragged batch, natural and batch-fastest."""
    torch = torch_cuda
    c = config1
    nq, K, B = c["nq"], c["K"], 200
    f = perturbed_maxwellians(c["L"], K, B, seed=98)
    Qref = oracle.angular_first_batch(np.zeros(B * nq * K), f, c["lab"], c["val"], c["ic"], c["R"], nq, K,
                                      nthreads=8)
    fd = torch.from_numpy(f).cuda()
    fT = fd.reshape(B, -1).t().contiguous()
    for kc, nb, pf, wpb in ((K, 1, 0, 1), ((K + 1) // 2, 1, 0, 1), (K, 2, 0, 1), (K, 1, 0, 2), (K, 1, 0, 4),
                            (K, 2, 0, 4), (0, 1, 1, 1), (0, 1, 1, 2), (0, 1, 1, 4)):
        c["plan"].set_tile("mixed", kc, nb, pf)
        c["plan"].set_waves("mixed", wpb)
        c["plan"].set_split("mixed", 0)
        print("mixed tile", kc, nb, pf, wpb, flush=True)
        Qd = torch.zeros(B * nq * K, dtype=torch.float64, device="cuda")
        c["plan"].execute(Qd, fd, B, precision="mixed")
        QT = torch.zeros_like(fT)
        c["plan"].execute(QT, fT, B, precision="mixed", layout="batch_fastest")
        torch.cuda.synchronize()
        assert rel_err(Qd.cpu().numpy(), Qref, nq, K) <= TOL_MIXED, (kc, nb, pf)
        Qn = QT.t().contiguous().cpu().numpy().ravel()
        assert rel_err(Qn, Qref, nq, K) <= TOL_MIXED, (kc, nb, pf)
        Qn = Qn.reshape(B, K, nq)
        assert float(np.abs(Qn[:, 0, :4]).max()) == 0.0 and float(np.abs(Qn[:, 1, 0]).max()) == 0.0
    c["plan"].set_tile("mixed")
    c["plan"].set_waves("mixed")
    c["plan"].set_split("mixed")


def test_waves_per_workgroup_keep_parity(config1, torch_cuda):
    """FP64 and symmetrized FP64 at 1, 2 and 4 waves per workgroup on a ragged batch (the last
    workgroup holds tiles past the batch end) against the oracle."""
    torch = torch_cuda
    c = config1
    nq, K, B = c["nq"], c["K"], 64 * 5 + 17
    f = perturbed_maxwellians(c["L"], K, B, seed=97)
    Qref = oracle.angular_first_batch(np.zeros(B * nq * K), f, c["lab"], c["val"], c["ic"], c["R"], nq, K,
                                      nthreads=8)
    fd = torch.from_numpy(f).cuda()
    for prec in ("fp64", "fp64_sym"):
        for wpb in (1, 2, 4):
            c["plan"].set_waves(prec, wpb)
            assert c["plan"].waves(prec) == wpb
            print("waves", prec, wpb, flush=True)
            Qd = torch.zeros(B * nq * K, dtype=torch.float64, device="cuda")
            c["plan"].execute(Qd, fd, B, precision=prec)
            torch.cuda.synchronize()
            assert rel_err(Qd.cpu().numpy(), Qref, nq, K) <= TOL64, (prec, wpb)
        c["plan"].set_waves(prec)


def test_row_split_keeps_parity_and_is_deterministic(config1, torch_cuda):
    """Rows split between the two waves of a workgroup (symmetrized FP64 and mixed) on a ragged batch:
    against the oracle, bit-identical on a second run, and accumulating into a nonzero Q."""
    torch = torch_cuda
    c = config1
    nq, K, B = c["nq"], c["K"], 64 * 2 + 5   # 3 batch tiles: with 2 tiles per workgroup one is past the end
    f = perturbed_maxwellians(c["L"], K, B, seed=96)
    Qref = oracle.angular_first_batch(np.zeros(B * nq * K), f, c["lab"], c["val"], c["ic"], c["R"], nq, K,
                                      nthreads=8)
    fd = torch.from_numpy(f).cuda()
    fT = fd.reshape(B, -1).t().contiguous()
    for prec, tol, wpb, pf in (("fp64_sym", TOL64, 1, 0), ("mixed", TOL_MIXED, 1, 0), ("fp64_sym", TOL64, 2, 0),
                               ("mixed", TOL_MIXED, 2, 0), ("mixed", TOL_MIXED, 1, 1), ("mixed", TOL_MIXED, 2, 1)):
        if prec == "mixed":
            c["plan"].set_tile(prec, 0, 1, pf)   # pf = 1: radial stage on the matrix cores
        c["plan"].set_split(prec, 1)
        c["plan"].set_waves(prec, wpb)   # 2: two batch tiles per workgroup
        assert c["plan"].split(prec) == 1
        print("row split", prec, wpb, flush=True)
        Qd = torch.zeros(B * nq * K, dtype=torch.float64, device="cuda")
        c["plan"].execute(Qd, fd, B, precision=prec)
        Q1 = torch.zeros_like(fT)
        c["plan"].execute(Q1, fT, B, precision=prec, layout="batch_fastest")
        Q2 = torch.full_like(fT, 0.25)
        c["plan"].execute(Q2, fT, B, precision=prec, layout="batch_fastest")
        torch.cuda.synchronize()
        assert rel_err(Qd.cpu().numpy(), Qref, nq, K) <= tol, prec
        Qn = Q1.t().contiguous().cpu().numpy().ravel()
        assert rel_err(Qn, Qref, nq, K) <= tol, prec
        Q1b = torch.zeros_like(fT)
        c["plan"].execute(Q1b, fT, B, precision=prec, layout="batch_fastest")
        torch.cuda.synchronize()
        assert bool((Q1b == Q1).all()), prec
        assert bool(((Q2 - 0.25) - Q1).abs().max() <= 1e-12 * Q1.abs().max()), prec
        c["plan"].set_split(prec)
        c["plan"].set_waves(prec)
        c["plan"].set_tile(prec)
