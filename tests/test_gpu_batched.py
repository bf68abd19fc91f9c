"""Batched uneven sizes (cfk_multi_*): distributions with L in {4,6,8,12}, K = L+2, hard spheres,
mixed precision, one item kernel per K captured in a hipGraph (or enqueued directly).  Checked per
group against the plain FP64 GPU operator (itself parity-tested against the oracle), on several
graph replays with eager work in between.
"""
import numpy as np
import pytest

import cfk
from cfk.inputs import perturbed_maxwellians

pytestmark = pytest.mark.gpu

GROUPS = [(4, 6), (6, 8), (8, 10), (12, 14)]
TOL_MIXED = 1e-6


@pytest.fixture(scope="module")
def groups(torch_cuda):
    out = []
    for L, K in GROUPS:
        print("group L=%d K=%d: geometry + R (sumfac, pad 0)" % (L, K), flush=True)
        lab, val, ic = cfk.real_gaunt(L)
        ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
        R = cfk.build_rtensor(K - 1, L, 1.0, cfk.sumfac_grid_sizes(K - 1, L, 0, 0), geometry=(lab, val, ic))
        plan = cfk.CollisionPlan(ls, vs, ics, R, (L + 1) ** 2, K)
        out.append(dict(L=L, K=K, nq=(L + 1) ** 2, plan=plan))
    return out


@pytest.mark.parametrize("use_graph", [True, False])
def test_multi_matches_fp64_on_replays(groups, torch_cuda, use_graph):
    torch = torch_cuda
    Bs = [200, 150, 130, 70]
    fs, Qs, refs = [], [], []
    for i, (g, B) in enumerate(zip(groups, Bs)):
        f = perturbed_maxwellians(g["L"], g["K"], B, seed=100 + i)
        fT = torch.from_numpy(f).cuda().reshape(B, -1).t().contiguous()
        ref = torch.zeros_like(fT)
        g["plan"].execute(ref, fT, B, precision="fp64", layout="batch_fastest")
        fs.append(fT)
        Qs.append(torch.zeros_like(fT))
        refs.append(ref)
    torch.cuda.synchronize()
    mp = cfk.MultiPlan([g["plan"] for g in groups])
    mp.set_graph(use_graph)
    for rep in range(3):
        for Q in Qs:
            Q.zero_()
        print("multi replay %d graph=%s" % (rep, use_graph), flush=True)
        mp.execute(Qs, fs, Bs)
        torch.cuda.synchronize()
        for g, Q, ref in zip(groups, Qs, refs):
            err = float(((Q - ref).abs().amax(dim=0) / ref.abs().amax(dim=0)).max())
            assert err <= TOL_MIXED, (g["L"], rep, err)
            nq = g["nq"]
            assert float(Q[[0, 1, 2, 3, nq]].abs().max()) == 0.0
        # eager work on the same buffers between replays
        extra = torch.zeros_like(Qs[2])
        groups[2]["plan"].execute(extra, fs[2], Bs[2], precision="mixed", layout="batch_fastest")
        torch.cuda.synchronize()
    # accumulation semantics: a second replay without zeroing doubles Q
    for Q in Qs:
        Q.zero_()
    mp.execute(Qs, fs, Bs)
    mp.execute(Qs, fs, Bs)
    torch.cuda.synchronize()
    for Q, ref in zip(Qs, refs):
        err = float(((Q - 2 * ref).abs().amax(dim=0) / (2 * ref).abs().amax(dim=0)).max())
        assert err <= TOL_MIXED


def test_multi_rejects_unsupported_groups(torch_cuda):
    lab, val, ic = cfk.real_gaunt(2)
    ls, vs, ics = cfk.sort_t_q1(lab, val, ic)
    R = np.random.default_rng(0).standard_normal(int(ic.max()) * 5 ** 3)
    p = cfk.CollisionPlan(ls, vs, ics, R, 9, 5)    # K=5 has no fused-kernel path
    with pytest.raises(cfk.CfkError):
        cfk.MultiPlan([p])


def test_matrix_core_radial_stage_per_k(groups, torch_cuda):
    """The mixed path with the radial stage on the matrix cores, for every supported K, against the
    FP64 GPU operator (ragged batches)."""
    torch = torch_cuda
    for i, g in enumerate(groups):
        B = 130 + 7 * i
        f = perturbed_maxwellians(g["L"], g["K"], B, seed=200 + i)
        fT = torch.from_numpy(f).cuda().reshape(B, -1).t().contiguous()
        ref = torch.zeros_like(fT)
        g["plan"].execute(ref, fT, B, precision="fp64", layout="batch_fastest")
        g["plan"].set_tile("mixed", 0, 1, 1)
        print("matrix-core radial stage L=%d K=%d" % (g["L"], g["K"]), flush=True)
        Q = torch.zeros_like(fT)
        g["plan"].execute(Q, fT, B, precision="mixed", layout="batch_fastest")
        torch.cuda.synchronize()
        g["plan"].set_tile("mixed")
        err = float(((Q - ref).abs().amax(dim=0) / ref.abs().amax(dim=0)).max())
        assert err <= TOL_MIXED, (g["L"], err)
