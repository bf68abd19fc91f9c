"""GPU network kernels (cfk_mlp_*, csrc/mlp.hip) against the numpy oracle (oracle/mlp.py).

The oracle emulates the BF16 roundings of the GPU path (weights, input, hidden activations,
output and hidden gradients), so the comparison isolates FP32 accumulation order; the bars are
relative Frobenius errors per layer (a transposed operand, a wrong MFMA lane mapping or a dropped
term gives O(1) errors).  Forward 2e-3, gradients 5e-3, Adam 1e-5 (FP32 arithmetic).
"""
import numpy as np
import pytest

import cfk
from oracle import mlp as M

pytestmark = pytest.mark.gpu

SHAPES = [(130, 200, 2, 70, 300),      # ragged: no size is a multiple of the 128 x 128 x 32 tiles
          (810, 1024, 3, 810, 256)]    # the surrogate's network


def _net(shape, seed):
    n_in, width, depth, n_out, _ = shape
    net = cfk.make_mlp(n_in, width, depth, n_out, seed=seed)
    rng = np.random.default_rng(seed)
    p = net.get_params()
    p += (rng.standard_normal(p.size) * 0.02).astype(np.float32)   # nonzero biases too
    net.set_params(p)
    return net, net.get_params().astype(np.float64)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("shape", SHAPES)
def test_forward_matches_oracle(shape, torch_cuda):
    torch = torch_cuda
    n_in, width, depth, n_out, B = shape
    print("mlp forward", shape, flush=True)
    net, p = _net(shape, 11)
    X = np.random.default_rng(1).standard_normal((n_in, B)).astype(np.float32)
    Y = net.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    Yo, _ = M.forward(p, X, shape[:4], emulate=True)
    assert _rel(Y, Yo) < 2e-3


@pytest.mark.parametrize("shape", SHAPES)
def test_gradients_match_oracle(shape, torch_cuda):
    torch = torch_cuda
    n_in, width, depth, n_out, B = shape
    print("mlp gradients", shape, flush=True)
    net, p = _net(shape, 12)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((n_in, B)).astype(np.float32)
    T = rng.standard_normal((n_out, B)).astype(np.float32)
    loss = net.loss_grad(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    g = net.get_grads().astype(np.float64)
    lo, go = M.loss_grad(p, X, T, shape[:4], emulate=True)
    assert abs(loss - lo) <= 1e-4 * lo
    for (gW, gb), (oW, ob) in zip(cfk.split_params(g, *shape[:4]), cfk.split_params(go, *shape[:4])):
        assert _rel(gW, oW) < 5e-3
        assert _rel(gb, ob) < 5e-3


def test_adam_step_matches_oracle(torch_cuda):
    torch = torch_cuda
    shape = SHAPES[0]
    n_in, width, depth, n_out, B = shape
    net, p = _net(shape, 13)
    rng = np.random.default_rng(3)
    X = torch.from_numpy(rng.standard_normal((n_in, B)).astype(np.float32)).cuda()
    T = torch.from_numpy(rng.standard_normal((n_out, B)).astype(np.float32)).cuda()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in (1, 2, 3):
        net.loss_grad(X, T)
        g = net.get_grads().astype(np.float64)
        net.adam_step(1e-3)
        p, m, v = M.adam(p, m, v, g, step, 1e-3)
        q = net.get_params().astype(np.float64)
        assert np.max(np.abs(q - p)) <= 1e-5 * max(1.0, np.max(np.abs(p)))
        p = q


def test_training_reduces_loss(torch_cuda):
    torch = torch_cuda
    n_in, n_out, B = 64, 32, 512
    net = cfk.make_mlp(n_in, 128, 2, n_out, seed=4)
    rng = np.random.default_rng(4)
    A = rng.standard_normal((n_out, n_in)) / np.sqrt(n_in)
    X = rng.standard_normal((n_in, B))
    T = np.tanh(A @ X) + 0.1 * (A @ X) ** 2
    Xd = torch.from_numpy(X.astype(np.float32)).cuda()
    Td = torch.from_numpy(T.astype(np.float32)).cuda()
    first = net.loss_grad(Xd, Td)
    for _ in range(300):
        net.loss_grad(Xd, Td)
        net.adam_step(2e-3)
    last = net.loss_grad(Xd, Td)
    assert last < 0.05 * first, (first, last)
