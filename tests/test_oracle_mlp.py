"""Pins of the network oracle (oracle/mlp.py) used to check the GPU training kernels.

* float64 gradients of the loss against central finite differences (a transposed operand, a
  dropped bias gradient or a wrong sign fails);
* gelu' against a finite difference of gelu;
* the BF16 rounding against torch's float32 -> bfloat16 conversion (round to nearest, ties to
  even) and against hand-picked ties;
* the BF16-emulating gradient stays within BF16 accuracy of the float64 one;
* the Adam step against a hand-derived first step (|step| = lr for every nonzero gradient).
"""
import numpy as np
import pytest

from oracle import mlp as M

SHAPE = (5, 7, 2, 4)   # n_in, width, depth, n_out


def _params(seed, shape=SHAPE):
    n_in, width, depth, n_out = shape
    rng = np.random.default_rng(seed)
    n = sum((n_in if l == 0 else width) * (n_out if l == depth else width) + (n_out if l == depth else width)
            for l in range(depth + 1))
    return rng.standard_normal(n) * 0.5


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(0)
    X = rng.standard_normal((SHAPE[0], 3))
    T = rng.standard_normal((SHAPE[3], 3))
    _, g = M.loss_grad(p, X, T, SHAPE, emulate=False)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (M.loss_grad(p + e, X, T, SHAPE, emulate=False)[0] - M.loss_grad(p - e, X, T, SHAPE, emulate=False)[0]) / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_gelu_grad_matches_finite_difference():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(M.gelu_grad(z) - fd)) < 1e-8
    # gelu(z) = z Phi(z): known values
    assert abs(M.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15


def test_bf16_rounding_is_round_to_nearest_even():
    import torch
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-20, 20, 20000),
                        [0.0, -0.0, 1.0, 65504.0, 1e-30]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16_round(x), ref)
    # ties: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7 -> even (1); 1 + 3*2^-8 -> 1 + 2^-6
    assert M.bf16_round(np.float32(1 + 2.0 ** -8)) == 1.0
    assert M.bf16_round(np.float32(1 + 3 * 2.0 ** -8)) == 1 + 2.0 ** -6


def test_emulated_gradient_is_close_to_float64():
    shape = (40, 64, 3, 30)
    rng = np.random.default_rng(5)
    p = _params(2, shape) * 0.3
    X = rng.standard_normal((40, 32))
    T = rng.standard_normal((30, 32))
    _, g64 = M.loss_grad(p, X, T, shape, emulate=False)
    _, g16 = M.loss_grad(p, X, T, shape, emulate=True)
    assert np.linalg.norm(g16 - g64) <= 3e-2 * np.linalg.norm(g64)


def test_adam_first_step():
    p = np.array([1.0, -2.0, 0.5, 3.0])
    g = np.array([0.1, -4.0, 0.0, 1e-3])
    q, m, v = M.adam(p, np.zeros(4), np.zeros(4), g, 1, lr=0.01, eps=0.0 + 1e-12)
    # after bias correction m_hat = g, v_hat = g^2: each nonzero coordinate moves by lr * sign(g)
    assert np.allclose(q[[0, 1, 3]], p[[0, 1, 3]] - 0.01 * np.sign(g[[0, 1, 3]]), rtol=0, atol=1e-9)
    assert q[2] == p[2]
