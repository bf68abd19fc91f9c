"""ORACLE (test infrastructure only) -- numpy restatement of the surrogate network's training step.

Only tests may import this module, and only as a checker of the GPU kernels; the product path
never calls it.  There is no reference counterpart (the network is an addition); the oracle
restates the textbook MLP algebra of the library's kernels:

    Z_l = W_l H_{l-1} + b_l,  H_l = gelu(Z_l),  Y = W_L H_{L-1} + b_L,  loss = mean((Y - T)^2)
    gW_l = dZ_l H_{l-1}^T,  gb_l = sum_batch dZ_l,  dH_{l-1} = W_l^T dZ_l,  dZ_{l-1} = dH_{l-1} gelu'(Z_{l-1})

with gelu(z) = z Phi(z) (exact, erf form).  ``emulate=True`` inserts the BF16 roundings the GPU
path applies (weights, input, hidden activations, output gradient, hidden gradients) so that the
comparison isolates accumulation order; ``emulate=False`` is plain float64.

Pins (tests/test_oracle_mlp.py): float64 gradients against central finite differences of the
loss; gelu' against finite differences of gelu; the BF16 rounding against exhaustive
round-to-nearest-even on hand-picked bit patterns.
"""
import numpy as np
from scipy.special import erf


def bf16_round(x):
    """Round float values to the nearest BF16 (ties to even); returned as float64."""
    x32 = np.asarray(x, dtype=np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def split(p, n_in, width, depth, n_out):
    """[(W_l [out, in], b_l [out])] of the flat layer-by-layer parameter vector."""
    out, off = [], 0
    for l in range(depth + 1):
        fi = n_in if l == 0 else width
        fo = n_out if l == depth else width
        W = np.asarray(p[off:off + fi * fo], dtype=np.float64).reshape(fo, fi)
        off += fi * fo
        b = np.asarray(p[off:off + fo], dtype=np.float64)
        off += fo
        out.append((W, b))
    return out


def forward(p, X, shape, emulate=True):
    """Y [n_out, B] and the cache of pre-activations/activations; X is [n_in, B]."""
    n_in, width, depth, n_out = shape
    layers = split(p, *shape)
    r = bf16_round if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    H = [r(X)]
    Zs = []
    for l in range(depth):
        W, b = layers[l]
        Z = r(W) @ H[-1] + b[:, None]
        Zs.append(Z)
        H.append(r(gelu(Z)))
    W, b = layers[depth]
    Y = r(W) @ H[-1] + b[:, None]
    return Y, (H, Zs)


def loss_grad(p, X, T, shape, emulate=True):
    """(loss, flat gradient) of mean((forward(X) - T)^2), gradient in the parameter layout."""
    n_in, width, depth, n_out = shape
    layers = split(p, *shape)
    r = bf16_round if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    Y, (H, Zs) = forward(p, X, shape, emulate)
    n = Y.size
    D = Y - np.asarray(T, dtype=np.float64)
    loss = float(np.sum(D * D) / n)
    dZ = r(2.0 * D / n)
    grads = [None] * (depth + 1)
    for l in range(depth, -1, -1):
        gW = dZ @ H[l].T
        gb = dZ.sum(axis=1)
        grads[l] = (gW, gb)
        if l == 0:
            break
        W, _ = layers[l]
        dH = r(W).T @ dZ
        dZ = r(dH * gelu_grad(Zs[l - 1]))
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba, Algorithm 1); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
