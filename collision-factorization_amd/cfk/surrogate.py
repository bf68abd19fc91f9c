"""Learned surrogate of the collision operator and its data (cfk_mlp_*, cfk_project_maxwellians,
cfk_surrogate_* in include/cfk.h).  No counterpart in the reference.

Device arrays are torch tensors on ``cuda:0`` (plumbing only: every compute call goes through
libcfk).  Distributions and collision terms are batch-fastest ``[K*N_Q, B]`` float64 arrays, row
``k*N_Q + q``; network arrays are feature-major float32 ``[features, B]``.
"""
import ctypes
import math

import numpy as np

from ._lib import CfkError, RelaxStats, _vp, check, lib
from .plan import _ptr, _stream_ptr

RELAX_MODES = {"exact": 0, "surrogate": 1, "linear": 2}


def maxwellian_mixture_params(B, seed, max_components=3, drift_max=0.5, temp_range=(0.5, 2.0)):
    """[B, 3, 7] float64 parameters (n, ux, uy, uz, Tx, Ty, Tz) of B seeded mixtures of 1..3
    anisotropic Maxwellians: component count uniform in 1..max_components, weights from a flat
    Dirichlet, drifts uniform in the ball |u| <= drift_max, axis temperatures uniform in temp_range.
    Densities are scaled to sum to pi^(3/4), which makes the mass coefficient c_000 equal to 1.
    Unused component slots have n = 0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ncomp = rng.integers(1, max_components + 1, size=B)
    p = np.zeros((B, 3, 7))
    p[:, :, 4:7] = 1.0
    w = rng.gamma(1.0, size=(B, 3))
    active = np.arange(3)[None, :] < ncomp[:, None]
    w = np.where(active, w, 0.0)
    w /= w.sum(axis=1, keepdims=True)
    p[:, :, 0] = w * math.pi ** 0.75
    d = rng.standard_normal((B, 3, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    rad = drift_max * rng.random((B, 3)) ** (1.0 / 3.0)
    p[:, :, 1:4] = np.where(active[:, :, None], d * rad[:, :, None], 0.0)
    T = rng.uniform(temp_range[0], temp_range[1], size=(B, 3, 3))
    p[:, :, 4:7] = np.where(active[:, :, None], T, 1.0)
    return p


def project_maxwellians(params, L, K, grid=(60, 30, 30), stream=None):
    """Basis coefficients [K*N_Q, B] (float64, cuda) of the mixtures in ``params`` [B, n_comp, 7]
    on the projection grid (n_rad, n_pol, n_azi) of tutorial_collisions_spectral_properties.m."""
    import torch
    params = np.ascontiguousarray(params, dtype=np.float64)
    B, n_comp = params.shape[0], params.shape[1]
    dp = torch.from_numpy(params).cuda()
    c = torch.empty(((L + 1) ** 2 * K, B), dtype=torch.float64, device="cuda")
    check(lib().cfk_project_maxwellians(_ptr(c), _ptr(dp), B, n_comp, L, K, int(grid[0]), int(grid[1]), int(grid[2]),
                                        _stream_ptr(stream)))
    return c


class Mlp:
    """The network of include/cfk.h's cfk_mlp (owning, or a view of a surrogate's network)."""

    def __init__(self, n_in=None, width=None, depth=None, n_out=None, seed=0, handle=None, owner=None):
        if handle is not None:
            self._h, self._own, self._owner = handle, False, owner
        else:
            h = _vp()
            check(lib().cfk_mlp_create(int(n_in), int(width), int(depth), int(n_out), ctypes.c_uint64(seed),
                                       ctypes.byref(h)))
            self._h, self._own, self._owner = h, True, None
        n = ctypes.c_int64()
        check(lib().cfk_mlp_num_params(self._h, ctypes.byref(n)))
        self.n_params = n.value

    def get_params(self):
        out = np.empty(self.n_params, dtype=np.float32)
        check(lib().cfk_mlp_get_params(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError("expected %d parameters" % self.n_params)
        check(lib().cfk_mlp_set_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def get_grads(self):
        out = np.empty(self.n_params, dtype=np.float32)
        check(lib().cfk_mlp_get_grads(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def forward(self, X, stream=None):
        """Y [n_out, B] float32 of X [n_in, B] float32 (device tensors)."""
        import torch
        B = X.shape[1]
        Y = torch.empty((self._n_out(), B), dtype=torch.float32, device=X.device)
        check(lib().cfk_mlp_forward(self._h, _ptr(Y), _ptr(X.contiguous()), B, _stream_ptr(stream)))
        return Y

    def loss_grad(self, X, T, stream=None):
        """Mean squared error of forward(X) against T; leaves the gradients for get_grads/adam_step."""
        loss = ctypes.c_double()
        check(lib().cfk_mlp_loss_grad(self._h, _ptr(X.contiguous()), _ptr(T.contiguous()), X.shape[1],
                                      ctypes.byref(loss), _stream_ptr(stream)))
        return loss.value

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        check(lib().cfk_mlp_adam_step(self._h, lr, beta1, beta2, eps, _stream_ptr(stream)))

    def _n_out(self):
        # the output layer's bias is the last n_out parameters; recover n_out from the layout
        return self._shape[3]

    def shape_from(self, n_in, width, depth, n_out):
        self._shape = (n_in, width, depth, n_out)
        return self

    def close(self):
        if self._own and getattr(self, "_h", None) is not None and self._h.value:
            lib().cfk_mlp_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_mlp(n_in, width, depth, n_out, seed=0):
    return Mlp(n_in, width, depth, n_out, seed).shape_from(n_in, width, depth, n_out)


def split_params(p, n_in, width, depth, n_out):
    """[(W_l [out, in], b_l [out])] views of a flat parameter vector (layer by layer)."""
    out, off = [], 0
    for l in range(depth + 1):
        fi = n_in if l == 0 else width
        fo = n_out if l == depth else width
        W = p[off:off + fi * fo].reshape(fo, fi)
        off += fi * fo
        b = p[off:off + fo]
        off += fo
        out.append((W, b))
    return out


class Surrogate:
    """Q~(f) = n^2 (J d + N(d)) with the exact linear part and a trained network (cfk_surrogate_*)."""

    def __init__(self, plan, width=1024, depth=3, seed=0, anchored=False):
        """anchored=True: decomposition about each cell's own Maxwellian (CFK_SURROGATE_ANCHORED in
        include/cfk.h), the network seeing 2 K N_Q + 1 inputs."""
        self.plan = plan   # the plan must outlive the surrogate
        self.K, self.N_Q = plan.K, plan.N_Q
        self.L = int(round(math.sqrt(plan.N_Q))) - 1
        self.NK = self.K * self.N_Q
        self.width, self.depth = width, depth
        self.anchored = bool(anchored)
        self.n_in = 2 * self.NK + 1 if self.anchored else self.NK
        h = _vp()
        check(lib().cfk_surrogate_create_ex(plan._h, int(width), int(depth), ctypes.c_uint64(seed),
                                            1 if self.anchored else 0, ctypes.byref(h)))
        self._h = h
        m = _vp()
        check(lib().cfk_surrogate_mlp(self._h, ctypes.byref(m)))
        self.mlp = Mlp(handle=m, owner=self).shape_from(self.n_in, width, depth, self.NK)

    def linear_blocks(self):
        out = np.empty((self.L + 1, self.K, self.K))
        check(lib().cfk_surrogate_linear_blocks(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def normalization(self):
        arrs = [np.empty(n, dtype=np.float32) for n in (self.n_in, self.n_in, self.NK, self.NK)]
        fp = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for a in arrs]
        check(lib().cfk_surrogate_get_normalization(self._h, *fp))
        return dict(zip(("mu_x", "inv_s_x", "mu_y", "s_y"), arrs))

    def fit(self, f, Q, epochs=40, batch=1024, lr=1e-3, lr_final=1e-5, seed=0, global_input_scale=False,
            global_target_scale=True, stream=None):
        """Train on pairs (f, Q(f)), [K*N_Q, n] float64 device tensors; returns the epoch losses.

        One scale for all targets (the default) weights the loss by the targets' own magnitudes,
        which is what the relative error of Q~ measures (held-out error 2.7% against 5.5% with
        per-feature target scales, hard spheres L=8, K=10, 2e5 mixtures, 40 epochs)."""
        n = f.shape[1]
        if f.shape != Q.shape or f.shape[0] != self.NK:
            raise ValueError("f and Q must both be [%d, n]" % self.NK)
        hist = np.zeros(epochs)
        check(lib().cfk_surrogate_fit(self._h, _ptr(f.contiguous()), _ptr(Q.contiguous()), n, int(epochs), int(batch),
                                      float(lr), float(lr_final), ctypes.c_uint64(seed),
                                      (1 if global_input_scale else 0) | (2 if global_target_scale else 0),
                                      hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _stream_ptr(stream)))
        return hist

    def apply(self, f, use_network=True, out=None, stream=None):
        """Q~(f) [K*N_Q, B] float64 (projected)."""
        import torch
        B = f.shape[1]
        if out is None:
            out = torch.empty_like(f)
        check(lib().cfk_surrogate_apply(self._h, _ptr(out), _ptr(f.contiguous()), B, 1 if use_network else 0,
                                        _stream_ptr(stream)))
        return out

    def anchors(self, f, stream=None):
        """(c_M, Q(c_M)) [K*N_Q, B] float64: each cell's isotropic Maxwellian on the basis (the cell's n, u,
        T) and its exact operator value."""
        import torch
        if f.dtype != torch.float64 or not f.is_cuda or f.dim() != 2:
            raise ValueError("f must be a 2-D float64 cuda tensor [K*N_Q, B]")
        f = f.contiguous()
        cM, QM = torch.empty_like(f), torch.empty_like(f)
        check(lib().cfk_surrogate_anchors(self._h, _ptr(cM), _ptr(QM), _ptr(f), f.shape[1], _stream_ptr(stream)))
        return cM, QM

    def relax(self, c, dt, n_steps, mode="surrogate", tol=1e-2, sample_stride=100, stream=None):
        """Forward-Euler relaxation of c [K*N_Q, B] in place; returns (stats dict, residual per step)."""
        if not c.is_contiguous():
            raise ValueError("c must be contiguous")
        hist = np.zeros(max(n_steps, 1))
        st = RelaxStats()
        check(lib().cfk_surrogate_relax(self._h, _ptr(c), c.shape[1], float(dt), int(n_steps), RELAX_MODES[mode],
                                        float(tol), int(sample_stride),
                                        hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(st),
                                        _stream_ptr(stream)))
        stats = {k: getattr(st, k) for k, _ in RelaxStats._fields_}
        return stats, hist[:n_steps]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cfk_surrogate_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["CfkError", "Mlp", "Surrogate", "make_mlp", "maxwellian_mixture_params", "project_maxwellians",
           "split_params", "RELAX_MODES"]
