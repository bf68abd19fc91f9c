"""Device-resident collision plans (batched contraction on the GPU).

A ``CollisionPlan`` uploads the Gaunt routing table and the radial tensor once; ``execute``
accumulates ``Q += Q(f)`` for a batch of distributions already resident in device memory
(contiguous float64 torch tensors on the GPU, checked before each call; the library only sees raw
device pointers).
"""
import ctypes

import numpy as np

from ._lib import LAYOUT, PRECISION, VARIANT, CfkError, _vp, check, dptr, f64, lib


def _ptr(t):
    return ctypes.c_void_p(int(t.data_ptr()))


def _check_dev_f64(t, need, name):
    """A device buffer the kernels read or write as `need` contiguous float64 values: a torch tensor
    must be float64, on the GPU and contiguous (the library sees only its data pointer and would
    otherwise read or write past its end)."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor on the GPU" % name)
    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError("%s must be a contiguous float64 tensor on the GPU (got %s, %s, contiguous=%s)"
                         % (name, t.dtype, t.device, t.is_contiguous()))
    if t.numel() < need:
        raise ValueError("%s too small: %d values, need %d" % (name, t.numel(), need))


def _stream_ptr(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(int(stream.cuda_stream))
    return ctypes.c_void_p(int(stream))


class CollisionPlan:
    """Routing tables + R on the current HIP device (cf. cfk_plan_create in include/cfk.h)."""

    def __init__(self, labels, vals, ic_map, R, N_Q, K):
        labels = np.asarray(labels)
        self.N_G = labels.shape[0]
        self.N_Q = int(N_Q)
        self.K = int(K)
        Rf = f64(R)
        self.N_L = Rf.size // (self.K ** 3)
        h = _vp()
        check(lib().cfk_plan_create(dptr(f64(labels)), dptr(f64(vals)), dptr(f64(ic_map)), self.N_G, dptr(Rf),
                                    Rf.size, self.N_Q, self.K, ctypes.byref(h)))
        self._h = h

    def execute(self, Q, f, B=None, variant="angular_first", precision="fp64", layout="natural", stream=None):
        """Q += Q(f) for B distributions (device buffers of float64); Q = Q(f) after set_accumulate(False)."""
        if self._h is None:
            raise CfkError(-1, "plan destroyed")
        n = self.N_Q * self.K
        if B is None:
            B = Q.numel() // n
        _check_dev_f64(Q, B * n, "Q")
        _check_dev_f64(f, B * n, "f")
        check(lib().cfk_plan_execute(self._h, VARIANT[variant], PRECISION[precision], LAYOUT[layout], _ptr(Q),
                                     _ptr(f), int(B), _stream_ptr(stream)))
        return Q

    def execute_static(self, Q, f, B=None, variant="angular_first", precision="fp64", layout="natural", stream=None):
        """As execute, through cfk_plan_execute_static: no device allocation, staging in the library's
        fixed per-GPU area (CFK_STATIC_SCRATCH_BYTES); overlapping calls on one GPU race on it."""
        if self._h is None:
            raise CfkError(-1, "plan destroyed")
        n = self.N_Q * self.K
        if B is None:
            B = Q.numel() // n
        _check_dev_f64(Q, B * n, "Q")
        _check_dev_f64(f, B * n, "f")
        check(lib().cfk_plan_execute_static(self._h, VARIANT[variant], PRECISION[precision], LAYOUT[layout], _ptr(Q),
                                            _ptr(f), int(B), _stream_ptr(stream)))
        return Q

    def linearize(self, f0=None, stream=None):
        """J = dQ/dc at f0 (default: the absolute Maxwellian), [N, N] with N = N_Q*K, index k*N_Q + q."""
        n = self.N_Q * self.K
        J = np.empty((n, n), order="F")
        f0a = f64(f0) if f0 is not None else None   # kept alive across the call
        f0p = dptr(f0a) if f0a is not None else None
        check(lib().cfk_plan_linearize(self._h, f0p, J.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       _stream_ptr(stream)))
        return J

    def set_tile(self, precision, kc=0, nb=0, pf=None):
        """Fix one precision's tile (0 = default).  For "mixed", pf=1 runs the radial stage on the matrix
        cores, pf=0 on the VALU, pf=None (-1) picks by K (matrix cores from K = 12 up)."""
        if pf is None:
            pf = -1 if precision == "mixed" else 0
        check(lib().cfk_plan_set_tile(self._h, PRECISION[precision], int(kc), int(nb), int(pf)))

    def set_waves(self, precision, wpb=0):
        """Waves per workgroup (1, 2 or 4 batch tiles of one output row per CU; 0 = default)."""
        check(lib().cfk_plan_set_waves(self._h, PRECISION[precision], int(wpb)))

    def waves(self, precision):
        """The waves per workgroup the plan uses for `precision`."""
        w = ctypes.c_int32(0)
        check(lib().cfk_plan_get_waves(self._h, PRECISION[precision], ctypes.byref(w)))
        return int(w.value)

    def set_accumulate(self, accumulate=True):
        """Output mode of execute: Q += Q(f) (True, the MEX semantics) or Q = Q(f) (False)."""
        check(lib().cfk_plan_set_accumulate(self._h, 1 if accumulate else 0))

    def set_split(self, precision, split=-1):
        """Row split between two waves per batch tile (symmetrized precisions; 0 = off, 1 = on, -1 = the
        selected kernel's default)."""
        check(lib().cfk_plan_set_split(self._h, PRECISION[precision], int(split)))

    def split(self, precision):
        """Whether the plan splits rows between two waves for `precision`."""
        v = ctypes.c_int32(0)
        check(lib().cfk_plan_get_split(self._h, PRECISION[precision], ctypes.byref(v)))
        return int(v.value)

    def autotune(self, B, stream=None):
        """Pick the fastest tile per precision for batches of B; returns {precision: (KC, NB, PF)}."""
        t = (ctypes.c_int32 * 9)()
        check(lib().cfk_plan_autotune(self._h, int(B), _stream_ptr(stream), t))
        return {"fp64": tuple(t[0:3]), "mixed": tuple(t[3:6]), "fp64_sym": tuple(t[6:9])}

    def stats(self):
        """Routing-table counts: N_G, slices, and for the symmetrized tables nonzeros, slices, q2 groups,
        symmetric (l2 == l3) slices, precise (l2 == 0) slices; N_L (cfk_plan_get_stats)."""
        v = (ctypes.c_int64 * 8)()
        check(lib().cfk_plan_get_stats(self._h, v))
        keys = ("N_G", "S", "sym_N_G", "sym_S", "sym_groups", "sym_symmetric", "sym_precise", "N_L")
        return dict(zip(keys, (int(x) for x in v)))

    def set_timing(self, enable=True):
        """Bracket each execute's main kernel with HIP events (see cfk_plan_set_timing)."""
        check(lib().cfk_plan_set_timing(self._h, 1 if enable else 0))

    def read_timing(self):
        """(total kernel ms, launches) since the last read; waits for the recorded events."""
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        check(lib().cfk_plan_read_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cfk_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dense_assemble(C, labels, vals, ic_map, R, N_Q, K, stream=None):
    """Write the dense tensor C[(k1 N_Q + q1) + (k2 N_Q + q2) N + (k3 N_Q + q3) N^2] = R_t[k1,k2,k3] g of the
    COO (host arrays, 1-based) into the device tensor C (N^3 float64 values, N = N_Q K), as
    assemble_full_tensor (src/GeneralCollisionTensor.m:283-311)."""
    labels = np.asarray(labels)
    if labels.ndim != 2 or labels.shape[1] != 3:
        raise ValueError("labels must be N_G x 3")
    n_g = labels.shape[0]
    if np.asarray(vals).size != n_g or np.asarray(ic_map).size != n_g:
        raise ValueError("vals and ic_map must have N_G = %d entries" % n_g)
    N = int(N_Q) * int(K)
    _check_dev_f64(C, N ** 3, "C")
    Rf = f64(R)
    check(lib().cfk_dense_assemble(_ptr(C), dptr(f64(labels)), dptr(f64(vals)), dptr(f64(ic_map)), n_g,
                                   dptr(Rf), Rf.size, int(N_Q), int(K), _stream_ptr(stream)))
    return C


def dense_execute(Q, f, C, N, B=1, stream=None):
    _check_dev_f64(Q, int(B) * int(N), "Q")
    _check_dev_f64(f, int(B) * int(N), "f")
    _check_dev_f64(C, int(N) ** 3, "C")
    check(lib().cfk_dense_execute(_ptr(Q), _ptr(f), _ptr(C), int(N), int(B), _stream_ptr(stream)))
    return Q


def conservation_project(Q, B, N_Q, K, layout="natural", stream=None):
    _check_dev_f64(Q, int(B) * int(N_Q) * int(K), "Q")
    check(lib().cfk_conservation_project(_ptr(Q), int(B), int(N_Q), int(K), LAYOUT[layout], _stream_ptr(stream)))
    return Q


class MultiPlan:
    """Several CollisionPlans (different L, K) executed together in mixed precision: heaviest-first
    work items, one item kernel per radial size K on parallel graph branches, captured into a
    hipGraph and replayed (cfk_multi_*)."""

    def __init__(self, plans):
        self.plans = list(plans)   # keep the plans alive
        arr = (_vp * len(self.plans))(*[p._h for p in self.plans])
        h = _vp()
        check(lib().cfk_multi_create(arr, len(self.plans), ctypes.byref(h)))
        self._h = h

    def set_graph(self, use_graph):
        """Replay a captured hipGraph (True, default) or enqueue the launches directly (False)."""
        check(lib().cfk_multi_set_graph(self._h, 1 if use_graph else 0))

    def execute(self, Qs, fs, Bs, stream=None):
        """Q_g += Q(f_g) for every group g; buffers are device tensors in batch-fastest layout."""
        n = len(self.plans)
        if len(Qs) != n or len(fs) != n or len(Bs) != n:
            raise ValueError("expected %d groups" % n)
        for i, (p, q, f, b) in enumerate(zip(self.plans, Qs, fs, Bs)):
            need = p.N_Q * p.K * int(b)
            _check_dev_f64(q, need, "Q[%d]" % i)
            _check_dev_f64(f, need, "f[%d]" % i)
        Qa = (_vp * n)(*[int(q.data_ptr()) for q in Qs])
        fa = (_vp * n)(*[int(f.data_ptr()) for f in fs])
        Ba = (ctypes.c_int64 * n)(*[int(b) for b in Bs])
        check(lib().cfk_multi_execute(self._h, Qa, fa, Ba, _stream_ptr(stream)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cfk_multi_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

