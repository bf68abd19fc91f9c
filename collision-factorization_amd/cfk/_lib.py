"""ctypes binding of libcfk.so (the HIP library built from ../csrc by ``make -C collision-factorization_amd``).

There is no CPU fallback anywhere in this package: if the library is missing, or no HIP device is
visible, the compute calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcfk.so")

CFK_OK = 0
CFK_EINVAL = -1
CFK_EHIP = -2
CFK_ENOMEM = -3
CFK_ENODEV = -4
CFK_EUNSUP = -5

VARIANT = {"angular_first": 0, "radial_first": 1, "naive": 2}
PRECISION = {"fp64": 0, "mixed": 1, "fp64_sym": 2}
LAYOUT = {"natural": 0, "batch_fastest": 1}


class CfkError(RuntimeError):
    """A libcfk call returned a non-zero status; ``code`` is the CFK_E* value."""

    def __init__(self, code, msg):
        super().__init__("libcfk error %d: %s" % (code, msg))
        self.code = code


_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_vp = ctypes.c_void_p

class RelaxStats(ctypes.Structure):
    """cfk_relax_stats of include/cfk.h."""
    _fields_ = [("steps_network", ctypes.c_int64), ("steps_exact", ctypes.c_int64), ("exact_evals", ctypes.c_int64),
                ("max_residual", ctypes.c_double), ("last_residual", ctypes.c_double)]


# (name, argtypes) of every exported entry point of include/cfk.h
SIGNATURES = {
    "cfk_last_error": ([], ctypes.c_char_p),
    "cfk_version": ([], ctypes.c_int),
    "cfk_device_count": ([], ctypes.c_int),
    "cfk_angular_first_collision_kernel": ([_dp, _dp, _dp, _dp, _dp, _dp, _i64, _i64, _i32, _i32], ctypes.c_int),
    "cfk_radial_first_collision_kernel": ([_dp, _dp, _dp, _dp, _dp, _dp, _i64, _i64, _i32, _i32], ctypes.c_int),
    "cfk_naive_collision_kernel": ([_dp, _dp, _dp, _dp, _dp, _dp, _i64, _i64, _i32, _i32], ctypes.c_int),
    "cfk_dense_tensor_kernel": ([_dp, _dp, _dp, _i64], ctypes.c_int),
    "cfk_compute_rtensor_sumfac": ([_dp, _i32, _i32, _i32, ctypes.c_double] + [_dp, _dp, _i32] * 7 + [_dp] * 8
                                   + [_i32, _i32], ctypes.c_int),
    "cfk_real_gaunt": ([_i32, ctypes.POINTER(_i64), _dp, _dp, _dp], ctypes.c_int),
    "cfk_gauss_legendre": ([_i32, ctypes.c_double, ctypes.c_double, _dp, _dp], ctypes.c_int),
    "cfk_gauss_gen_laguerre": ([_i32, ctypes.c_double, _dp, _dp], ctypes.c_int),
    "cfk_sumfac_set_kernel": ([_i32], ctypes.c_int),
    "cfk_sumfac_set_chunk_nodes": ([_i32], ctypes.c_int),
    "cfk_sumfac_last_chunks": ([ctypes.POINTER(_i64)], ctypes.c_int),
    "cfk_h5_read_f64": ([ctypes.c_char_p, ctypes.c_char_p, _dp, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_i32),
                         ctypes.POINTER(_i64)], ctypes.c_int),
    "cfk_mat_write_f64": ([ctypes.c_char_p, _i32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_i32),
                           ctypes.POINTER(_i64), ctypes.POINTER(_dp)], ctypes.c_int),
    "cfk_sumfac_last_timing": ([ctypes.POINTER(_i32), ctypes.POINTER(_i64), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "cfk_sumfac_grid_sizes": ([_i32, _i32, _i32, _i32, ctypes.POINTER(_i32)], ctypes.c_int),
    "cfk_build_rtensor": ([_dp, _i32, _i32, ctypes.c_double, ctypes.POINTER(_i32), _dp, _dp, _dp, _i64],
                          ctypes.c_int),
    "cfk_plan_create": ([_dp, _dp, _dp, _i64, _dp, _i64, _i32, _i32, ctypes.POINTER(_vp)], ctypes.c_int),
    "cfk_plan_destroy": ([_vp], ctypes.c_int),
    "cfk_plan_execute_static": ([_vp, _i32, _i32, _i32, _vp, _vp, _i64, _vp], ctypes.c_int),
    "cfk_plan_get_stats": ([_vp, ctypes.POINTER(_i64)], ctypes.c_int),
    "cfk_plan_execute": ([_vp, _i32, _i32, _i32, _vp, _vp, _i64, _vp], ctypes.c_int),
    "cfk_plan_autotune": ([_vp, _i64, _vp, ctypes.POINTER(_i32)], ctypes.c_int),
    "cfk_plan_set_tile": ([_vp, _i32, _i32, _i32, _i32], ctypes.c_int),
    "cfk_plan_set_waves": ([_vp, _i32, _i32], ctypes.c_int),
    "cfk_plan_get_waves": ([_vp, _i32, ctypes.POINTER(_i32)], ctypes.c_int),
    "cfk_plan_set_accumulate": ([_vp, _i32], ctypes.c_int),
    "cfk_plan_set_split": ([_vp, _i32, _i32], ctypes.c_int),
    "cfk_plan_get_split": ([_vp, _i32, ctypes.POINTER(_i32)], ctypes.c_int),
    "cfk_plan_linearize": ([_vp, _dp, _dp, _vp], ctypes.c_int),
    "cfk_plan_set_timing": ([_vp, _i32], ctypes.c_int),
    "cfk_plan_read_timing": ([_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i64)], ctypes.c_int),
    "cfk_multi_create": ([ctypes.POINTER(_vp), _i32, ctypes.POINTER(_vp)], ctypes.c_int),
    "cfk_multi_destroy": ([_vp], ctypes.c_int),
    "cfk_multi_set_graph": ([_vp, _i32], ctypes.c_int),
    "cfk_multi_execute": ([_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_i64), _vp], ctypes.c_int),
    "cfk_dense_execute": ([_vp, _vp, _vp, _i64, _i64, _vp], ctypes.c_int),
    "cfk_dense_assemble": ([_vp, _dp, _dp, _dp, _i64, _dp, _i64, _i32, _i32, _vp], ctypes.c_int),
    "cfk_basis_evaluate": ([_vp, _vp, _i64, _i32, _i32, _vp], ctypes.c_int),
    "cfk_moments": ([_vp, _vp, _i64, _i32, _i32, _i32, _vp], ctypes.c_int),
    "cfk_moment_weights": ([_i32, _i32, ctypes.POINTER(_i32), ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "cfk_conservation_project": ([_vp, _i64, _i32, _i32, _i32, _vp], ctypes.c_int),
    "cfk_mlp_create": ([_i32, _i32, _i32, _i32, ctypes.c_uint64, ctypes.POINTER(_vp)], ctypes.c_int),
    "cfk_mlp_destroy": ([_vp], ctypes.c_int),
    "cfk_mlp_num_params": ([_vp, ctypes.POINTER(_i64)], ctypes.c_int),
    "cfk_mlp_get_params": ([_vp, _fp], ctypes.c_int),
    "cfk_mlp_set_params": ([_vp, _fp], ctypes.c_int),
    "cfk_mlp_get_grads": ([_vp, _fp], ctypes.c_int),
    "cfk_mlp_forward": ([_vp, _vp, _vp, _i64, _vp], ctypes.c_int),
    "cfk_mlp_loss_grad": ([_vp, _vp, _vp, _i64, _dp, _vp], ctypes.c_int),
    "cfk_mlp_adam_step": ([_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp], ctypes.c_int),
    "cfk_project_maxwellians": ([_vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _i32, _vp], ctypes.c_int),
    "cfk_surrogate_create": ([_vp, _i32, _i32, ctypes.c_uint64, ctypes.POINTER(_vp)], ctypes.c_int),
    "cfk_surrogate_create_ex": ([_vp, _i32, _i32, ctypes.c_uint64, _i32, ctypes.POINTER(_vp)], ctypes.c_int),
    "cfk_surrogate_destroy": ([_vp], ctypes.c_int),
    "cfk_surrogate_mlp": ([_vp, ctypes.POINTER(_vp)], ctypes.c_int),
    "cfk_surrogate_linear_blocks": ([_vp, _dp], ctypes.c_int),
    "cfk_surrogate_get_normalization": ([_vp, _fp, _fp, _fp, _fp], ctypes.c_int),
    "cfk_surrogate_fit": ([_vp, _vp, _vp, _i64, _i32, _i32, ctypes.c_float, ctypes.c_float, ctypes.c_uint64, _i32, _dp,
                           _vp],
                          ctypes.c_int),
    "cfk_surrogate_apply": ([_vp, _vp, _vp, _i64, _i32, _vp], ctypes.c_int),
    "cfk_surrogate_anchors": ([_vp, _vp, _vp, _vp, _i64, _vp], ctypes.c_int),
    "cfk_surrogate_relax": ([_vp, _vp, _i64, ctypes.c_double, _i32, _i32, ctypes.c_double, _i32, _dp,
                             ctypes.POINTER(RelaxStats), _vp], ctypes.c_int),
}

_LIB = None


def _bind_hip_runtime():
    """Make sure one HIP runtime serves both libcfk and torch.

    torch ships its own libamdhip64.so.7 (same SONAME as ROCm's).  The dynamic linker binds
    libcfk's dependency to whichever copy is already loaded, so torch (when installed) is imported
    first and libcfk runs on torch's runtime; loading libcfk first would pull in the system copy and
    leave torch without a usable device.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def hip_runtimes_loaded():
    """Paths of the libamdhip64 copies mapped into this process (should be exactly one)."""
    import re
    with open("/proc/self/maps") as fh:
        return sorted(set(re.findall(r"(/\S*libamdhip64\S*)", fh.read())))


def lib():
    """Load libcfk.so once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        _bind_hip_runtime()
        if not os.path.exists(LIB_PATH):
            raise CfkError(CFK_ENODEV, "libcfk.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def check(rc):
    if rc != CFK_OK:
        raise CfkError(rc, lib().cfk_last_error().decode(errors="replace"))
    return rc


def dptr(a):
    """double* of a C-contiguous float64 numpy array."""
    assert isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def f64(a, order="F"):
    """Contiguous float64 copy/view of ``a`` flattened in MATLAB column-major order."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order=order))
