"""CPU checks of the C-ABI library: it loads, exports every entry point include/cfk.h declares,
and rejects bad arguments on the host before touching a device.  No compute call is made here."""
import os
import re

import numpy as np
import pytest

import cfk
from cfk._lib import SIGNATURES, CfkError, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "cfk.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cfk_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_what_the_binding_knows():
    assert declared_symbols() == sorted(SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = lib()   # loads torch's HIP runtime first, then libcfk
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert lib().cfk_version() >= 1


def test_host_validation_rejects_bad_indices_without_a_device():
    nq, K = 4, 2
    Q = np.zeros((nq, K), order="F")
    f = np.ones((nq, K), order="F")
    lab = np.array([[1.0, 1.0, 1.0], [5.0, 1.0, 1.0]])   # q1 = 5 > N_Q
    with pytest.raises(CfkError) as e:
        cfk.angular_first_collision_kernel_mex(Q, f, lab, [0.5, 0.5], [1.0, 1.0], np.ones(K ** 3), nq, K)
    assert e.value.code == -1 and "labels" in str(e.value)
    lab = np.array([[1.0, 1.0, 1.0]])
    with pytest.raises(CfkError) as e:   # ic_map beyond N_L = numel(R)/K^3
        cfk.naive_collision_kernel_mex(Q, f, lab, [0.5], [2.0], np.ones(K ** 3), nq, K)
    assert "ic_map" in str(e.value)
    with pytest.raises(CfkError):        # R not a multiple of K^3
        cfk.radial_first_collision_kernel_mex(Q, f, lab, [0.5], [1.0], np.ones(K ** 3 + 1), nq, K)
    with pytest.raises(CfkError):        # non-integral index
        cfk.angular_first_collision_kernel_mex(Q, f, np.array([[1.5, 1.0, 1.0]]), [0.5], [1.0], np.ones(K ** 3), nq, K)
    with pytest.raises(ValueError):      # Q not in MATLAB layout cannot be updated in place
        cfk.angular_first_collision_kernel_mex(np.zeros((nq, K)), f, lab, [0.5], [1.0], np.ones(K ** 3), nq, K)


def test_empty_coo_leaves_q_unchanged_without_a_device():
    """N_G = 0: the reference's loops add nothing, so Q keeps its contents (no device is touched)."""
    nq, K = 4, 2
    before = np.arange(nq * K, dtype=np.float64).reshape(nq, K)
    Q = np.asfortranarray(before)
    f = np.ones((nq, K), order="F")
    empty = np.zeros((0, 3))
    for fn in (cfk.angular_first_collision_kernel_mex, cfk.radial_first_collision_kernel_mex,
               cfk.naive_collision_kernel_mex):
        fn(Q, f, empty, np.zeros(0), np.zeros(0), np.zeros(0), nq, K)
        assert np.array_equal(Q, before)


def test_sumfac_arity_error_mirrors_the_mex():
    # compute_rtensor_sumfac_mex.cpp:82-84
    with pytest.raises(cfk.MexError) as e:
        cfk.compute_rtensor_sumfac_mex(*([1.0] * 25))
    assert e.value.identifier == "R_tensor:InvalidInput"


def test_single_hip_runtime_in_process():
    # libcfk must bind to the same libamdhip64 as torch (one runtime per process)
    from cfk._lib import hip_runtimes_loaded
    lib()
    assert len(hip_runtimes_loaded()) == 1, hip_runtimes_loaded()


def test_multiplan_and_plan_release_their_handles():
    """close() and __del__ are methods of both wrappers (the native objects own graphs, streams and
    device buffers that must be freed)."""
    for cls in (cfk.CollisionPlan, cfk.MultiPlan):
        assert callable(getattr(cls, "close", None)), cls
        assert callable(getattr(cls, "__del__", None)), cls


def test_device_buffer_checks_run_before_the_library():
    """Q/f buffers are checked for dtype, device, contiguity and size on the host: a float32, host or
    strided tensor of the right element count would otherwise be read past its end."""
    import torch
    from cfk.plan import _check_dev_f64
    with pytest.raises(TypeError):
        _check_dev_f64(np.zeros(8), 8, "Q")
    with pytest.raises(ValueError):
        _check_dev_f64(torch.zeros(8, dtype=torch.float64), 8, "Q")   # not on the GPU
    with pytest.raises(ValueError):
        _check_dev_f64(torch.zeros(8, dtype=torch.float32), 8, "f")
    with pytest.raises(ValueError):
        _check_dev_f64(torch.zeros(4, 4, dtype=torch.float64).t(), 16, "f")
