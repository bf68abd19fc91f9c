"""The native MATLAB v7.3 (.mat / HDF5) reader against the reference's precalc files.

Two of the reference's files (src/precalc, copied by tests/golden/extract_precalc.py) are read
through cfk_h5_read_f64 and compared bit for bit with what h5dump extracted from them (the
precalc_k4_l*_g*.npz fixtures): L=2 (R in one chunk, the Gaunt COO in compact datasets) and L=4
(R and the Gaunt labels in deflate-compressed chunks).  Error paths: a file that is not HDF5, a missing
dataset, a dataset that is not float64, a too-small output buffer."""
import ctypes
import os

import numpy as np
import pytest

import cfk
from conftest import GOLDEN

CASES = [("collisiontensor_k4_l2_gamma0.00.mat", "precalc_k4_l2_g0.npz"),
         ("collisiontensor_k4_l4_gamma1.00.mat", "precalc_k4_l4_g1.npz")]


@pytest.mark.parametrize("mat,npz", CASES)
def test_load_precalc_matches_h5dump_extraction(mat, npz):
    t = cfk.load_precalc(os.path.join(GOLDEN, mat))
    d = np.load(os.path.join(GOLDEN, npz))
    assert np.array_equal(t["R"], d["R"])
    assert np.array_equal(t["labels"], d["labels"].astype(np.float64))
    assert np.array_equal(t["vals"], d["vals"])
    assert np.array_equal(t["ic_map"], d["ic_map"].astype(np.float64))
    assert [t["K_max"], t["L_max"], t["D_max"], t["alpha"]] == list(d["meta"])


def test_dataset_dims_are_matlab_order():
    R, dims = cfk.read_mat_dataset(os.path.join(GOLDEN, CASES[1][0]), "/#refs#/t")
    assert dims == (5, 5, 5, 42) and R.size == 5 * 5 * 5 * 42
    lab, dims = cfk.read_mat_dataset(os.path.join(GOLDEN, CASES[1][0]), "/#refs#/u")
    assert dims == (1158, 3)


def test_loaded_geometry_equals_native_real_gaunt():
    """The precalc COO read from the file equals the library's own Gaunt COO for L=4."""
    t = cfk.load_precalc(os.path.join(GOLDEN, CASES[1][0]))
    lab, val, ic = cfk.real_gaunt(4)
    assert np.array_equal(t["labels"], lab) and np.array_equal(t["ic_map"], ic)
    assert np.max(np.abs(t["vals"] - val)) <= 1e-15


def test_errors_are_reported(tmp_path):
    bad = tmp_path / "plain.mat"
    bad.write_bytes(b"MATLAB 5.0 MAT-file" + bytes(600))
    with pytest.raises(cfk.CfkError, match="superblock"):
        cfk.read_mat_dataset(bad, "/x")
    path = os.path.join(GOLDEN, CASES[0][0])
    with pytest.raises(cfk.CfkError, match="no object"):
        cfk.read_mat_dataset(path, "/#refs#/no_such_dataset")
    with pytest.raises(cfk.CfkError, match="float64"):
        cfk.read_mat_dataset(path, "/#refs#/B")        # a uint64 MCOS dataset
    n = ctypes.c_int64(0)
    out = np.zeros(4)
    rc = cfk.lib().cfk_h5_read_f64(path.encode(), b"/#refs#/t", out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   4, ctypes.byref(n), None, None)
    assert rc == -5 and "too small" in cfk.lib().cfk_last_error().decode()
    with pytest.raises(cfk.CfkError):
        cfk.read_mat_dataset(tmp_path / "missing.mat", "/x")
