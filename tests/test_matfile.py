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


H5DUMP = "/opt/conda/bin/h5dump"


def _h5dump_values(path, name, tmp_path):
    """The dataset's raw float64 values as h5dump writes them (little endian binary)."""
    import subprocess
    out = tmp_path / ("%s.bin" % name)
    subprocess.run([H5DUMP, "-d", "/" + name, "-b", "LE", "-o", str(out), str(path)], check=True,
                   capture_output=True)
    return np.fromfile(out, dtype="<f8")


def test_writer_round_trip(tmp_path):
    """cfk_mat_write_f64 (the save(..., '-v7.3') of precompute_collision_operator.m:61 for plain arrays):
    every variable reads back bit for bit, in MATLAB dimensions, through the native reader."""
    rng = np.random.default_rng(3)
    vars_ = {"A": rng.standard_normal((3, 4)), "R_tensor": rng.standard_normal((5, 5, 5, 42)),
             "v": rng.standard_normal(7), "K_max": 4.0, "neg_zero": np.array([[-0.0, np.inf, -np.inf]])}
    p = tmp_path / "w.mat"
    cfk.save_mat(p, vars_)
    head = open(p, "rb").read(128)
    assert head.startswith(b"MATLAB 7.3 MAT-file") and head[124:128] == b"\x00\x02IM"
    for name, a in vars_.items():
        a = np.asarray(a, dtype=np.float64)
        shape = a.shape if a.ndim >= 2 else ((a.size, 1) if a.ndim == 1 else (1, 1))
        d, dims = cfk.read_mat_dataset(p, name)
        assert dims == shape, name
        assert np.array_equal(d.view(np.uint64), np.asfortranarray(a.reshape(shape)).ravel(order="F").view(np.uint64))


def test_writer_output_is_read_by_hdf5(tmp_path):
    """The HDF5 library reads the file: datasets with MATLAB's MATLAB_class attribute, dimensions
    reversed from MATLAB's, the same values; 20 variables span three symbol-table nodes."""
    import subprocess
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    rng = np.random.default_rng(4)
    vars_ = {"x%02d" % i: rng.standard_normal((i + 1, 2)) for i in range(20)}
    p = tmp_path / "many.mat"
    cfk.save_mat(p, vars_)
    txt = subprocess.run([H5DUMP, "-A", str(p)], check=True, capture_output=True, text=True).stdout
    for name, a in vars_.items():
        assert 'DATASET "%s"' % name in txt
        assert np.array_equal(_h5dump_values(p, name, tmp_path), a.ravel(order="F"))
    assert txt.count('"double"') == 20 and "SIMPLE { ( 2, 20 ) / ( 2, 20 ) }" in txt


def test_operator_round_trip_of_a_reference_fixture(tmp_path):
    """The reference's L=4, gamma=1 precalc operator saved as plain variables and read back."""
    src = os.path.join(GOLDEN, "collisiontensor_k4_l4_gamma1.00.mat")
    if not os.path.exists(src):
        pytest.skip("fixture .mat not present")
    op = cfk.load_precalc(src)
    p = tmp_path / "op.mat"
    cfk.save_operator(p, op["R"], op["labels"], op["vals"], op["ic_map"], op["K_max"], op["L_max"], op["alpha"])
    back = cfk.load_operator(p)
    assert np.array_equal(back["R"], op["R"]) and np.array_equal(back["labels"], op["labels"])
    assert np.array_equal(back["vals"], op["vals"]) and np.array_equal(back["ic_map"], op["ic_map"])
    assert (back["K_max"], back["L_max"], back["alpha"]) == (op["K_max"], op["L_max"], op["alpha"])
    assert cfk.read_mat_dataset(p, "R_tensor")[1] == (5, 5, 5, int(op["ic_map"].max()))


def test_writer_rejects_bad_variables(tmp_path):
    """Names that are not MATLAB identifiers, and a name given twice, are rejected before writing."""
    from cfk._lib import _i32, _i64, check, lib
    p = tmp_path / "bad.mat"
    for bad in ({"1x": 1.0}, {"a b": 1.0}, {"": 1.0}):
        with pytest.raises(cfk.CfkError):
            cfk.save_mat(p, bad)
    a = np.zeros(1)
    ap = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    with pytest.raises(cfk.CfkError):
        check(lib().cfk_mat_write_f64(str(p).encode(), 2, (ctypes.c_char_p * 2)(b"a", b"a"), (_i32 * 2)(1, 1),
                                      (_i64 * 2)(1, 1), (ctypes.POINTER(ctypes.c_double) * 2)(ap, ap)))
