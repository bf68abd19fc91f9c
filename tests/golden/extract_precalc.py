"""Extract golden fixtures from the reference's precomputed tensor files.

The reference ships nine MATLAB v7.3 files ``src/precalc/collisiontensor_k4_l{L}_gamma{g}.mat``
(written by ``precompute_collision_operator.m:61``).  A v7.3 file is HDF5; the ``TensorObj``
properties live under ``/#refs#`` (MCOS object storage):

    /#refs#/t  R_tensor      HDF5 dims (N_L,5,5,5)  == MATLAB R[k1,k2,k3,t] column-major
    /#refs#/u  gaunt_labels  HDF5 dims (3,N_G)      == MATLAB [N_G x 3] column-major, 1-based
    /#refs#/v  gaunt_vals    HDF5 dims (1,N_G)
    /#refs#/w  ic_map        HDF5 dims (1,N_G)      1-based l-triplet rank
    /#refs#/q,r,s  K_max, L_max, D_max;  /#refs#/j  alpha (gamma)

Only ``h5dump`` (HDF5 command-line tool) is used, as a byte reader: nothing in the file is
executed.  Output: one ``precalc_k4_l{L}_g{gamma}.npz`` per file with

    labels  int16 [N_G,3]  (1-based, MATLAB find order: sorted by (q3,q2,q1))
    vals    float64 [N_G]
    ic_map  int16 [N_G]    (1-based)
    R       float64, flat in MATLAB column-major order, length 125*N_L
    meta    float64 [K_max, L_max, D_max, alpha]

Two of the files (L=2 gamma=0: R in one chunk, the Gaunt COO compact; L=4 gamma=1: R and the
labels in deflate-compressed chunks) are also copied byte for byte next to the npz files, as data for the
native .mat reader's test (tests/test_matfile.py).

Run this in a container where the reference is mounted at /root/reference; the fixtures
are committed so nothing at test time reads the reference.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference/src/precalc"
H5DUMP = "/opt/conda/bin/h5dump"
HERE = os.path.dirname(os.path.abspath(__file__))

FILES = [
    (2, 0.0), (4, 0.0), (6, 0.0),
    (2, 1.0), (4, 1.0), (6, 1.0), (8, 1.0), (10, 1.0), (12, 1.0),
]


def dump(path, name, tmpdir):
    out = os.path.join(tmpdir, name.strip("/").replace("/", "_").replace("#", "") + ".bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path],
                   check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return np.fromfile(out, dtype="<f8")


def extract(L, gamma, tmpdir):
    fn = "collisiontensor_k4_l%d_gamma%.2f.mat" % (L, gamma)
    path = os.path.join(REF, fn)
    R = dump(path, "/#refs#/t", tmpdir)
    lab = dump(path, "/#refs#/u", tmpdir)
    vals = dump(path, "/#refs#/v", tmpdir)
    ic = dump(path, "/#refs#/w", tmpdir)
    meta = np.array([dump(path, "/#refs#/" + c, tmpdir)[0] for c in "qrsj"])
    n_g = vals.size
    labels = lab.reshape(3, n_g).T  # column-major [N_G x 3]
    assert np.all(labels == np.round(labels)) and np.all(ic == np.round(ic))
    assert R.size % 125 == 0 and int(ic.max()) == R.size // 125
    assert int(meta[1]) == L and abs(meta[3] - gamma) < 1e-12
    out = os.path.join(HERE, "precalc_k4_l%d_g%d.npz" % (L, int(gamma)))
    np.savez_compressed(out, labels=labels.astype(np.int16), vals=vals,
                        ic_map=ic.astype(np.int16), R=R, meta=meta)
    print("%s: N_G=%d N_L=%d -> %s (%d bytes)" % (fn, n_g, R.size // 125, os.path.basename(out),
                                                   os.path.getsize(out)))


MAT_COPIES = ["collisiontensor_k4_l2_gamma0.00.mat", "collisiontensor_k4_l4_gamma1.00.mat"]


def main():
    with tempfile.TemporaryDirectory() as tmpdir:
        for L, g in FILES:
            extract(L, g, tmpdir)
    for fn in MAT_COPIES:
        shutil.copyfile(os.path.join(REF, fn), os.path.join(HERE, fn))


if __name__ == "__main__":
    sys.exit(main())
