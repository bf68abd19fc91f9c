import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "collision-factorization_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libcfk.so")


def load_precalc(L, gamma):
    """Reference fixture: (labels int [N_G,3] 1-based, vals, ic_map 1-based, R flat col-major, meta)."""
    d = np.load(os.path.join(GOLDEN, "precalc_k4_l%d_g%d.npz" % (L, int(gamma))))
    return (d["labels"].astype(np.int64), d["vals"], d["ic_map"].astype(np.int64), d["R"], d["meta"])


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible HIP device")
    return torch
