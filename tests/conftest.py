import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); runs on the GPU box")


def golden_cases():
    """The conversion / MTTKRP fixtures (dense_helpers.npz holds the CP-ALS
    helpers and dumps instead)."""
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not p.endswith("dense_helpers.npz"))


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_digests():
    out = {}
    with open(os.path.join(GOLDEN, "digests.txt")) as f:
        for line in f:
            k, v = line.split()
            out[k] = v
    return out


def rel_fro(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    if nb == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)


HICOO_KEYS = ["bptr", "binds", "einds", "values", "kptr", "cptr", "knnzs", "nkiters", "nkpars"]


def hicoo_keys(N):
    keys = list(HICOO_KEYS)
    for m in range(N):
        keys += ["kschr_ptr%d" % m, "kschr_dat%d" % m, "bal_ptr%d" % m, "bal_dat%d" % m,
                 "pos_ptr%d" % m, "pos_dat%d" % m, "rest%d" % m]
    return keys


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu_available():
    from hiparti_amd import lib
    return lib().ptiDevCount() > 0
