"""CP-ALS helpers and text dumps against the reference's own outputs (CPU).

tests/golden/dense_helpers.npz holds what the reference build wrote
(tests/golden/make_goldens.py): ptiMatrix2Norm and ptiMatrixMaxNorm
(src/matrix/matrix.c:389-553), GetFinalLambda (:556-572) and the text of
ptiDumpMatrix (src/matrix/dump.c:33-49), ptiDumpKruskalTensor
(src/ktensor/dump.c:23-51) and ptiDumpSparseTensorHiCOO
(src/sptensor/hicoo/dump.c:30-74).  Checked against them:
  - the oracle's column statistics (oracle/cpd_als.py col_2norm,
    col_maxnorm), which the device dense step is tested against;
  - the library's host ptiMatrix2Norm / ptiMatrixMaxNorm / GetFinalLambda;
  - the library's dumps, byte for byte.
The matrix has an all-negative column and a column below 1, so a max-norm
taken over absolute values, or one without the clamp to 1, fails here.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, rel_fro
from hiparti_amd import COOTensor, DenseMatrix, lib
from hiparti_amd import _ffi as F
from hiparti_amd.tensor import _CFile

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import cpd_als as A  # noqa: E402


@pytest.fixture(scope="module")
def g():
    with np.load(os.path.join(GOLDEN, "dense_helpers.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_oracle_column_statistics_match_the_reference(g):
    a = g["A"].astype(np.float64)
    an, lam = A.col_2norm(a)
    assert rel_fro(lam, g["2norm_lambda"]) <= 1e-6
    assert rel_fro(an, g["2norm_A"]) <= 1e-6
    am, lam = A.col_maxnorm(a)
    assert rel_fro(lam, g["maxnorm_lambda"]) <= 1e-7
    assert rel_fro(am, g["maxnorm_A"]) <= 1e-6
    # The fixture tells the rule apart from its plausible misreadings.
    assert not np.allclose(np.maximum(np.abs(a).max(axis=0), 1.0), g["maxnorm_lambda"])
    assert not np.allclose(a.max(axis=0), g["maxnorm_lambda"])
    assert g["maxnorm_lambda"][3] == 1.0 and g["maxnorm_lambda"][5] == 1.0


@pytest.mark.parametrize("kind", ["2norm", "maxnorm"])
def test_library_host_norms_match_the_reference(g, kind):
    m = DenseMatrix.from_numpy(g["A"])
    lam = (F.ptiValue * 10)()
    fn = lib().ptiMatrix2Norm if kind == "2norm" else lib().ptiMatrixMaxNorm
    assert fn(C.byref(m.address), lam) == 0
    assert rel_fro(np.array(lam[:10], np.float64), g[kind + "_lambda"]) <= 1e-6
    assert rel_fro(m.numpy(), g[kind + "_A"]) <= 1e-6


def test_library_get_final_lambda_matches_the_reference(g):
    mats = [DenseMatrix.from_numpy(g["final_in%d" % i]) for i in range(3)]
    arr = (C.POINTER(F.ptiMatrix) * 3)(*[C.pointer(m.address) for m in mats])
    lam = (F.ptiValue * 6)(*g["final_lambda_in"])
    lib().GetFinalLambda(6, 3, arr, lam)
    assert rel_fro(np.array(lam[:6], np.float64), g["final_lambda_out"]) <= 1e-6
    for i, m in enumerate(mats):
        assert rel_fro(m.numpy(), g["final_out%d" % i]) <= 1e-6


def _dump(fn, obj, tmp_path):
    path = str(tmp_path / "dump.txt")
    with _CFile(path, "w") as fp:
        assert fn(C.byref(obj), fp) == 0
    with open(path, "rb") as f:
        return f.read()


def test_dump_matrix_is_byte_identical(g, tmp_path):
    m = DenseMatrix.from_numpy(g["A"])
    assert _dump(lib().ptiDumpMatrix, m.address, tmp_path) == g["dump_matrix"].tobytes()


def test_dump_kruskal_tensor_is_byte_identical(g, tmp_path):
    mats = [DenseMatrix.from_numpy(g["final_out%d" % i]) for i in range(3)]
    k = F.ptiKruskalTensor()
    k.nmodes = 3
    k.rank = 6
    nd = (F.ptiIndex * 3)(20, 15, 9)
    k.ndims = nd
    lam = (F.ptiValue * 6)(*g["final_lambda_out"])
    k.lambda_ = lam
    k.fit = 0.123456789
    arr = (C.POINTER(F.ptiMatrix) * 3)(*[C.pointer(m.address) for m in mats])
    k.factors = arr
    assert _dump(lib().ptiDumpKruskalTensor, k, tmp_path) == g["dump_ktensor"].tobytes()


def test_dump_hicoo_is_byte_identical(g, tmp_path):
    t = COOTensor.from_arrays([int(x) for x in g["hicoo_ndims"]], g["hicoo_in_inds"], g["hicoo_in_vals"])
    h = t.toHiCOO(*[int(x) for x in g["hicoo_params"]])
    path = str(tmp_path / "hicoo.txt")
    h.dump(path)
    with open(path, "rb") as f:
        assert f.read() == g["dump_hicoo"].tobytes()
