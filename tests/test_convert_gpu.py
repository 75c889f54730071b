"""GPU COO -> HiCOO conversion (ptiDevSparseTensorToHiCOO) against the host
conversion and the reference.

Every array of the HiCOO struct, max_nnzb and the caller's re-sorted COO must
be bit-identical to ptiSparseTensorToHiCOO's (itself bit-exact against the
reference in test_convert_host.py): on the reference fixtures, on
benchmark-shaped 3-D Zipf and 4-D tensors (GPU path), against the reference's
SHA-256 digests at 2M nonzeros, and on the cases the GPU path hands to the
host code (duplicates, 3-D indices >= 2^24, orders 2 and 5).
"""
import hashlib

import numpy as np
import pytest

from conftest import golden_cases, golden_digests, hicoo_keys, load_golden
from hiparti_amd import COOTensor

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu(gpu_available):
    if not gpu_available:
        pytest.fail("no GPU visible: the GPU tests must run on the GPU box")


def both(ndims, inds, vals, sb, sk, sc):
    t1 = COOTensor.from_arrays(ndims, inds, vals)
    t2 = COOTensor.from_arrays(ndims, inds, vals)
    h1 = t1.toHiCOO(sb, sk, sc)
    h2 = t2.toHiCOO(sb, sk, sc, device=True)
    return (h1, t1), (h2, t2)


def assert_same(a, b, N, what):
    (h1, t1), (h2, t2) = a, b
    x, y = h1.arrays(), h2.arrays()
    for k in hicoo_keys(N):
        np.testing.assert_array_equal(y[k], x[k], err_msg="%s %s" % (what, k))
    assert h1.max_nnzb == h2.max_nnzb, what
    np.testing.assert_array_equal(t2.inds(), t1.inds(), err_msg=what)
    np.testing.assert_array_equal(t2.values(), t1.values(), err_msg=what)


@pytest.mark.parametrize("case", golden_cases())
def test_gpu_conversion_matches_reference_fixtures(case):
    g = load_golden(case)
    N = len(g["ndims"])
    for p, (sb, sk, sc) in enumerate(g["params"]):
        t = COOTensor.from_arrays(g["ndims"], g["in_inds"], g["in_vals"])
        h = t.toHiCOO(int(sb), int(sk), int(sc), device=True)
        a = h.arrays()
        for k in hicoo_keys(N):
            np.testing.assert_array_equal(a[k], g["p%d_%s" % (p, k)], err_msg="%s p%d %s" % (case, p, k))
        assert h.max_nnzb == int(g["p%d_max_nnzb" % p])
        np.testing.assert_array_equal(t.inds(), g["p%d_sorted_inds" % p])
        np.testing.assert_array_equal(t.values(), g["p%d_sorted_vals" % p])


@pytest.mark.parametrize("name,ndims,nnz,dist,s,permute,params", [
    ("zipf3", [1 << 20, 1 << 20, 1 << 16], 3_000_000, [1, 1, 1], [1.1] * 3, [1, 1, 1], (7, 14, 14)),
    ("zipf3_small_blocks", [1 << 20, 1 << 20, 1 << 16], 1_000_000, [1, 1, 1], [1.1] * 3, [1, 1, 1], (3, 9, 5)),
    ("four_d", [10_000_000, 50_000, 2_000, 100], 2_000_000, [1, 1, 0, 0], [1.2, 1.2, 1, 1], [0] * 4, (7, 14, 14)),
    ("uniform3", [1000, 1000, 1000], 1_000_000, [0, 0, 0], None, [0, 0, 0], (7, 10, 12)),
])
def test_gpu_conversion_matches_host(name, ndims, nnz, dist, s, permute, params):
    t = COOTensor.synthetic(ndims, nnz, dist=dist, zipf_s=s, permute=permute, seed=5)
    a, b = both(ndims, t.inds(), t.values(), *params)
    assert_same(a, b, len(ndims), name)


def _digest(arrs):
    d = hashlib.sha256()
    for a in arrs:
        d.update(np.ascontiguousarray(a).tobytes())
    return d.hexdigest()


@pytest.mark.parametrize("name,ndims,nnz,dist,s,permute", [
    ("c0", [1000, 1000, 1000], 1000000, [0, 0, 0], None, [0, 0, 0]),
    ("c1s", [1 << 20, 1 << 20, 1 << 16], 2000000, [1, 1, 1], [1.1] * 3, [1, 1, 1]),
])
def test_gpu_conversion_matches_reference_digests(name, ndims, nnz, dist, s, permute):
    dg = golden_digests()
    t = COOTensor.synthetic(ndims, nnz, dist=dist, zipf_s=s, permute=permute, seed=42)
    h = t.toHiCOO(7, 14, 14, device=True)
    a = h.arrays()
    assert str(len(a["bptr"]) - 1) == dg[name + "_nb"]
    assert _digest([a["bptr"], a["binds"], a["einds"], a["values"]]) == dg[name + "_blocks"]
    assert _digest([a["kptr"], a["cptr"], a["knnzs"], a["nkiters"]] +
                   [a["kschr_dat%d" % m] for m in range(len(ndims))]) == dg[name + "_sched"]


def test_cases_handed_to_the_host_code():
    rng = np.random.default_rng(8)
    # Duplicate coordinates: the order of equal keys is the reference quicksort's.
    inds = rng.integers(0, 30, (3, 4000)).astype(np.uint32)
    inds = np.concatenate([inds, inds[:, :500]], axis=1)
    vals = rng.random(inds.shape[1]).astype(np.float32)
    assert_same(*both([30, 30, 30], inds, vals, 2, 3, 4), 3, "duplicates")
    # 3-D indices >= 2^24 (the reference's Morton key drops their top bits).
    nd = [1 << 25, 1 << 25, 64]
    inds = np.stack([rng.integers(0, d, 3000) for d in nd]).astype(np.uint32)
    vals = rng.random(3000).astype(np.float32)
    assert_same(*both(nd, inds, vals, 7, 14, 14), 3, "wide indices")
    # Orders 2 and 5 (no Morton order in the reference).
    for nd in ([500, 400], [30, 20, 10, 8, 6]):
        inds = np.stack([rng.integers(0, d, 3000) for d in nd]).astype(np.uint32)
        vals = rng.random(3000).astype(np.float32)
        assert_same(*both(nd, inds, vals, 2, 4, 4), len(nd), "order %d" % len(nd))
