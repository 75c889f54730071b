"""libHiParTI's host-side COO -> HiCOO conversion and sorting (CPU only).

The conversion is product code (hiparti_amd/csrc/host/hicoo.cpp) that must
reproduce the reference's arrays bit for bit: checked against the reference's
own fixtures, against the oracle on seeded random inputs (orders 2-6, ties,
duplicates, shuffled input order, several block / superblock / chunk sizes),
and against SHA-256 digests of the reference's output at 0.3M-2M nonzeros.
"""
import hashlib

import numpy as np
import pytest

from conftest import golden_cases, golden_digests, hicoo_keys, load_golden
from hiparti_amd import COOTensor, PtiError


def convert(ndims, inds, vals, sb, sk, sc, nthreads=0):
    t = COOTensor.from_arrays(ndims, inds, vals)
    h = t.toHiCOO(sb, sk, sc, nthreads=nthreads)
    a = h.arrays()
    a["max_nnzb"] = h.max_nnzb
    return a, t.inds(), t.values()


@pytest.mark.parametrize("case", golden_cases())
def test_product_conversion_matches_reference(case):
    g = load_golden(case)
    N = len(g["ndims"])
    for p, (sb, sk, sc) in enumerate(g["params"]):
        a, si, sv = convert(g["ndims"], g["in_inds"], g["in_vals"], int(sb), int(sk), int(sc))
        for k in hicoo_keys(N):
            np.testing.assert_array_equal(a[k], g["p%d_%s" % (p, k)], err_msg="%s p%d %s" % (case, p, k))
        assert a["max_nnzb"] == int(g["p%d_max_nnzb" % p])
        np.testing.assert_array_equal(si, g["p%d_sorted_inds" % p])
        np.testing.assert_array_equal(sv, g["p%d_sorted_vals" % p])


def _random_case(rng, N, with_dups):
    ndims = [int(x) for x in rng.integers(3, 60, N)]
    cells = int(np.prod(ndims, dtype=np.float64))
    nnz = int(min(rng.integers(1, 3000), cells))
    inds = np.stack([rng.integers(0, d, nnz) for d in ndims]).astype(np.uint32)
    if with_dups and nnz > 2:
        k = int(rng.integers(1, nnz))
        extra = rng.integers(0, nnz, k)
        inds = np.concatenate([inds, inds[:, extra]], axis=1)
    vals = rng.random(inds.shape[1]).astype(np.float32)
    return ndims, inds, vals


@pytest.mark.parametrize("seed", range(24))
def test_product_conversion_matches_oracle_random(seed, oracle):
    rng = np.random.default_rng(1000 + seed)
    N = int(rng.integers(2, 7))
    ndims, inds, vals = _random_case(rng, N, with_dups=bool(seed % 3 == 0))
    sb = int(rng.integers(1, 4))
    sk = sb + int(rng.integers(0, 3))
    sc = sb + int(rng.integers(0, 4))
    h, si, sv = oracle.hicoo_convert(ndims, inds, vals, sb, sk, sc)
    for nthreads in (1, 4):
        a, pi, pv = convert(ndims, inds, vals, sb, sk, sc, nthreads=nthreads)
        for k in hicoo_keys(N):
            np.testing.assert_array_equal(a[k], h[k], err_msg="seed %d %s" % (seed, k))
        np.testing.assert_array_equal(pi, si)
        np.testing.assert_array_equal(pv, sv)


def _digest(arrs):
    d = hashlib.sha256()
    for a in arrs:
        d.update(np.ascontiguousarray(a).tobytes())
    return d.hexdigest()


@pytest.mark.parametrize("name,ndims,nnz,dist,s,permute", [
    ("c0", [1000, 1000, 1000], 1000000, [0, 0, 0], None, [0, 0, 0]),
    ("c1s", [1 << 20, 1 << 20, 1 << 16], 2000000, [1, 1, 1], [1.1] * 3, [1, 1, 1]),
    ("c4s", [1000000, 500000, 100000, 1000, 24], 300000, [1, 1, 1, 0, 0], [1.1, 1.1, 1.1, 1, 1], [0] * 5),
])
def test_product_conversion_matches_reference_digests(name, ndims, nnz, dist, s, permute):
    dg = golden_digests()
    t = COOTensor.synthetic(ndims, nnz, dist=dist, zipf_s=s, permute=permute, seed=42)
    assert _digest([t.inds(), t.values()]) == dg[name + "_input"], "generator output changed"
    h = t.toHiCOO(7, 14, 14)
    a = h.arrays()
    assert str(len(a["bptr"]) - 1) == dg[name + "_nb"]
    assert _digest([a["bptr"], a["binds"], a["einds"], a["values"]]) == dg[name + "_blocks"]
    assert _digest([a["kptr"], a["cptr"], a["knnzs"], a["nkiters"]] +
                   [a["kschr_dat%d" % m] for m in range(len(ndims))]) == dg[name + "_sched"]


def test_conversion_argument_errors():
    t = COOTensor.from_arrays([4, 4, 4], np.zeros((3, 1), np.uint32), np.ones(1, np.float32))
    with pytest.raises(PtiError) as e:
        t.toHiCOO(3, 2, 3)      # sk_bits < sb_bits
    assert e.value.code == 3
    with pytest.raises(PtiError):
        t.toHiCOO(8, 8, 8)      # element offsets are 8-bit: sb_bits <= 7


def test_empty_tensor_conversion(oracle):
    ndims = [5, 6, 7]
    inds = np.zeros((3, 0), np.uint32)
    vals = np.zeros(0, np.float32)
    a, _, _ = convert(ndims, inds, vals, 1, 2, 2)
    h, _, _ = oracle.hicoo_convert(ndims, inds, vals, 1, 2, 2)
    for k in hicoo_keys(3):
        np.testing.assert_array_equal(a[k], h[k], err_msg=k)


def test_lexicographic_sort_matches_reference_quicksort(oracle):
    g = load_golden("syn3_dup")
    t = COOTensor.from_arrays(g["ndims"], g["in_inds"], g["in_vals"])
    t.sort()
    si, sv = oracle.sort_lexicographic(g["in_inds"], g["in_vals"])
    np.testing.assert_array_equal(t.inds(), si)
    np.testing.assert_array_equal(t.values(), sv)


def test_load_tns(tmp_path):
    g = load_golden("tns_3D_12031")
    p = tmp_path / "x.tns"
    with open(p, "w") as f:
        f.write("3\n%s\n" % " ".join(str(int(d)) for d in g["ndims"]))
        for z in range(g["in_inds"].shape[1]):
            f.write("%d %d %d %.9e\n" % (g["in_inds"][0, z] + 1, g["in_inds"][1, z] + 1, g["in_inds"][2, z] + 1,
                                         g["in_vals"][z]))
    t = COOTensor().load(str(p))
    np.testing.assert_array_equal(t.inds(), g["in_inds"])
    np.testing.assert_array_equal(t.values(), g["in_vals"])
    with pytest.raises(PtiError) as e:
        COOTensor().load(str(tmp_path / "missing.tns"))
    assert e.value.code == 0x10000 + 2   # PTIERR_OS_ERROR + ENOENT
