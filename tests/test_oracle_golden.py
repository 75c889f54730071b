"""Pins the oracle (oracle/pti_oracle.c) before its results are trusted.

Against the reference's own outputs (tests/golden/*.npz, written by
tests/golden/make_goldens.py from the compiled reference): the restated
conversion must reproduce every HiCOO array and the sorted COO bit for bit,
and the restated sequential FP32 MTTKRPs must reproduce the reference's FP32
results bit for bit.  The FP64 closed form is pinned by mathematics:
all-ones factors reduce it to a segmented sum, a rank-1 tensor has a closed
form, it is linear in the values and independent of the nonzero order.
"""
import numpy as np
import pytest

from conftest import golden_cases, hicoo_keys, load_golden, rel_fro

CASES = golden_cases()


@pytest.mark.parametrize("case", CASES)
def test_conversion_matches_reference(case, oracle):
    g = load_golden(case)
    N = len(g["ndims"])
    for p, (sb, sk, sc) in enumerate(g["params"]):
        h, si, sv = oracle.hicoo_convert(g["ndims"], g["in_inds"], g["in_vals"], int(sb), int(sk), int(sc))
        for k in hicoo_keys(N):
            np.testing.assert_array_equal(h[k], g["p%d_%s" % (p, k)], err_msg="%s p%d %s" % (case, p, k))
        assert int(h["max_nnzb"]) == int(g["p%d_max_nnzb" % p])
        np.testing.assert_array_equal(si, g["p%d_sorted_inds" % p])
        np.testing.assert_array_equal(sv, g["p%d_sorted_vals" % p])


def _factor_sets(g):
    N = len(g["ndims"])
    R = int(g["R"])
    sets = {"unif": [g["fac%d" % m] for m in range(N)],
            "const": [np.ones((int(d), R), np.float32) for d in g["ndims"]]}
    if "coo_ipj_m0" in g:
        sets["ipj"] = [(np.arange(int(d))[:, None] + np.arange(R)[None, :] + 1).astype(np.float32)
                       for d in g["ndims"]]
    return sets


MTTKRP_CASES = [c for c in CASES if "coo_unif_m0" in load_golden(c)]


@pytest.mark.parametrize("case", MTTKRP_CASES)
def test_sequential_mttkrp_matches_reference_bitwise(case, oracle):
    g = load_golden(case)
    N = len(g["ndims"])
    for kind, facs in _factor_sets(g).items():
        for mode in range(N):
            out = oracle.mttkrp(g["in_inds"], g["in_vals"], facs, mode, kind="seq")
            np.testing.assert_array_equal(out, g["coo_%s_m%d" % (kind, mode)], err_msg="%s %s m%d" % (case, kind, mode))


@pytest.mark.parametrize("case", MTTKRP_CASES)
def test_hicoo_sequential_mttkrp_matches_reference_bitwise(case, oracle):
    g = load_golden(case)
    N = len(g["ndims"])
    sb = int(g["params"][0][0])
    h = {k: g["p0_" + k] for k in ("binds", "einds", "values", "kptr", "bptr")}
    facs = _factor_sets(g)["unif"]
    for mode in range(N):
        out = oracle.mttkrp_hicoo(h, sb, facs, mode, kind="seq")
        np.testing.assert_array_equal(out, g["hicoo_unif_m%d" % mode])


@pytest.mark.parametrize("case", MTTKRP_CASES)
def test_fp64_and_omp_close_to_reference(case, oracle):
    g = load_golden(case)
    N = len(g["ndims"])
    facs = _factor_sets(g)["unif"]
    for mode in range(N):
        ref = g["coo_unif_m%d" % mode]
        f64 = oracle.mttkrp(g["in_inds"], g["in_vals"], facs, mode, kind="f64")
        assert rel_fro(ref, f64) < 1e-5
        omp = oracle.mttkrp(g["in_inds"], g["in_vals"], facs, mode, kind="omp", nthreads=4)
        assert rel_fro(omp, f64) < 1e-5
        h = {k: g["p0_" + k] for k in ("binds", "einds", "values", "kptr", "bptr")}
        homp = oracle.mttkrp_hicoo(h, int(g["params"][0][0]), facs, mode, kind="omp", nthreads=4)
        assert rel_fro(homp, f64) < 1e-5


def _random_tensor(rng, ndims, nnz):
    inds = np.stack([rng.integers(0, d, nnz) for d in ndims]).astype(np.uint32)
    vals = rng.random(nnz).astype(np.float32)
    return inds, vals


def test_fp64_closed_forms(oracle):
    rng = np.random.default_rng(5)
    ndims = [37, 29, 23, 11]
    inds, vals = _random_tensor(rng, ndims, 3000)
    R = 6
    # All-ones factors: a segmented sum of the values.
    ones = [np.ones((d, R), np.float32) for d in ndims]
    for mode in range(4):
        out = oracle.mttkrp(inds, vals, ones, mode, kind="f64")
        seg = np.bincount(inds[mode], weights=vals.astype(np.float64), minlength=ndims[mode])
        np.testing.assert_allclose(out, np.repeat(seg[:, None], R, axis=1), rtol=1e-13, atol=1e-13)
    # Rank-1 tensor X = v (x) w (x) y (x) u on its nonzeros only:
    # M[i, r] = sum_{z: i_0 = i} v_z prod_k A_k[i_k, r], checked by a direct numpy formula.
    facs = [rng.random((d, R)).astype(np.float32) for d in ndims]
    for mode in range(4):
        out = oracle.mttkrp(inds, vals, facs, mode, kind="f64")
        prod = vals.astype(np.float64)[:, None] * np.ones((1, R))
        for k in range(4):
            if k != mode:
                prod = prod * facs[k][inds[k]].astype(np.float64)
        exp = np.zeros((ndims[mode], R))
        np.add.at(exp, inds[mode], prod)
        np.testing.assert_allclose(out, exp, rtol=1e-12, atol=1e-12)
    # Linearity in the values and independence of the nonzero order.
    a = oracle.mttkrp(inds, vals, facs, 1, kind="f64")
    b = oracle.mttkrp(inds, 2 * vals, facs, 1, kind="f64")
    np.testing.assert_allclose(b, 2 * a, rtol=1e-14)
    perm = rng.permutation(inds.shape[1])
    c = oracle.mttkrp(inds[:, perm], vals[perm], facs, 1, kind="f64")
    np.testing.assert_allclose(c, a, rtol=1e-12)


def test_mats_order_changes_only_rounding(oracle):
    rng = np.random.default_rng(6)
    ndims = [20, 30, 40, 10]
    inds, vals = _random_tensor(rng, ndims, 2000)
    facs = [rng.random((d, 8)).astype(np.float32) for d in ndims]
    a = oracle.mttkrp(inds, vals, facs, 2, order=[2, 3, 0, 1], kind="seq")
    b = oracle.mttkrp(inds, vals, facs, 2, order=[2, 0, 1, 3], kind="seq")
    f = oracle.mttkrp(inds, vals, facs, 2, kind="f64")
    assert rel_fro(a, f) < 1e-6 and rel_fro(b, f) < 1e-6


def test_comparison_detects_planted_mistakes(oracle):
    """The bitwise comparison above is sensitive to the plausible mistakes:
    a dropped factor, an index off by one, a sign flip."""
    g = load_golden("tns_3D_12031")
    facs = [g["fac%d" % m] for m in range(3)]
    ref = g["coo_unif_m0"]
    dropped = oracle.mttkrp(g["in_inds"], g["in_vals"], [facs[0], facs[1], np.ones_like(facs[2])], 0, kind="seq")
    assert not np.array_equal(dropped, ref)
    shifted = g["in_inds"].copy()
    shifted[1] = (shifted[1] + 1) % g["ndims"][1]
    assert not np.array_equal(oracle.mttkrp(shifted, g["in_vals"], facs, 0, kind="seq"), ref)
    flipped = g["in_vals"].copy()
    flipped[7] = -flipped[7]
    assert not np.array_equal(oracle.mttkrp(g["in_inds"], flipped, facs, 0, kind="seq"), ref)


def test_lexicographic_sort(oracle):
    g = load_golden("syn3_dup")
    si, sv = oracle.sort_lexicographic(g["in_inds"], g["in_vals"])
    keys = si[0].astype(np.int64) * 10**8 + si[1].astype(np.int64) * 10**4 + si[2]
    assert np.all(np.diff(keys) >= 0)
    assert sorted(zip(map(tuple, g["in_inds"].T), g["in_vals"])) == sorted(zip(map(tuple, si.T), sv))
