"""The seeded synthetic-tensor generator (host code, CPU only)."""
import numpy as np

from hiparti_amd import COOTensor, DenseMatrix


def test_deterministic_unique_sorted_in_range():
    ndims = [300, 200, 50]
    a = COOTensor.synthetic(ndims, 20000, dist=[1, 1, 0], zipf_s=[1.1, 1.2, 1], permute=[1, 0, 0], seed=9,
                            nthreads=1)
    b = COOTensor.synthetic(ndims, 20000, dist=[1, 1, 0], zipf_s=[1.1, 1.2, 1], permute=[1, 0, 0], seed=9,
                            nthreads=7)
    ia, ib = a.inds(), b.inds()
    np.testing.assert_array_equal(ia, ib)
    np.testing.assert_array_equal(a.values(), b.values())
    assert ia.shape == (3, 20000)
    for m, d in enumerate(ndims):
        assert ia[m].max() < d
    key = (ia[0].astype(np.int64) * 200 + ia[1]) * 50 + ia[2]
    assert np.all(np.diff(key) > 0), "lexicographically sorted without duplicates"
    v = a.values()
    assert v.min() > 0 and v.max() < 1
    c = COOTensor.synthetic(ndims, 20000, dist=[1, 1, 0], zipf_s=[1.1, 1.2, 1], permute=[1, 0, 0], seed=10)
    assert not np.array_equal(c.inds(), ia)


def test_zipf_marginal_head():
    # Unpermuted Zipf(1.2) mode: index 0 is the most frequent, frequencies fall.
    t = COOTensor.synthetic([1000, 100000], 50000, dist=[1, 0], zipf_s=[1.2, 1.0], seed=3)
    cnt = np.bincount(t.inds()[0], minlength=1000)
    assert cnt[0] == cnt.max()
    assert cnt[:10].sum() > cnt[10:100].sum() * 0.5
    assert cnt[0] > cnt[9] > cnt[99]


def test_lognormal_values_and_uniform_matrix():
    t = COOTensor.synthetic([50, 60, 70], 5000, lognormal=True, seed=4)
    lv = np.log(t.values().astype(np.float64))
    assert abs(lv.mean()) < 0.1 and abs(lv.std() - 1) < 0.1
    m = DenseMatrix(100, 12)
    m.uniform(5)
    a = m.numpy()
    assert a.min() > 0 and a.max() < 1 and abs(a.mean() - 0.5) < 0.05
