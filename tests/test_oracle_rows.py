"""The oracle's full-size checkers against its pinned FP64 closed form (CPU).

orc_mttkrp_f64_rows (the FP64 MTTKRP restricted to chosen output rows, used
to check results on tensors too large for the full closed form) must equal
orc_mttkrp_f64 -- itself pinned by closed forms in test_oracle_golden.py --
on those rows, for orders 2-5, repeated and unused rows, strides above the
rank and any thread count.  orc_completion_f64 must give numpy's FP64 loss
and, on chosen rows, numpy's FP64 MTTKRP of the FP64 residuals.
"""
import numpy as np
import pytest

import oracle as O


def _case(ndims, nnz, R, seed, stride=None):
    rng = np.random.default_rng(seed)
    inds = np.stack([rng.integers(0, d, nnz) for d in ndims]).astype(np.uint32)
    vals = rng.uniform(-1, 1, nnz).astype(np.float32)
    stride = stride or R
    facs = []
    for d in ndims:
        f = np.full((d, stride), np.nan, np.float32)
        f[:, :R] = rng.uniform(-1, 1, (d, R))
        facs.append(f)
    return inds, vals, facs


@pytest.mark.parametrize("ndims,nnz,R,stride", [([60, 50], 3000, 5, 8), ([70, 60, 50], 8000, 16, 16),
                                                ([20, 15, 12, 9], 5000, 7, 8), ([9, 8, 7, 6, 5], 4000, 32, 40)])
@pytest.mark.parametrize("nthreads", [1, 4])
def test_mttkrp_rows_equals_the_closed_form(ndims, nnz, R, stride, nthreads):
    inds, vals, facs = _case(ndims, nnz, R, seed=nnz + R, stride=stride)
    rng = np.random.default_rng(1)
    for mode in range(len(ndims)):
        full = O.mttkrp(inds, vals, [f[:, :R] for f in facs], mode, kind="f64")
        rows = np.concatenate([rng.integers(0, ndims[mode], 6), [0, ndims[mode] - 1, 2, 2]]).astype(np.uint32)
        got = O.mttkrp_rows(inds, vals, [f[:, :R].copy() for f in facs], mode, rows, nthreads=nthreads)
        assert np.allclose(got, full[rows], rtol=1e-13, atol=1e-13 * np.abs(full).max())
        # Factors read with their own (padded) stride give the same rows.
        got_s = O.mttkrp_rows(inds, vals, facs, mode, rows, nthreads=nthreads)
        assert np.array_equal(got_s.shape, (len(rows), stride))
        assert np.allclose(got_s[:, :R], full[rows], rtol=1e-13, atol=1e-13 * np.abs(full).max())


def test_mttkrp_rows_empty_selection_and_empty_tensor():
    inds, vals, facs = _case([10, 9, 8], 100, 4, seed=3)
    assert O.mttkrp_rows(inds, vals, facs, 0, []).shape == (0, 4)
    out = O.mttkrp_rows(inds[:, :0], vals[:0], facs, 1, [0, 3])
    assert out.shape == (2, 4) and np.all(out == 0)


@pytest.mark.parametrize("ndims,nnz,R", [([50, 40, 30], 6000, 8), ([12, 11, 10, 9], 3000, 5)])
def test_completion_f64_matches_numpy(ndims, nnz, R):
    inds, vals, facs = _case(ndims, nnz, R, seed=R)
    f64 = [f.astype(np.float64) for f in facs]
    prod = np.ones((nnz, R))
    for m, f in enumerate(f64):
        prod *= f[inds[m]]
    e = vals.astype(np.float64) - prod.sum(axis=1)
    for mode in range(len(ndims)):
        rows = np.array([0, 5, ndims[mode] - 1], np.uint32)
        loss, got = O.completion_f64(inds, vals, facs, mode, rows, nthreads=3)
        assert abs(loss - 0.5 * float(e @ e)) <= 1e-12 * loss
        p = e[:, None] * np.ones((1, R))
        for m, f in enumerate(f64):
            if m != mode:
                p = p * f[inds[m]]
        full = np.zeros((ndims[mode], R))
        np.add.at(full, inds[mode], p)
        assert np.allclose(got, full[rows], rtol=1e-12, atol=1e-12 * np.abs(full).max())
