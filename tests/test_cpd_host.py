"""CP-ALS oracle pins and the host-side CP-ALS helpers (CPU only).

The FP64 ALS restatement in oracle/cpd_als.py is pinned by mathematics, not
by retyping its formulas:
  - its MTTKRP equals the C oracle's (an independent restatement of
    src/sptensor/mttkrp.c:58-91);
  - the normal-equation solve satisfies A V = M (Cholesky and LU paths) and
    leaves M unchanged for an exactly singular V;
  - the fit formula equals 1 - ||X - K||_F / ||X||_F computed on the dense
    reconstruction of a small tensor (ktensor.c:77-99 closed form);
  - ALS recovers an exactly rank-R dense tensor (fit -> 1), and its fit
    never decreases after the first iteration on such a tensor.
The host helpers of the library (ptiMatrixDotMulSeqTriangle, KruskalTensorFit,
GetFinalLambda, ptiRankMatrix, ptiRankKruskalTensor) are checked against
those pinned functions.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, rel_fro
from hiparti_amd import COOTensor, DenseMatrix, lib
from hiparti_amd import _ffi as F

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import cpd_als as A  # noqa: E402


def dense_tensor(ndims, R, seed, noise=0.0):
    rng = np.random.default_rng(seed)
    facs = [rng.uniform(0.1, 1.0, (d, R)) for d in ndims]
    X = np.einsum("ir,jr,kr->ijk", *facs) if len(ndims) == 3 else None
    if noise:
        X = X + noise * rng.standard_normal(X.shape)
    inds = np.array(np.unravel_index(np.arange(X.size), X.shape), np.uint32)
    vals = X.ravel().astype(np.float32)
    return X, facs, inds, vals


def reconstruct(lam, facs):
    return np.einsum("r,ir,jr,kr->ijk", lam, *facs)


def test_oracle_mttkrp_matches_c_oracle(oracle):
    rng = np.random.default_rng(0)
    ndims = [30, 20, 10]
    inds = np.array([rng.integers(0, d, 500) for d in ndims], np.uint32)
    vals = rng.uniform(-1, 1, 500).astype(np.float32)
    facs = [rng.uniform(-1, 1, (d, 6)).astype(np.float32) for d in ndims]
    for m in range(3):
        a = A.mttkrp_f64(inds, vals, facs, m)
        b = oracle.mttkrp(inds, vals, facs, m, kind="f64")
        assert rel_fro(a, b) <= 1e-12


def test_oracle_solve_normals_satisfies_the_equations():
    rng = np.random.default_rng(1)
    R = 12
    grams = []
    for _ in range(3):
        B = rng.standard_normal((40, R))
        grams.append(B.T @ B)
    M = rng.standard_normal((25, R))
    X, st = A.solve_normals(grams, 1, M)
    V = grams[0] * grams[2]
    assert st == 0
    assert np.abs(X @ V - M).max() <= 1e-9 * np.abs(M).max() * np.linalg.cond(V)
    # Symmetric indefinite V: LU path.
    S = [rng.standard_normal((R, R)) for _ in range(3)]
    S = [s + s.T for s in S]
    X, st = A.solve_normals(S, 0, M)
    assert st == 1
    assert np.abs(X @ (S[1] * S[2]) - M).max() <= 1e-8 * np.abs(M).max() * np.linalg.cond(S[1] * S[2])
    # Exactly singular V: M is returned unchanged.
    Z = [np.zeros((R, R)) for _ in range(3)]
    X, st = A.solve_normals(Z, 0, M)
    assert st == 2 and np.array_equal(X, M)


def test_oracle_fit_equals_dense_residual():
    X, facs, inds, vals, = dense_tensor([7, 6, 5], 3, 2, noise=0.05)
    rng = np.random.default_rng(5)
    K = [rng.uniform(-1, 1, (d, 3)) for d in X.shape]
    lam = rng.uniform(0.5, 2.0, 3)
    M = A.mttkrp_f64(inds, vals, K, 2)
    grams = [k.T @ k for k in K]
    x64 = vals.astype(np.float64)
    fit = A.kruskal_fit(float(x64 @ x64), grams, lam, K[2], M)
    Xf = x64.reshape(X.shape)
    ref = 1 - np.linalg.norm(Xf - reconstruct(lam, K)) / np.linalg.norm(Xf)
    assert abs(fit - ref) <= 1e-10


def test_oracle_als_recovers_exact_low_rank_tensor():
    X, facs, inds, vals = dense_tensor([12, 10, 8], 3, 3)
    rng = np.random.default_rng(4)
    init = [rng.uniform(0, 1, (d, 3)) for d in X.shape]
    lam, fs, fits = A.cp_als(inds, vals, init, 200, 1e-12, round_factors=None)
    assert fits[-1] > 0.9999
    assert all(b >= a - 1e-9 for a, b in zip(fits[1:], fits[2:]))
    rec = reconstruct(lam, fs)
    assert np.linalg.norm(rec - vals.astype(np.float64).reshape(X.shape)) <= 1e-3 * np.linalg.norm(rec)


# ---- host helpers of the library ------------------------------------------


def mat_ptrs(mats):
    return (C.POINTER(F.ptiMatrix) * len(mats))(*[C.pointer(m.address) for m in mats])


def test_dot_mul_seq_triangle_reads_the_upper_triangle():
    rng = np.random.default_rng(6)
    R, N = 9, 3
    full = []
    mats = []
    for _ in range(N):
        B = rng.standard_normal((30, R)).astype(np.float32)
        g = (B.T.astype(np.float64) @ B).astype(np.float32)
        full.append(g.astype(np.float64))
        junk = g.copy()
        junk[np.tril_indices(R, -1)] = 1e30   # the unwritten triangle
        mats.append(DenseMatrix.from_numpy(junk))
    mats.append(DenseMatrix(R, R))
    assert lib().ptiMatrixDotMulSeqTriangle(1, N, mat_ptrs(mats)) == 0
    got = mats[N].numpy().astype(np.float64)
    ref = full[0] * full[2]
    assert rel_fro(got, ref) <= 1e-6
    assert np.array_equal(got, got.T)


def test_kruskal_fit_matches_dense_residual():
    X, facs, inds, vals = dense_tensor([6, 5, 4], 2, 8, noise=0.1)
    rng = np.random.default_rng(9)
    K = [rng.uniform(-1, 1, (d, 2)).astype(np.float32) for d in X.shape]
    lam = rng.uniform(0.5, 2.0, 2).astype(np.float32)
    t = COOTensor.from_arrays(list(X.shape), inds, vals)
    M = A.mttkrp_f64(inds, vals, K, 2).astype(np.float32)
    mats = [DenseMatrix.from_numpy(k) for k in K] + [DenseMatrix.from_numpy(M)]
    ata = [DenseMatrix.from_numpy((k.T.astype(np.float64) @ k).astype(np.float32)) for k in K] + [DenseMatrix(2, 2)]
    lam_c = (F.ptiValue * 2)(*lam)
    fit = lib().KruskalTensorFit(C.byref(t.address), lam_c, mat_ptrs(mats), mat_ptrs(ata))
    x64 = vals.astype(np.float64).reshape(X.shape)
    ref = 1 - np.linalg.norm(x64 - reconstruct(lam.astype(np.float64), [k.astype(np.float64) for k in K])) / \
        np.linalg.norm(x64)
    assert abs(fit - ref) <= 1e-5


def test_get_final_lambda_normalises_and_accumulates():
    rng = np.random.default_rng(10)
    K = [rng.uniform(0.1, 1, (d, 4)).astype(np.float32) for d in (5, 6, 7)]
    mats = [DenseMatrix.from_numpy(k) for k in K]
    lam = (F.ptiValue * 4)(*([2.0] * 4))
    lib().GetFinalLambda(4, 3, mat_ptrs(mats), lam)
    ref = 2.0 * np.prod([np.linalg.norm(k.astype(np.float64), axis=0) for k in K], axis=0)
    assert np.allclose(np.array(lam[:]), ref, rtol=1e-5)
    for m in mats:
        assert np.allclose(np.linalg.norm(m.numpy().astype(np.float64), axis=0), 1.0, rtol=1e-5)


def test_rank_matrix_and_rank_ktensor_containers():
    m = F.ptiRankMatrix()
    assert lib().ptiNewRankMatrix(C.byref(m), 10, 13) == 0
    assert (m.nrows, m.ncols, m.cap, m.stride) == (10, 13, 10, 16)
    lib().ptiRandomizeRankMatrix.argtypes = [C.POINTER(F.ptiRankMatrix), F.ptiIndex, F.ptiElementIndex]
    assert lib().ptiRandomizeRankMatrix(C.byref(m), 10, 13) == 0
    v = np.ctypeslib.as_array(m.values, shape=(10, 16))[:, :13]
    i, j = np.meshgrid(np.arange(10), np.arange(13), indexing="ij")
    assert np.array_equal(v, (i + j + 1).astype(np.float32))
    lib().ptiFreeRankMatrix(C.byref(m))
    assert lib().ptiNewRankMatrix(C.byref(m), 10, 250) != 0   # stride would not fit in 8 bits
    k = F.ptiRankKruskalTensor()
    nd = (F.ptiIndex * 3)(4, 5, 6)
    assert lib().ptiNewRankKruskalTensor(C.byref(k), 3, nd, 7) == 0
    assert k.nmodes == 3 and k.rank == 7 and list(k.ndims[:3]) == [4, 5, 6]
    lib().ptiFreeRankKruskalTensor(C.byref(k))
    assert k.nmodes == 0


def test_oracle_bf16_rounding_is_round_to_nearest_even():
    import torch
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(20000) * 10 ** rng.uniform(-20, 20, 20000),
                        # ties: the low 16 bits exactly 0x8000, even and odd upper halves
                        np.array([0x3F808000, 0x3F818000, 0xBF808000], np.uint32).view(np.float32)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(A.to_bf16(x).astype(np.float32), ref)
    assert A.to_bf16(np.array([0x3F808000], np.uint32).view(np.float32))[0] == 1.0


def test_host_omp_cpd_als_follows_the_fp64_oracle(oracle):
    """orc_omp_cpd_als, the CPU baseline's restatement of ptiOmpCpdAls
    (src/sptensor/cpd_omp.c:32-244), follows the FP64 ALS oracle from the same
    factors: fits, lambda and factors to FP32 rounding, at 1 and 4 threads."""
    rng = np.random.default_rng(7)
    for ndims, R in (((30, 25, 20), 8), ((12, 10, 9, 8, 6), 5)):
        nnz = 3000
        flat = rng.choice(int(np.prod(ndims)), size=nnz, replace=False)
        inds = np.array(np.unravel_index(np.sort(flat), ndims), np.uint32)
        vals = rng.uniform(0, 1, nnz).astype(np.float32)
        init = [rng.uniform(0, 1, (d, R)).astype(np.float32) for d in ndims]
        lam_o, fs_o, fits_o = A.cp_als(inds, vals, [f.astype(np.float64) for f in init], 5, 0.0)
        for nt in (1, 4):
            lam, fs, fits, tm, to = oracle.omp_cpd_als(inds, vals, ndims, init, 5, 0.0, nthreads=nt)
            assert np.allclose(fits, fits_o, atol=1e-5, rtol=0), (fits, fits_o)
            assert rel_fro(lam, lam_o) <= 1e-4
            for f, g in zip(fs, fs_o):
                assert rel_fro(f, g) <= 1e-4
            assert tm >= 0 and to >= 0


def test_host_omp_cpd_als_recovers_exact_low_rank(oracle):
    shape = (20, 15, 10)
    rng = np.random.default_rng(2)
    P = [rng.standard_normal((d, 3)) for d in shape]
    vals = np.einsum("ir,jr,kr->ijk", *P).ravel().astype(np.float32)
    inds = np.array(np.unravel_index(np.arange(vals.size), shape), np.uint32)
    init = [rng.uniform(0, 1, (d, 3)).astype(np.float32) for d in shape]
    lam, fs, fits, _, _ = oracle.omp_cpd_als(inds, vals, shape, init, 300, 1e-10, nthreads=2)
    assert fits[-1] > 0.999, fits[-5:]
