"""GPU checks of the FP64 dense step and of CP-ALS on the device.

Dense kernels against the FP64 restatement (oracle/cpd_als.py, itself
pinned in test_cpd_host.py) and numpy:
  Gram (ranks 1..256, padded strides, empty/one-row matrices), the
  normal-equation solve on its three paths (Cholesky, LU for an indefinite V,
  unchanged rows for a singular V), column 2-norm and max-norm, the fit
  against the dense residual of a small tensor.
CP-ALS (COO and HiCOO, device API and the drop-in ptiCpdAls /
ptiCpdAlsHiCOO) against the oracle's ALS trajectory from the same initial
factors, recovery of an exactly low-rank tensor, and the BF16-then-FP32
schedule against the FP32 run.  Drop-in ptiCpdAls starts from the
reference's i+j+1 factors, which have rank 2 and make the normal equations singular for
R > 2; the drop-in comparison uses R = 1, where that initialisation is well
posed.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest
import torch

from conftest import ROOT, rel_fro
from hiparti_amd import COOTensor, DenseMatrix, lib
from hiparti_amd import _ffi as F
from hiparti_amd import cpd as D
from hiparti_amd.device import DevCOO, DevHiCOO

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import cpd_als as A  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu(gpu_available):
    if not gpu_available:
        pytest.fail("no GPU visible: the GPU tests must run on the GPU box")


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("nrows,R,stride", [(0, 8, 8), (1, 5, 8), (1000, 1, 8), (5000, 16, 16), (70001, 16, 16),
                                            (3000, 10, 16), (4097, 32, 40), (20000, 33, 40), (3000, 64, 64),
                                            (777, 100, 104), (300, 256, 256)])
def test_gram_matches_fp64(nrows, R, stride):
    rng = np.random.default_rng(nrows + R)
    a = np.zeros((nrows, stride), np.float32)
    a[:, :R] = rng.uniform(-1, 1, (nrows, R))
    a[:, R:] = np.nan     # padding columns are never read
    G = D.gram(cuda(a), R).cpu().numpy()
    ref = a[:, :R].astype(np.float64).T @ a[:, :R].astype(np.float64)
    assert np.array_equal(G, G.T)
    if nrows == 0:
        assert np.all(G == 0)
    else:
        assert rel_fro(G, ref) <= 1e-12


def spd_grams(rng, N, R, rows=50):
    gs = []
    for _ in range(N):
        B = rng.uniform(0, 1, (rows, R))
        gs.append(B.T @ B)
    return np.stack(gs)


@pytest.mark.parametrize("R,mode", [(1, 0), (7, 1), (16, 2), (32, 0), (64, 1), (130, 2)])
def test_solve_normals_cholesky(R, mode):
    rng = np.random.default_rng(R)
    grams = spd_grams(rng, 3, R, rows=3 * R + 10)
    M = rng.uniform(-1, 1, (2000, R)).astype(np.float32)
    X, st = D.solve_normals(cuda(grams), mode, cuda(M))
    ref, st_ref = A.solve_normals(list(grams), mode, M.astype(np.float64))
    assert st == 0 == st_ref
    assert rel_fro(X.cpu().numpy(), ref) <= 1e-6


@pytest.mark.parametrize("R", [24, 12])
def test_solve_normals_lu_and_singular(R):
    rng = np.random.default_rng(3)
    S = rng.standard_normal((3, R, R))
    S = S + S.transpose(0, 2, 1)            # symmetric, indefinite Hadamard product
    M = rng.uniform(-1, 1, (500, R)).astype(np.float32)
    X, st = D.solve_normals(cuda(S), 0, cuda(M))
    ref, st_ref = A.solve_normals(list(S), 0, M.astype(np.float64))
    assert st == 1 == st_ref
    assert rel_fro(X.cpu().numpy(), ref) <= 1e-6
    Z = np.zeros((3, R, R))
    X, st = D.solve_normals(cuda(Z), 2, cuda(M))
    assert st == 2
    assert torch.equal(X.cpu(), torch.from_numpy(M))


@pytest.mark.parametrize("nrows,R", [(1, 4), (5000, 16), (3333, 33), (100, 256)])
def test_normalize_columns(nrows, R):
    rng = np.random.default_rng(nrows)
    a = rng.uniform(-2, 3, (nrows, R)).astype(np.float32)
    a[:, 0] = -1.0            # max-norm: the column maximum is not an absolute value
    t = cuda(a)
    lam = D.normalize_columns(t).cpu().numpy()
    ref, lref = A.col_2norm(a.astype(np.float64))
    assert rel_fro(lam, lref) <= 1e-12
    assert rel_fro(t.cpu().numpy(), ref) <= 1e-7
    t = cuda(a)
    lam = D.normalize_columns(t, max_norm=True).cpu().numpy()
    ref, lref = A.col_maxnorm(a.astype(np.float64))
    assert lam[0] == 1.0
    assert rel_fro(lam, lref) <= 1e-12
    assert rel_fro(t.cpu().numpy(), ref) <= 1e-7


def test_kruskal_fit_matches_dense_residual():
    rng = np.random.default_rng(11)
    shape = (9, 8, 7)
    X = rng.uniform(0, 1, shape)
    inds = np.array(np.unravel_index(np.arange(X.size), shape), np.uint32)
    vals = X.ravel().astype(np.float32)
    K = [rng.uniform(-1, 1, (d, 4)).astype(np.float32) for d in shape]
    lam = rng.uniform(0.5, 2, 4)
    M = A.mttkrp_f64(inds, vals, K, 2).astype(np.float32)
    grams = np.stack([k.astype(np.float64).T @ k.astype(np.float64) for k in K])
    x64 = vals.astype(np.float64)
    fit = D.kruskal_fit(cuda(grams), cuda(lam), cuda(K[2]), cuda(M), float(x64 @ x64))
    Xf = x64.reshape(shape)
    rec = np.einsum("r,ir,jr,kr->ijk", lam, *[k.astype(np.float64) for k in K])
    ref = 1 - np.linalg.norm(Xf - rec) / np.linalg.norm(Xf)
    assert abs(fit - ref) <= 1e-6


def random_sparse(ndims, nnz, seed, lowrank=None):
    rng = np.random.default_rng(seed)
    flat = rng.choice(int(np.prod(ndims)), size=nnz, replace=False)
    inds = np.array(np.unravel_index(np.sort(flat), ndims), np.uint32)
    if lowrank is None:
        vals = rng.uniform(0, 1, nnz).astype(np.float32)
    else:
        P = [rng.uniform(0.2, 1, (d, lowrank)) for d in ndims]
        prod = np.ones((nnz, lowrank))
        for m, p in enumerate(P):
            prod *= p[inds[m]]
        vals = prod.sum(axis=1).astype(np.float32)
    return inds, vals


def device_tensor(kind, ndims, inds, vals):
    t = COOTensor.from_arrays(ndims, inds, vals)
    if kind == "coo":
        return DevCOO(t), t
    h = t.toHiCOO(3, 5, 5)
    return DevHiCOO(h), h


@pytest.mark.parametrize("kind", ["coo", "hicoo"])
@pytest.mark.parametrize("ndims,nnz,R", [((30, 25, 20), 4000, 8), ((12, 10, 9, 8), 3000, 5)])
def test_cp_als_follows_fp64_oracle(kind, ndims, nnz, R):
    inds, vals = random_sparse(ndims, nnz, 7)
    dev, _keep = device_tensor(kind, list(ndims), inds, vals)
    init = D.init_factors(ndims, R, "uniform", seed=3)
    init_np = [f.cpu().numpy().astype(np.float64) for f in init]
    res = D.cp_als(dev, R, niters=4, tol=0.0, factors=[f.clone() for f in init])
    lam, fs, fits = A.cp_als(inds, vals, init_np, 4, 0.0)
    assert res.iters == 4 and len(res.fits) == 4
    assert np.allclose(res.fits, fits, atol=1e-4, rtol=0), (res.fits, fits)
    assert rel_fro(res.lambda_.cpu().numpy(), lam) <= 1e-3
    for f, g in zip(res.factors, fs):
        assert rel_fro(f.cpu().numpy(), g) <= 1e-3


@pytest.mark.parametrize("kind", ["coo", "hicoo"])
def test_cp_als_recovers_planted_low_rank(kind):
    ndims = (40, 35, 30)
    inds = np.array(np.unravel_index(np.arange(int(np.prod(ndims))), ndims), np.uint32)
    rng = np.random.default_rng(2)
    # Gaussian factors: well conditioned, so ALS leaves the swamp quickly
    # (all-positive planted factors converge far more slowly).
    P = [rng.standard_normal((d, 3)) for d in ndims]
    vals = np.einsum("ir,jr,kr->ijk", *P).ravel().astype(np.float32)
    dev, _keep = device_tensor(kind, list(ndims), inds, vals)
    res = D.cp_als(dev, 3, niters=200, tol=1e-9, seed=5)
    assert res.fit > 0.999, res.fits[-5:]
    rec = np.einsum("r,ir,jr,kr->ijk", res.lambda_.cpu().numpy(), *[f.cpu().numpy().astype(np.float64)
                                                                      for f in res.factors])
    assert rel_fro(rec.ravel(), vals.astype(np.float64)) <= 2e-3


def test_bf16_then_fp32_schedule_reaches_the_fp32_fit():
    ndims = (60, 50, 40)
    inds, vals = random_sparse(ndims, 30000, 4, lowrank=4)
    dev, _keep = device_tensor("hicoo", list(ndims), inds, vals)
    r32 = D.cp_als(dev, 4, niters=30, tol=0.0, seed=1)
    rmix = D.cp_als(dev, 4, niters=30, tol=0.0, seed=1, bf16_iters=10)
    assert rmix.iters == 30
    assert abs(rmix.fit - r32.fit) <= 1e-3, (rmix.fit, r32.fit)
    # The BF16 iterations are not the FP32 ones.
    assert rmix.fits[:10] != r32.fits[:10]


@pytest.mark.parametrize("R", [65, 75, 130])
def test_cp_als_coo_wide_odd_ranks_follow_fp64_oracle(R):
    """Ranks above 64 that are odd or 2 mod 4 run the COO kernel in column
    tiles of 64 lanes x 1 or 2 ranks.  Gaussian initial factors keep the
    Hadamard of the Grams well conditioned, so the FP32 device run and the
    FP64 oracle stay close over the iterations."""
    ndims = (400, 300, 250)
    inds, vals = random_sparse(ndims, 30000, R)
    dev, _keep = device_tensor("coo", list(ndims), inds, vals)
    rng = np.random.default_rng(R)
    init = [rng.standard_normal((d, R)).astype(np.float32) for d in ndims]
    res = D.cp_als(dev, R, niters=3, tol=0.0, factors=[cuda(f) for f in init])
    lam, fs, fits = A.cp_als(inds, vals, [f.astype(np.float64) for f in init], 3, 0.0)
    assert np.allclose(res.fits, fits, atol=1e-4, rtol=0), (res.fits, fits)
    assert rel_fro(res.lambda_.cpu().numpy(), lam) <= 1e-3
    for f, g in zip(res.factors, fs):
        assert rel_fro(f.cpu().numpy(), g) <= 1e-3


def test_dropin_cpd_als_runs_at_odd_rank_above_64():
    # The reference's i+j+1 start makes the normal equations singular at this
    # rank, so only the call's completion and finite results are checked here;
    # the COO MTTKRP it runs is checked against the oracle above.
    ndims = [200, 150, 100]
    inds, vals = random_sparse(ndims, 5000, 21)
    t = COOTensor.from_arrays(ndims, inds, vals)
    k = F.ptiKruskalTensor()
    assert lib().ptiNewKruskalTensor(C.byref(k), 3, (F.ptiIndex * 3)(*ndims), 65) == 0
    assert lib().ptiCpdAls(C.byref(t.address), 65, 2, 0.0, C.byref(k)) == 0
    assert np.isfinite(k.fit)
    assert np.all(np.isfinite(np.array(k.lambda_[:65])))
    lib().ptiFreeKruskalTensor(C.byref(k))


C4_DIMS = [1000000, 500000, 100000, 1000, 24]


def c4_tensor(nnz, seed):
    """The end-to-end CP-ALS configuration's shape and index distributions
    (order 5, Zipf(1.1) on modes 0-2, uniform on modes 3-4, U(0,1) values) at
    a smaller nonzero count."""
    return COOTensor.synthetic(C4_DIMS, nnz, dist=[1, 1, 1, 0, 0], zipf_s=[1.1, 1.1, 1.1, 1, 1],
                               permute=[0] * 5, seed=seed)


@pytest.mark.parametrize("bf16_iters", [0, 3])
def test_cp_als_order5_c4_shape_follows_fp64_oracle(bf16_iters, oracle):
    """Device CP-ALS (HiCOO, stream kernel of order 5, R = 16) on a 1.5M-nonzero
    tensor of the CP-ALS configuration's shape against the oracle's FP64 ALS
    from the same factors over 6 iterations, all-FP32 and with the first 3
    iterations' MTTKRPs on BF16 factor copies (the oracle rounds its factor
    copies the same way).  Fits must agree within 1e-4 either way.  Lambda
    and the factors must agree within 1e-3 for the all-FP32 run; with BF16
    copies the bar is 1e-2: an FP32-versus-FP64 difference of ~1e-7 in a
    factor entry can round it to the neighbouring BF16 value (a 2^-8 relative
    step) on one side only, and ALS on this tensor, which has no low-rank
    structure, carries such steps into lambda (measured 1.3e-3 after 6
    iterations while the fits agreed to 5e-8)."""
    R = 16
    X = c4_tensor(1_500_000, 44)
    inds, vals = X.inds(), X.values()
    H = X.toHiCOO(7, 14, 14)
    dev = DevHiCOO(H)
    init = D.init_factors(C4_DIMS, R, "uniform", seed=4)
    init_np = [f.cpu().numpy().astype(np.float64) for f in init]
    res = D.cp_als(dev, R, niters=6, tol=0.0, factors=[f.clone() for f in init], bf16_iters=bf16_iters)

    def mttkrp(i, v, f, m):
        return oracle.mttkrp(i, v, f, m, kind="f64")
    lam, fs, fits = A.cp_als(inds, vals, init_np, 6, 0.0, bf16_iters=bf16_iters, mttkrp=mttkrp)
    print("order-5 CP-ALS, bf16_iters=%d: device fits %s, oracle fits %s" % (bf16_iters, res.fits, fits), flush=True)
    assert res.iters == 6
    assert np.allclose(res.fits, fits, atol=1e-4, rtol=0), (res.fits, fits)
    tol = 1e-3 if bf16_iters == 0 else 1e-2
    assert rel_fro(res.lambda_.cpu().numpy(), lam) <= tol
    for f, g in zip(res.factors, fs):
        assert rel_fro(f.cpu().numpy(), g) <= tol


@pytest.mark.parametrize("seed", [8, 10])
def test_bf16_then_fp32_reaches_the_fp32_fit_on_a_planted_order5_tensor(seed):
    """Every entry of a 12 x 10 x 9 x 8 x 6 tensor observed, values from a
    planted rank-3 model (factors U(0.2, 1)) plus 5% Gaussian noise, ALS
    started from the planted factors perturbed by 30%: all-FP32 ALS and the
    schedule with 10 BF16 iterations first must end at the same fit (within
    1e-4) after 60 iterations, both at the noise level (fit ~0.95).  The noise
    keeps the residual norm well above the fit formula's cancellation floor
    (sqrt(||X||^2 + ||K||^2 - 2<X,K>) is only good to ~sqrt(eps) * ||X|| when
    the model is exact), and the warm start keeps ALS out of the poor local
    optima that random starts reach on such small tensors."""
    shape = (12, 10, 9, 8, 6)
    rng = np.random.default_rng(seed)
    P = [rng.uniform(0.2, 1.0, (d, 3)) for d in shape]
    inds = np.array(np.unravel_index(np.arange(int(np.prod(shape))), shape), np.uint32)
    prod = np.ones((inds.shape[1], 3))
    for m, p in enumerate(P):
        prod *= p[inds[m]]
    clean = prod.sum(axis=1)
    noise = 0.05 * np.linalg.norm(clean) / np.sqrt(clean.size)
    vals = (clean + noise * rng.standard_normal(clean.size)).astype(np.float32)
    init = [torch.from_numpy((p * (1 + 0.3 * rng.standard_normal(p.shape))).astype(np.float32)).cuda()
            for p in P]
    dev, _keep = device_tensor("hicoo", list(shape), inds, vals)
    r32 = D.cp_als(dev, 3, niters=60, tol=0.0, factors=[f.clone() for f in init])
    rmix = D.cp_als(dev, 3, niters=60, tol=0.0, factors=[f.clone() for f in init], bf16_iters=10)
    print("planted order 5 + 5%% noise: FP32 fit %.8f, BF16-then-FP32 fit %.8f" % (r32.fit, rmix.fit), flush=True)
    assert 0.94 < r32.fit < 0.96
    assert abs(rmix.fit - r32.fit) <= 1e-4, (rmix.fit, r32.fit)


def test_dropin_cpd_als_with_reference_init():
    # Rank 1: the reference's initial factor (i + 1) is well posed.
    ndims = [20, 15, 10]
    inds, vals = random_sparse(ndims, 1500, 9)
    t = COOTensor.from_arrays(ndims, inds, vals)
    k = F.ptiKruskalTensor()
    assert lib().ptiNewKruskalTensor(C.byref(k), 3, (F.ptiIndex * 3)(*ndims), 1) == 0
    assert lib().ptiCpdAls(C.byref(t.address), 1, 5, 0.0, C.byref(k)) == 0
    init = [np.arange(1, d + 1, dtype=np.float64)[:, None] for d in ndims]
    lam, fs, fits = A.cp_als(inds, vals, init, 5, 0.0)
    assert abs(k.fit - fits[-1]) <= 1e-5, (k.fit, fits)
    assert rel_fro(np.array(k.lambda_[:1], np.float64), lam) <= 1e-5
    for m in range(3):
        f = k.factors[m].contents
        got = np.ctypeslib.as_array(f.values, shape=(f.nrows, f.stride))[:, :1]
        assert rel_fro(got, fs[m]) <= 1e-5
    lib().ptiFreeKruskalTensor(C.byref(k))


def test_dropin_cpd_als_hicoo_with_reference_init():
    ndims = [20, 15, 10]
    inds, vals = random_sparse(ndims, 1500, 9)
    t = COOTensor.from_arrays(ndims, inds, vals)
    h = t.toHiCOO(3, 4, 4)
    k = F.ptiRankKruskalTensor()
    assert lib().ptiNewRankKruskalTensor(C.byref(k), 3, (F.ptiIndex * 3)(*ndims), 1) == 0
    assert lib().ptiCpdAlsHiCOO(C.byref(h.address), 1, 5, 0.0, C.byref(k)) == 0
    init = [np.arange(1, d + 1, dtype=np.float64)[:, None] for d in ndims]
    lam, fs, fits = A.cp_als(inds, vals, init, 5, 0.0)
    assert abs(k.fit - fits[-1]) <= 1e-5, (k.fit, fits)
    assert rel_fro(np.array(k.lambda_[:1], np.float64), lam) <= 1e-5
    lib().ptiFreeRankKruskalTensor(C.byref(k))


def test_dropin_solve_normals_on_host_matrices():
    rng = np.random.default_rng(13)
    R, N = 10, 3
    grams = spd_grams(rng, N, R)
    ata = []
    for g in grams:
        junk = g.astype(np.float32)
        junk[np.tril_indices(R, -1)] = 7.0    # only the upper triangle is read
        ata.append(DenseMatrix.from_numpy(junk))
    ata.append(DenseMatrix(R, R))
    M = rng.uniform(-1, 1, (300, R)).astype(np.float32)
    rhs = DenseMatrix.from_numpy(M)
    ptrs = (C.POINTER(F.ptiMatrix) * (N + 1))(*[C.pointer(m.address) for m in ata])
    assert lib().ptiMatrixSolveNormals(0, N, ptrs, C.byref(rhs.address)) == 0
    g32 = [g.astype(np.float32).astype(np.float64) for g in grams]
    ref, _ = A.solve_normals(g32, 0, M.astype(np.float64))
    assert rel_fro(rhs.numpy(), ref) <= 1e-6


def test_mttkrp_variants_match_fp64(oracle):
    ndims = [25, 20, 15]
    inds, vals = random_sparse(ndims, 2000, 12)
    rng = np.random.default_rng(1)
    facs = [rng.uniform(-1, 1, (d, 8)).astype(np.float32) for d in ndims]
    t = COOTensor.from_arrays(ndims, inds, vals)
    h = t.toHiCOO(3, 4, 4)
    order = (F.ptiIndex * 3)(1, 2, 0)
    for mode in range(3):
        order = (F.ptiIndex * 3)(*([mode] + [(mode + i) % 3 for i in (1, 2)]))
        ref = oracle.mttkrp(inds, vals, facs, mode, kind="f64")
        mats = [DenseMatrix.from_numpy(f) for f in facs] + [DenseMatrix(max(ndims), 8)]
        ptrs = (C.POINTER(F.ptiMatrix) * 4)(*[C.pointer(m.address) for m in mats])
        assert lib().ptiOmpMTTKRP(C.byref(t.address), ptrs, order, mode, 16) == 0
        assert rel_fro(mats[3].numpy()[:ndims[mode]], ref) <= 1e-5
        assert lib().ptiOmpMTTKRPHiCOO(C.byref(h.address), ptrs, order, mode, 16, 1) == 0
        assert rel_fro(mats[3].numpy()[:ndims[mode]], ref) <= 1e-5
        rms = []
        for f in facs + [np.zeros((max(ndims), 8), np.float32)]:
            rm = F.ptiRankMatrix()
            assert lib().ptiNewRankMatrix(C.byref(rm), f.shape[0], 8) == 0
            np.ctypeslib.as_array(rm.values, shape=(f.shape[0], rm.stride))[:, :8] = f
            rms.append(rm)
        rptrs = (C.POINTER(F.ptiRankMatrix) * 4)(*[C.pointer(r) for r in rms])
        assert lib().ptiMTTKRPHiCOO_MatrixTiling(C.byref(h.address), rptrs, order, mode) == 0
        got = np.ctypeslib.as_array(rms[3].values, shape=(max(ndims), rms[3].stride))[:ndims[mode], :8]
        assert rel_fro(got, ref) <= 1e-5
        for r in rms:
            lib().ptiFreeRankMatrix(C.byref(r))
