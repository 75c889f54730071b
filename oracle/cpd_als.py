"""ORACLE -- test infrastructure only.

FP64 numpy restatement of the reference's CP-ALS (src/sptensor/cpd.c:32-160)
and its dense helpers, used by tests/ to check the device CP-ALS.  Only tests/
import it.

Per iteration, for each mode m (cpd.c:89-137):
  M    = MTTKRP(X, m) with mats_order = (m, m+1, ..., m-1)   (FP64 closed form)
  V    = Hadamard product of the Grams A_k^T A_k, k != m     (matrix.c:341-385)
  A_m  = M V^-1 (Cholesky; else LU; V exactly singular: A_m = M)
                                                             (matrix_solver.c:32-83)
  A_m /= lambda, lambda = column 2-norms at iteration 0      (matrix.c:389-469)
         else column max (not abs) clamped to >= 1           (matrix.c:472-553)
  G_m  = A_m^T A_m
fit = 1 - sqrt(||X||^2 + ||K||^2 - 2<X,K>)/||X|| (ktensor.c:77-99) with
||K||^2 from the full Grams (the reference reads an unwritten triangle,
ktensor.c:128-143) and <X,K> = sum_r lambda_r sum_i A_{N-1}[i,r] M[i,r]
from the last MTTKRP (ktensor.c:152-240).  Stop when |fit - oldfit| < tol
after the first iteration.  Finally every factor is normalised by its column
2-norms, which multiply lambda (matrix.c:556-572).
"""
import numpy as np


def mttkrp_f64(inds, vals, facs, mode):
    """FP64 closed form: M[i_mode, r] = sum_z v_z prod_{k != mode} A_k[i_k, r]."""
    N = len(facs)
    R = facs[0].shape[1]
    prod = np.asarray(vals, np.float64)[:, None] * np.ones((1, R))
    for k in range(N):
        if k != mode:
            prod = prod * np.asarray(facs[k], np.float64)[inds[k]]
    out = np.zeros((facs[mode].shape[0], R))
    np.add.at(out, inds[mode], prod)
    return out


def solve_normals(grams, mode, M):
    """(A, status): A = M V^-1 with V the Hadamard of grams[k], k != mode.
    status 0 Cholesky, 1 LU, 2 singular (A = M)."""
    V = np.ones_like(grams[0])
    for k, g in enumerate(grams):
        if k != mode:
            V = V * g
    try:
        L = np.linalg.cholesky(V)
        Y = np.linalg.solve(L, M.T)
        return np.linalg.solve(L.T, Y).T, 0
    except np.linalg.LinAlgError:
        pass
    try:
        return np.linalg.solve(V, M.T).T, 1
    except np.linalg.LinAlgError:
        return M.copy(), 2


def col_2norm(A):
    lam = np.sqrt((A * A).sum(axis=0))
    return A / lam, lam


def col_maxnorm(A):
    lam = np.maximum(A.max(axis=0), 0.0)
    lam = np.where(lam < 1.0, 1.0, lam)
    return A / lam, lam


def kruskal_fit(normX2, grams, lam, last, M):
    H = np.ones_like(grams[0])
    for g in grams:
        H = H * g
    normK = abs(float(lam @ H @ lam))
    inner = float(((last * M).sum(axis=0) * lam).sum())
    residual = normX2 + normK - 2 * inner
    if residual > 0:
        residual = np.sqrt(residual)
    return 1 - residual / np.sqrt(normX2)


def to_bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64: the
    rounding of the device's BF16 factor copies."""
    x = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    x = (x + 0x7FFF + ((x >> 16) & 1)) & 0xFFFF0000
    return x.astype(np.uint32).view(np.float32).astype(np.float64)


def cp_als(inds, vals, facs, niters, tol, round_factors=np.float32, bf16_iters=0, mttkrp=None):
    """Run the reference's ALS from initial factors `facs` (list of arrays).
    With round_factors=np.float32 each updated factor is rounded to FP32, as
    FP32 factor storage does; None keeps FP64 throughout.  During the first
    bf16_iters iterations the MTTKRPs read bfloat16-rounded copies of the
    factors (the device's BF16-then-FP32 schedule; the solve, the norms and
    the fit are unchanged).  mttkrp(inds, vals, factors, mode) replaces the
    numpy FP64 MTTKRP (e.g. the C oracle's, for large tensors).
    Returns (lambda, factors, fits)."""
    N = len(facs)
    mttkrp = mttkrp or mttkrp_f64
    A = [np.asarray(f, np.float64) for f in facs]
    G = [a.T @ a for a in A]
    normX2 = float(np.dot(np.asarray(vals, np.float64), np.asarray(vals, np.float64)))
    fits = []
    oldfit = 0.0
    lam = None
    M = None
    for it in range(niters):
        for m in range(N):
            M = mttkrp(inds, vals, [to_bf16(a) for a in A] if it < bf16_iters else A, m)
            if round_factors is not None:
                M = M.astype(round_factors).astype(np.float64)
            a, _ = solve_normals(G, m, M)
            if round_factors is not None:
                a = a.astype(round_factors).astype(np.float64)
            a, lam = col_2norm(a) if it == 0 else col_maxnorm(a)
            if round_factors is not None:
                a = a.astype(round_factors).astype(np.float64)
            A[m] = a
            G[m] = a.T @ a
        fit = kruskal_fit(normX2, G, lam, A[N - 1], M)
        fits.append(fit)
        if it > 0 and abs(fit - oldfit) < tol:
            break
        oldfit = fit
    lam = lam.copy()
    for m in range(N):
        A[m], nrm = col_2norm(A[m])
        lam = lam * nrm
    return lam, A, fits
