"""CP-ALS and its FP64 dense step on device-resident data (torch tensors).

Thin wrappers over libHiParTI's device entry points (include/hiparti_dev.h):
the iteration, the MTTKRPs and the dense step all run in HIP kernels; torch
only owns the memory.  The algorithm is the reference's ptiCpdAls
(src/sptensor/cpd.c:32-208) with the normal equations solved in FP64 on the
device.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List

import torch

from ._lib import check, lib
from .device import DevCOO, DevHiCOO, _check_factor, _check_factors, _stream_ptr

_vp = C.c_void_p


def _vec_ptrs(ts):
    return (_vp * len(ts))(*[t.data_ptr() for t in ts])


def gram(A, R=None):
    """A^T A in FP64 (R x R) of a float32 CUDA matrix (rows, stride)."""
    _check_factor(A, torch.float32, "A")
    R = R or A.shape[1]
    G = torch.empty((R, R), dtype=torch.float64, device=A.device)
    check(lib().ptiDevGram(_vp(A.data_ptr()), A.shape[0], R, A.shape[1], _vp(G.data_ptr()), _stream_ptr()),
          "ptiDevGram")
    return G


def solve_normals(grams, mode, M):
    """(A, status): rows of M solved against V = Hadamard(grams[k], k != mode).
    grams: (N, R, R) float64 CUDA tensor; M: float32 (rows, R) CUDA tensor.
    status 0: Cholesky, 1: LU, 2: V singular (A = M)."""
    if not (grams.is_cuda and grams.dtype == torch.float64 and grams.dim() == 3 and grams.is_contiguous()):
        raise ValueError("grams must be a contiguous (N, R, R) float64 CUDA tensor")
    _check_factor(M, torch.float32, "M")
    N, R, _ = grams.shape
    A = torch.empty_like(M)
    st = C.c_int(-1)
    check(lib().ptiDevSolveNormals(_vp(grams.data_ptr()), N, mode, R, _vp(M.data_ptr()), _vp(A.data_ptr()),
                                   M.shape[0], M.shape[1], C.byref(st), _stream_ptr()), "ptiDevSolveNormals")
    return A, st.value


def normalize_columns(A, max_norm=False, R=None):
    """Divide the columns of A (in place) by their 2-norms, or by their maxima
    clamped to >= 1; returns the FP64 norms."""
    _check_factor(A, torch.float32, "A")
    R = R or A.shape[1]
    lam = torch.empty(R, dtype=torch.float64, device=A.device)
    check(lib().ptiDevNormalizeColumns(_vp(A.data_ptr()), A.shape[0], R, A.shape[1], int(bool(max_norm)),
                                       _vp(lam.data_ptr()), _stream_ptr()), "ptiDevNormalizeColumns")
    return lam


def kruskal_fit(grams, lam, last, M, normX2):
    """1 - sqrt(normX2 + ||K||^2 - 2<X,K>)/sqrt(normX2) from the Grams (N, R, R),
    lambda (R, float64), the last factor and its MTTKRP output."""
    N, R, _ = grams.shape
    fit = C.c_double()
    check(lib().ptiDevKruskalFit(_vp(grams.data_ptr()), N, R, _vp(lam.data_ptr()), _vp(last.data_ptr()),
                                 _vp(M.data_ptr()), last.shape[0], last.shape[1], float(normX2), C.byref(fit),
                                 _stream_ptr()), "ptiDevKruskalFit")
    return fit.value


@dataclass
class CPResult:
    lambda_: torch.Tensor                  # (R,) float64
    factors: List[torch.Tensor]            # (ndims[m], R) float32
    fit: float
    fits: List[float] = field(default_factory=list)
    iters: int = 0


def init_factors(ndims, R, init="uniform", seed=0, device="cuda"):
    """Initial factors: "uniform" U(0,1) from a seeded torch generator, or
    "reference": entry (i, j) = i + j + 1 (src/matrix/matrix.c:94-111)."""
    facs = []
    for m, d in enumerate(ndims):
        if init == "reference":
            i = torch.arange(d, dtype=torch.float32).unsqueeze(1)
            j = torch.arange(R, dtype=torch.float32).unsqueeze(0)
            f = i + j + 1
        elif init == "uniform":
            g = torch.Generator().manual_seed(int(seed) * 1000003 + m)
            f = torch.rand((d, R), generator=g, dtype=torch.float32)
        else:
            raise ValueError(init)
        facs.append(f.to(device).contiguous())
    return facs


def cp_als(tensor, R, niters=50, tol=1e-5, factors=None, init="uniform", seed=0, bf16_iters=0):
    """CP-ALS on a DevHiCOO or DevCOO tensor.  `factors` (float32 CUDA, updated
    in place) overrides `init`.  bf16_iters > 0 (HiCOO only): the first
    iterations' MTTKRPs gather BF16 copies of the factors."""
    ndims = tensor.ndims
    facs = factors if factors is not None else init_factors(ndims, R, init, seed)
    stride = _check_factors(facs, ndims, torch.float32, R)
    lam = torch.empty(R, dtype=torch.float64, device=facs[0].device)
    fit = C.c_double()
    trace = (C.c_double * max(niters, 1))()
    done = C.c_uint32()
    if isinstance(tensor, DevHiCOO):
        rc = lib().ptiDevCpdAlsHiCOO(tensor.handle, _vec_ptrs(facs), R, stride, niters, float(tol), bf16_iters,
                                     _vp(lam.data_ptr()), C.byref(fit), trace, C.byref(done), _stream_ptr())
        check(rc, "ptiDevCpdAlsHiCOO")
    elif isinstance(tensor, DevCOO):
        if bf16_iters:
            raise ValueError("the BF16 schedule needs a HiCOO tensor")
        rc = lib().ptiDevCpdAlsCOO(tensor.handle, _vec_ptrs(facs), R, stride, niters, float(tol),
                                   _vp(lam.data_ptr()), C.byref(fit), trace, C.byref(done), _stream_ptr())
        check(rc, "ptiDevCpdAlsCOO")
    else:
        raise TypeError("tensor must be a DevHiCOO or DevCOO")
    n = int(done.value)
    return CPResult(lam, facs, fit.value, [trace[i] for i in range(n)], n)
