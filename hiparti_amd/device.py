"""Device-resident tensors for torch callers (include/hiparti_dev.h).

PyTorch only provides device memory and streams here: factor matrices are
float32 (or bfloat16) CUDA tensors of shape (rows, stride) whose data
pointers are handed to libHiParTI's entry points, which enqueue the HIP
kernels on torch's current stream.
"""
import ctypes as C

import torch

from . import _ffi as F
from ._lib import check, lib

_vp = C.c_void_p


def _stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return _vp(s.cuda_stream)


def _check_mode(mode, N):
    if not 0 <= mode < N:
        raise ValueError("mode %d out of range for an order-%d tensor" % (mode, N))


def _order(mode, N, order):
    if order is None:
        order = [mode] + [(mode + i) % N for i in range(1, N)]
    return (F.ptiIndex * N)(*order)


def _check_factor(t, dtype, name):
    if not (t.is_cuda and t.dtype == dtype and t.dim() == 2 and t.is_contiguous()):
        raise ValueError("%s must be a contiguous 2-D %s CUDA tensor" % (name, dtype))


def _check_factors(facs, ndims, dtype, R, skip=None, allow_none=False):
    """Validates the factor list the kernels read: one per mode, contiguous
    2-D `dtype` CUDA tensors with at least ndims[m] rows and one shared row
    stride (their column count) >= R.  Entry `skip` (the target mode of an
    MTTKRP, which is not read) is not size-checked, and may be None when
    allow_none is set.  Returns
    the stride.  The kernels index rows and columns from these shapes, so a
    short or narrow tensor would be read past its end on the device."""
    if len(facs) != len(ndims):
        raise ValueError("expected %d factor matrices, got %d" % (len(ndims), len(facs)))
    stride = None
    for m, f in enumerate(facs):
        if f is None and m == skip and allow_none:
            continue
        _check_factor(f, dtype, "factor %d" % m)
        if stride is None:
            stride = f.shape[1]
        if f.shape[1] != stride:
            raise ValueError("factor %d has %d columns; every factor must share the row stride %d"
                             % (m, f.shape[1], stride))
        if m != skip and f.shape[0] < ndims[m]:
            raise ValueError("factor %d has %d rows; mode %d has %d" % (m, f.shape[0], m, ndims[m]))
    if stride is None:
        raise ValueError("no factor matrix given")
    if R is not None and not 1 <= R <= stride:
        raise ValueError("rank %d must be in [1, stride = %d]" % (R, stride))
    return stride


def _check_out(out, rows, cols):
    """The output must be a contiguous float32 CUDA tensor of at least `rows`
    rows and exactly `cols` columns (the row stride the kernels write with)."""
    _check_factor(out, torch.float32, "out")
    if out.shape[0] < rows or out.shape[1] != cols:
        raise ValueError("out must have shape (>= %d, %d), got %s" % (rows, cols, tuple(out.shape)))


class DevHiCOO:
    """A HiCOO tensor uploaded to HBM (ptiDevHiCOOUpload)."""

    def __init__(self, hicoo, stream=None):
        self.handle = _vp()
        check(lib().ptiDevHiCOOUpload(C.byref(self.handle), C.byref(hicoo.address), _stream_ptr(stream)),
              "ptiDevHiCOOUpload")
        self.nmodes = hicoo.nmodes()
        self.ndims = [hicoo.dim_mode(m) for m in range(self.nmodes)]
        nnz, nb, nm, by = C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_uint64()
        lib().ptiDevHiCOOInfo(self.handle, C.byref(nnz), C.byref(nb), C.byref(nm), C.byref(by))
        self.nnz, self.nblocks, self.bytes = int(nnz.value), int(nb.value), int(by.value)

    def mttkrp(self, facs, mode, R=None, out=None, order=None, stream=None):
        """facs: nmodes float32 CUDA tensors (rows, stride); returns the
        (ndims[mode], stride) output (columns >= R are zero)."""
        N = self.nmodes
        _check_mode(mode, N)
        stride = _check_factors(facs, self.ndims, torch.float32, R, skip=mode)
        R = R or stride
        if out is None:
            out = torch.empty((self.ndims[mode], stride), dtype=torch.float32, device=facs[0].device)
        _check_out(out, self.ndims[mode], stride)
        ptrs = (_vp * (N + 1))(*[f.data_ptr() for f in facs], out.data_ptr())
        check(lib().ptiDevMTTKRPHiCOO(self.handle, ptrs, R, stride, _order(mode, N, order), mode,
                                      _stream_ptr(stream)), "ptiDevMTTKRPHiCOO")
        return out

    def mttkrp_bf16(self, facs, mode, R=None, out=None, order=None, stream=None):
        """facs: bfloat16 CUDA tensors (entry `mode` may be None); FP32 output."""
        N = self.nmodes
        _check_mode(mode, N)
        stride = _check_factors(facs, self.ndims, torch.bfloat16, R, skip=mode, allow_none=True)
        ref = next(f for f in facs if f is not None)
        R = R or stride
        if out is None:
            out = torch.empty((self.ndims[mode], stride), dtype=torch.float32, device=ref.device)
        _check_factor(out, torch.float32, "out")
        if out.shape[0] < self.ndims[mode] or out.shape[1] < R:
            raise ValueError("out must have >= %d rows and >= %d columns" % (self.ndims[mode], R))
        ptrs = (_vp * N)(*[(f.data_ptr() if f is not None else None) for f in facs])
        check(lib().ptiDevMTTKRPHiCOO_BF16(self.handle, ptrs, _vp(out.data_ptr()), R, stride, out.shape[1],
                                           _order(mode, N, order), mode, _stream_ptr(stream)),
              "ptiDevMTTKRPHiCOO_BF16")
        return out

    def __del__(self):
        try:
            if self.handle:
                lib().ptiDevHiCOOFree(self.handle)
                self.handle = _vp()
        except Exception:
            pass


class DevCOO:
    """A COO tensor uploaded to HBM (ptiDevCOOUpload)."""

    def __init__(self, coo, stream=None):
        self.handle = _vp()
        check(lib().ptiDevCOOUpload(C.byref(self.handle), C.byref(coo.address), _stream_ptr(stream)),
              "ptiDevCOOUpload")
        self.nmodes = coo.nmodes()
        self.ndims = coo.ndims()
        self.nnz = coo.nnz()

    def prepare(self, mode, stream=None):
        check(lib().ptiDevCOOPrepareMode(self.handle, mode, _stream_ptr(stream)), "ptiDevCOOPrepareMode")

    def mttkrp(self, facs, mode, R=None, out=None, order=None, stream=None):
        N = self.nmodes
        _check_mode(mode, N)
        stride = _check_factors(facs, self.ndims, torch.float32, R, skip=mode)
        R = R or stride
        if out is None:
            out = torch.empty((self.ndims[mode], stride), dtype=torch.float32, device=facs[0].device)
        _check_out(out, self.ndims[mode], stride)
        ptrs = (_vp * (N + 1))(*[f.data_ptr() for f in facs], out.data_ptr())
        check(lib().ptiDevMTTKRPCOO(self.handle, ptrs, R, stride, _order(mode, N, order), mode, _stream_ptr(stream)),
              "ptiDevMTTKRPCOO")
        return out

    def __del__(self):
        try:
            if self.handle:
                lib().ptiDevCOOFree(self.handle)
                self.handle = _vp()
        except Exception:
            pass
