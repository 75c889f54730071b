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


def _order(mode, N, order):
    if order is None:
        order = [mode] + [(mode + i) % N for i in range(1, N)]
    return (F.ptiIndex * N)(*order)


def _check_factor(t, dtype, name):
    if not (t.is_cuda and t.dtype == dtype and t.dim() == 2 and t.is_contiguous()):
        raise ValueError("%s must be a contiguous 2-D %s CUDA tensor" % (name, dtype))


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
        stride = facs[0].shape[1]
        R = R or stride
        for i, f in enumerate(facs):
            _check_factor(f, torch.float32, "factor %d" % i)
        if out is None:
            out = torch.empty((self.ndims[mode], stride), dtype=torch.float32, device=facs[0].device)
        ptrs = (_vp * (N + 1))(*[f.data_ptr() for f in facs], out.data_ptr())
        check(lib().ptiDevMTTKRPHiCOO(self.handle, ptrs, R, stride, _order(mode, N, order), mode,
                                      _stream_ptr(stream)), "ptiDevMTTKRPHiCOO")
        return out

    def mttkrp_bf16(self, facs, mode, R=None, out=None, order=None, stream=None):
        """facs: bfloat16 CUDA tensors (entry `mode` may be None); FP32 output."""
        N = self.nmodes
        ref = next(f for f in facs if f is not None)
        stride = ref.shape[1]
        R = R or stride
        for i, f in enumerate(facs):
            if f is not None:
                _check_factor(f, torch.bfloat16, "factor %d" % i)
        if out is None:
            out = torch.empty((self.ndims[mode], stride), dtype=torch.float32, device=ref.device)
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
        stride = facs[0].shape[1]
        R = R or stride
        for i, f in enumerate(facs):
            _check_factor(f, torch.float32, "factor %d" % i)
        if out is None:
            out = torch.empty((self.ndims[mode], stride), dtype=torch.float32, device=facs[0].device)
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
