"""MI355X-native sparse MTTKRP / CPD-ALS with HiParTI's C API.

The compute path is libHiParTI.so (hand-written HIP kernels for gfx950 behind
a C ABI, hiparti_amd/csrc).  This package holds the Python front-end
(tensor.py, mirroring HiParTIPy), the torch/autograd integration (device.py)
and the build script (build.py).
"""
from ._lib import PtiError, lib  # noqa: F401
from .tensor import COOTensor, DenseMatrix, HiCOOTensor  # noqa: F401

__all__ = ["COOTensor", "HiCOOTensor", "DenseMatrix", "PtiError", "lib"]
