"""Loads the in-tree libHiParTI.so (built by hiparti_amd/build.py).

There is no fallback: if the library is missing the import fails with a
message saying how to build it, and the MTTKRP entry points return
PTIERR_CUDA_ERROR + code when no HIP device is usable.
"""
import ctypes
import os

from . import _ffi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libHiParTI.so")

_lib = None


class PtiError(RuntimeError):
    """A libHiParTI call returned a non-zero PTIERR_* code."""

    def __init__(self, code, func, reason=None):
        self.code = code
        msg = "%s failed with error 0x%x" % (func, code)
        if reason:
            msg += ": " + reason
        super().__init__(msg)


def _share_torch_hip_runtime():
    """Load torch's HIP runtime before libHiParTI.so when torch is present.

    torch's libraries request the runtime by the unversioned name
    libamdhip64.so from their own directory, while libHiParTI.so requests
    libamdhip64.so.7.  If libHiParTI.so were loaded first, torch would later
    load a second copy of the runtime, and one process would then hold two HIP
    runtimes (torch then sees no device).  Loaded in the other order, the
    dynamic loader resolves libHiParTI.so's libamdhip64.so.7 to torch's copy by
    its soname, so both share one runtime, one device context and one set of
    streams."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libHiParTI.so not found at %s; build it with "
                              "`python -m hiparti_amd.build`" % LIB_PATH)
        _share_torch_hip_runtime()
        _lib = _ffi.declare(ctypes.CDLL(LIB_PATH))
    return _lib


def last_error_reason():
    reason = ctypes.c_char_p()
    lib().ptiGetLastError(None, None, None, ctypes.byref(reason))
    return reason.value.decode() if reason.value else None


def check(rc, func):
    if rc != 0:
        raise PtiError(rc, func, last_error_reason())
    return rc
