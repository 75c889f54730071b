"""ORACLE -- test infrastructure only.

numpy front-end to oracle/liboracle.so (pti_oracle.c), the CPU restatement of
the reference's MTTKRP path.  Only tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py use it, and only to check results or as the
reported CPU baseline.  Build: make -C oracle.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_P = C.POINTER
_u32p = _P(C.c_uint32)
_u64p = _P(C.c_uint64)
_f32p = _P(C.c_float)
_f64p = _P(C.c_double)
_u8p = _P(C.c_uint8)


class OrcHiCOO(C.Structure):
    _fields_ = [("nmodes", C.c_uint32), ("nnz", C.c_uint64), ("nb", C.c_uint64), ("nk", C.c_uint64),
                ("nc", C.c_uint64), ("max_nnzb", C.c_uint64),
                ("bptr", _u64p), ("binds", _P(_u32p)), ("einds", _P(_u8p)), ("values", _f32p),
                ("kptr", _u64p), ("cptr", _u64p), ("knnzs", _u64p),
                ("nkiters", _u32p), ("nkpars", _u32p), ("kernel_ndim", _u32p),
                ("kschr_ptr", _P(_u64p)), ("kschr_dat", _P(_u32p)),
                ("bal_ptr", _P(_u64p)), ("bal_dat", _P(_u32p)),
                ("pos_ptr", _P(_u64p)), ("pos_dat", _P(_u32p)),
                ("rest_len", _u64p), ("rest_dat", _P(_u32p))]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        mt = [C.c_uint32, C.c_uint64, _P(_u32p), _f32p, C.c_uint32, C.c_uint32, _P(_f32p), _u32p, C.c_uint32,
              C.c_uint64]
        L.orc_mttkrp_f64.argtypes = mt + [_f64p]
        L.orc_mttkrp_coo_seq.argtypes = mt + [_f32p]
        L.orc_omp_mttkrp_coo.argtypes = mt + [_f32p, C.c_int]
        ht = [C.c_uint32, C.c_uint64, _u64p, _u64p, _P(_u32p), _P(_u8p), _f32p, C.c_uint32, C.c_uint32, C.c_uint32,
              _P(_f32p), _u32p, C.c_uint32, C.c_uint64, _f32p]
        L.orc_mttkrp_hicoo_seq.argtypes = ht
        L.orc_omp_mttkrp_hicoo.argtypes = ht + [C.c_int]
        L.orc_hicoo_convert.argtypes = [C.c_uint32, _u32p, C.c_uint64, _P(_u32p), _f32p, C.c_uint32, C.c_uint32,
                                        C.c_uint32, _P(OrcHiCOO)]
        L.orc_hicoo_convert.restype = C.c_int
        L.orc_hicoo_free.argtypes = [_P(OrcHiCOO)]
        L.orc_sort_lexicographic.argtypes = [C.c_uint32, C.c_uint64, _P(_u32p), _f32p]
        rt = [C.c_uint32, C.c_uint64, _P(_u32p), _f32p, C.c_uint32, C.c_uint32, _P(_f32p), C.c_uint32, C.c_uint64,
              C.c_uint32, _u32p, _f64p]
        L.orc_mttkrp_f64_rows.argtypes = rt + [C.c_int]
        L.orc_completion_f64.argtypes = rt + [_f64p, C.c_int]
        L.orc_omp_cpd_als.argtypes = [C.c_uint32, C.c_uint64, _P(_u32p), _f32p, _u32p, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_double, _P(_f32p), _f32p, _f64p, _u32p, _f64p, _f64p, C.c_int]
        L.orc_omp_cpd_als.restype = C.c_double
        L.orc_mttkrp_f64_rows.restype = None
        L.orc_completion_f64.restype = None
        for f in ("orc_mttkrp_f64", "orc_mttkrp_coo_seq", "orc_omp_mttkrp_coo", "orc_mttkrp_hicoo_seq",
                  "orc_omp_mttkrp_hicoo", "orc_hicoo_free", "orc_sort_lexicographic"):
            getattr(L, f).restype = None
        _lib = L
    return _lib


def _ptr_array(arrs, ctype):
    out = (_P(ctype) * len(arrs))()
    for i, a in enumerate(arrs):
        out[i] = a.ctypes.data_as(_P(ctype))
    return out


def _factors(facs, R):
    """Contiguous float32 copies with stride R."""
    return [np.ascontiguousarray(np.asarray(f, np.float32)[:, :R]) for f in facs]


def _default_order(mode, N):
    return [mode] + [(mode + i) % N for i in range(1, N)]


def mttkrp(inds, vals, facs, mode, order=None, nrows_out=None, kind="f64", nthreads=0):
    """COO MTTKRP on the CPU.  kind: "f64" (FP64 closed form), "seq" (the
    reference's sequential FP32 order) or "omp" (the OpenMP baseline)."""
    inds = np.ascontiguousarray(inds, np.uint32)
    vals = np.ascontiguousarray(vals, np.float32)
    N, nnz = inds.shape
    R = facs[0].shape[1]
    fs = _factors(facs, R)
    order = order or _default_order(mode, N)
    nrows_out = nrows_out if nrows_out is not None else fs[mode].shape[0]
    rows = [inds[m] for m in range(N)]
    args = [N, nnz, _ptr_array(rows, C.c_uint32), vals.ctypes.data_as(_f32p), R, R, _ptr_array(fs, C.c_float),
            (C.c_uint32 * N)(*order), mode, nrows_out]
    if kind == "f64":
        out = np.zeros((nrows_out, R), np.float64)
        lib().orc_mttkrp_f64(*args, out.ctypes.data_as(_f64p))
    elif kind == "seq":
        out = np.zeros((nrows_out, R), np.float32)
        lib().orc_mttkrp_coo_seq(*args, out.ctypes.data_as(_f32p))
    elif kind == "omp":
        out = np.zeros((nrows_out, R), np.float32)
        lib().orc_omp_mttkrp_coo(*args, out.ctypes.data_as(_f32p), int(nthreads or os.cpu_count()))
    else:
        raise ValueError(kind)
    return out


def _row_args(inds, vals, facs, mode, rows):
    inds = np.ascontiguousarray(inds, np.uint32)
    vals = np.ascontiguousarray(vals, np.float32)
    N, nnz = inds.shape
    R = facs[0].shape[1]
    fs = [np.ascontiguousarray(f, np.float32) for f in facs]
    stride = fs[0].shape[1]
    rows = np.ascontiguousarray(rows if rows is not None else np.zeros(0), np.uint32)
    keep = (inds, vals, fs, rows)
    args = [N, nnz, _ptr_array([inds[m] for m in range(N)], C.c_uint32), vals.ctypes.data_as(_f32p), R, stride,
            _ptr_array(fs, C.c_float), mode, fs[mode].shape[0], len(rows), rows.ctypes.data_as(_u32p)]
    return args, keep, R


def mttkrp_rows(inds, vals, facs, mode, rows, nthreads=0):
    """FP64 MTTKRP restricted to output rows `rows` (len(rows) x R), for
    accuracy checks on tensors too large for the full FP64 closed form.
    Factors are used with their own row stride (all must share it)."""
    args, keep, R = _row_args(inds, vals, facs, mode, rows)
    out = np.zeros((len(keep[3]), R), np.float64)
    lib().orc_mttkrp_f64_rows(*args, out.ctypes.data_as(_f64p), int(nthreads))
    return out


def completion_f64(inds, vals, facs, mode=0, rows=None, nthreads=0):
    """FP64 CP-completion loss 0.5 sum_z e_z^2 and, for output rows `rows` of
    `mode`, the MTTKRP of the residuals e_z (-dL/dA_mode there).  Returns
    (loss, rows x R array)."""
    args, keep, R = _row_args(inds, vals, facs, mode, rows)
    out = np.zeros((len(keep[3]), R), np.float64)
    loss = C.c_double()
    lib().orc_completion_f64(*args, out.ctypes.data_as(_f64p), C.byref(loss), int(nthreads))
    return loss.value, out


def omp_cpd_als(inds, vals, ndims, facs, niters, tol=0.0, nthreads=0):
    """Host CP-ALS (orc_omp_cpd_als, the restatement of ptiOmpCpdAls with
    use_reduce = 0) from initial factors `facs` (float32, rows x R).
    Returns (lambda, factors, fits, seconds in MTTKRPs, seconds elsewhere)."""
    inds = np.ascontiguousarray(inds, np.uint32)
    vals = np.ascontiguousarray(vals, np.float32)
    N, nnz = inds.shape
    R = facs[0].shape[1]
    fs = [np.array(f, np.float32, order="C", copy=True) for f in facs]
    lam = np.zeros(R, np.float32)
    trace = np.zeros(max(niters, 1), np.float64)
    done = C.c_uint32()
    tm, to = C.c_double(), C.c_double()
    lib().orc_omp_cpd_als(N, nnz, _ptr_array([inds[m] for m in range(N)], C.c_uint32), vals.ctypes.data_as(_f32p),
                          (C.c_uint32 * N)(*[int(d) for d in ndims]), R, R, niters, float(tol),
                          _ptr_array(fs, C.c_float), lam.ctypes.data_as(_f32p), trace.ctypes.data_as(_f64p),
                          C.byref(done), C.byref(tm), C.byref(to), int(nthreads or os.cpu_count()))
    return lam, fs, list(trace[:done.value]), tm.value, to.value


def mttkrp_hicoo(h, sb_bits, facs, mode, order=None, nrows_out=None, kind="seq", nthreads=0):
    """HiCOO MTTKRP on the CPU from a dict of HiCOO arrays (the keys of
    hicoo_convert / hiparti_amd.tensor.hicoo_struct_arrays)."""
    binds = np.ascontiguousarray(h["binds"], np.uint32)
    einds = np.ascontiguousarray(h["einds"], np.uint8)
    N = binds.shape[0]
    R = facs[0].shape[1]
    fs = _factors(facs, R)
    order = order or _default_order(mode, N)
    nrows_out = nrows_out if nrows_out is not None else fs[mode].shape[0]
    kptr = np.ascontiguousarray(h["kptr"], np.uint64)
    bptr = np.ascontiguousarray(h["bptr"], np.uint64)
    vals = np.ascontiguousarray(h["values"], np.float32)
    out = np.zeros((nrows_out, R), np.float32)
    args = [N, len(kptr) - 1, kptr.ctypes.data_as(_u64p), bptr.ctypes.data_as(_u64p),
            _ptr_array([binds[m] for m in range(N)], C.c_uint32), _ptr_array([einds[m] for m in range(N)], C.c_uint8),
            vals.ctypes.data_as(_f32p), sb_bits, R, R, _ptr_array(fs, C.c_float), (C.c_uint32 * N)(*order), mode,
            nrows_out, out.ctypes.data_as(_f32p)]
    if kind == "seq":
        lib().orc_mttkrp_hicoo_seq(*args)
    elif kind == "omp":
        lib().orc_omp_mttkrp_hicoo(*args, int(nthreads or os.cpu_count()))
    else:
        raise ValueError(kind)
    return out


def hicoo_convert(ndims, inds, vals, sb, sk, sc):
    """Restated reference conversion.  Returns (arrays dict, sorted inds,
    sorted vals); the inputs are not modified."""
    inds = np.array(inds, np.uint32, copy=True, order="C")
    vals = np.array(vals, np.float32, copy=True)
    N, nnz = inds.shape
    h = OrcHiCOO()
    rows = [inds[m] for m in range(N)]
    rc = lib().orc_hicoo_convert(N, (C.c_uint32 * N)(*ndims), nnz, _ptr_array(rows, C.c_uint32),
                                 vals.ctypes.data_as(_f32p), sb, sk, sc, C.byref(h))
    if rc:
        raise ValueError("orc_hicoo_convert failed (%d)" % rc)

    def arr(p, n, dtype):
        if n == 0:
            return np.empty(0, dtype)
        return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)

    out = {
        "bptr": arr(h.bptr, h.nb + 1, np.uint64),
        "binds": np.stack([arr(h.binds[m], h.nb, np.uint32) for m in range(N)]),
        "einds": np.stack([arr(h.einds[m], nnz, np.uint8) for m in range(N)]),
        "values": arr(h.values, nnz, np.float32),
        "kptr": arr(h.kptr, h.nk + 1, np.uint64),
        "cptr": arr(h.cptr, h.nc, np.uint64),
        "knnzs": arr(h.knnzs, h.nk, np.uint64),
        "nkiters": arr(h.nkiters, N, np.uint32),
        "nkpars": arr(h.nkpars, N, np.uint32),
        "max_nnzb": np.uint64(h.max_nnzb),
    }
    for m in range(N):
        kd = h.kernel_ndim[m]
        for name, P, D in (("kschr", h.kschr_ptr, h.kschr_dat), ("bal", h.bal_ptr, h.bal_dat),
                           ("pos", h.pos_ptr, h.pos_dat)):
            ptr = arr(P[m], kd + 1, np.uint64)
            out["%s_ptr%d" % (name, m)] = ptr
            out["%s_dat%d" % (name, m)] = arr(D[m], int(ptr[-1]), np.uint32)
        out["rest%d" % m] = arr(h.rest_dat[m], h.rest_len[m], np.uint32)
    lib().orc_hicoo_free(C.byref(h))
    return out, inds, vals


def sort_lexicographic(inds, vals):
    inds = np.array(inds, np.uint32, copy=True, order="C")
    vals = np.array(vals, np.float32, copy=True)
    N, nnz = inds.shape
    lib().orc_sort_lexicographic(N, nnz, _ptr_array([inds[m] for m in range(N)], C.c_uint32),
                                 vals.ctypes.data_as(_f32p))
    return inds, vals
