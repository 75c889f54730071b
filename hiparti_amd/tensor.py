"""Python front-end over libHiParTI.so, mirroring HiParTIPy's classes.

Class and method names follow the reference's Python layer
(HiParTIPy/Tensor.py:6-207, HiParTIPy/Matrix.py:281-): COOTensor with
load / dump / toHiCOO / MTTKRP, HiCOOTensor with MTTKRP, DenseMatrix.  Unlike
the reference's cffi bindings, which reach symbols the cffi cdef never
declares (HiParTIPy/PtiCffi.py:7-275), every call here goes to an exported
libHiParTI.so function.  MTTKRP runs on the GPU; there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import _ffi as F
from ._lib import check, lib

_P = C.POINTER


def _as_index_array(seq):
    arr = (F.ptiIndex * len(seq))(*[int(x) for x in seq])
    return arr


class DenseMatrix:
    """Row-major ptiMatrix (stride = ncols rounded up to 8)."""

    def __init__(self, nrows, ncols):
        self.address = F.ptiMatrix()
        check(lib().ptiNewMatrix(C.byref(self.address), int(nrows), int(ncols)), "ptiNewMatrix")
        self._owned = True

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        m = cls(a.shape[0], a.shape[1])
        m.full_view()[:, :a.shape[1]] = a
        m.full_view()[:, a.shape[1]:] = 0
        return m

    def nrows(self):
        return int(self.address.nrows)

    def ncols(self):
        return int(self.address.ncols)

    def stride(self):
        return int(self.address.stride)

    def full_view(self):
        """Writable view of all allocated rows with the padding columns."""
        n = int(self.address.cap) * int(self.address.stride)
        buf = np.ctypeslib.as_array(self.address.values, shape=(n,))
        return buf.reshape(int(self.address.cap), int(self.address.stride))

    def numpy(self):
        """Copy of the nrows x ncols values."""
        return self.full_view()[:self.nrows(), :self.ncols()].copy()

    def constant(self, val):
        check(lib().ptiConstantMatrix(C.byref(self.address), float(val)), "ptiConstantMatrix")

    def randomize(self):
        """Reference initialisation: entry (i, j) = i + j + 1."""
        check(lib().ptiRandomizeMatrix(C.byref(self.address)), "ptiRandomizeMatrix")

    def uniform(self, seed):
        check(lib().ptiSynthUniformMatrix(C.byref(self.address), int(seed)), "ptiSynthUniformMatrix")

    def free(self):
        if getattr(self, "_owned", False):
            lib().ptiFreeMatrix(C.byref(self.address))
            self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _mats_array(mats):
    arr = (_P(F.ptiMatrix) * len(mats))()
    for i, m in enumerate(mats):
        arr[i] = C.pointer(m.address)
    return arr


class _CFile:
    """A C FILE* for the dump functions (libc fopen / fclose)."""
    _libc = None

    def __init__(self, path, mode):
        if _CFile._libc is None:
            libc = C.CDLL(None)
            libc.fopen.restype = C.c_void_p
            libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
            libc.fclose.argtypes = [C.c_void_p]
            _CFile._libc = libc
        self.path, self.mode = path, mode

    def __enter__(self):
        self.fp = _CFile._libc.fopen(self.path.encode(), self.mode.encode())
        if not self.fp:
            raise OSError("cannot open %s" % self.path)
        return C.c_void_p(self.fp)

    def __exit__(self, *exc):
        _CFile._libc.fclose(C.c_void_p(self.fp))
        return False


def _kruskal_out(k, N, R):
    lam = np.array(k.lambda_[:R], np.float32)
    facs = []
    for m in range(N):
        f = k.factors[m].contents
        a = np.ctypeslib.as_array(f.values, shape=(int(f.nrows), int(f.stride)))[:, :R].copy()
        facs.append(a)
    return lam, facs, float(k.fit)


def _default_order(mode, nmodes):
    return [mode] + [(mode + i) % nmodes for i in range(1, nmodes)]


class COOTensor:
    """COO sparse tensor (ptiSparseTensor)."""

    def __init__(self):
        self.address = F.ptiSparseTensor()
        self._owned = False

    # ---- construction ----------------------------------------------------
    @classmethod
    def from_arrays(cls, ndims, inds, vals):
        """inds: (nmodes, nnz) integer array, vals: (nnz,) floats."""
        inds = np.ascontiguousarray(inds, dtype=np.uint32)
        vals = np.ascontiguousarray(vals, dtype=np.float32)
        N, nnz = inds.shape
        t = cls()
        check(lib().ptiNewSparseTensor(C.byref(t.address), N, _as_index_array(ndims)), "ptiNewSparseTensor")
        t._owned = True
        for m in range(N):
            check(lib().ptiResizeIndexVector(C.byref(t.address.inds[m]), nnz), "ptiResizeIndexVector")
            if nnz:
                C.memmove(t.address.inds[m].data, inds[m].ctypes.data, nnz * 4)
        check(lib().ptiResizeValueVector(C.byref(t.address.values), nnz), "ptiResizeValueVector")
        if nnz:
            C.memmove(t.address.values.data, vals.ctypes.data, nnz * 4)
        t.address.nnz = nnz
        return t

    @classmethod
    def synthetic(cls, ndims, nnz, dist=None, zipf_s=None, permute=None, lognormal=False, seed=0,
                  nthreads=0):
        """Seeded synthetic tensor (see include/hiparti_synth.h)."""
        N = len(ndims)
        dist = dist or [0] * N
        zipf_s = zipf_s or [1.0] * N
        permute = permute or [0] * N
        t = cls()
        rc = lib().ptiSynthSparseTensor(C.byref(t.address), N, _as_index_array(ndims), int(nnz),
                                        (C.c_int * N)(*dist), (C.c_double * N)(*zipf_s),
                                        (C.c_int * N)(*permute), 1 if lognormal else 0, int(seed),
                                        int(nthreads))
        check(rc, "ptiSynthSparseTensor")
        t._owned = True
        return t

    def load(self, filename, start_index=1):
        check(lib().ptiLoadSparseTensor(C.byref(self.address), int(start_index), filename.encode()),
              "ptiLoadSparseTensor")
        self._owned = True
        return self

    def dump(self, filename, start_index=1):
        """Text coordinate dump (ptiDumpSparseTensor, HiParTIPy/Tensor.py:21-23)."""
        with _CFile(filename, "w") as fp:
            check(lib().ptiDumpSparseTensor(C.byref(self.address), int(start_index), fp), "ptiDumpSparseTensor")

    # ---- accessors --------------------------------------------------------
    def nmodes(self):
        return int(self.address.nmodes)

    def nnz(self):
        return int(self.address.nnz)

    def dim_mode(self, mode):
        return int(self.address.ndims[mode])

    def ndims(self):
        return [int(self.address.ndims[m]) for m in range(self.nmodes())]

    def inds(self):
        n = self.nnz()
        out = np.empty((self.nmodes(), n), dtype=np.uint32)
        for m in range(self.nmodes()):
            if n:
                out[m] = np.ctypeslib.as_array(self.address.inds[m].data, shape=(n,))
        return out

    def values(self):
        n = self.nnz()
        if n == 0:
            return np.empty(0, np.float32)
        return np.ctypeslib.as_array(self.address.values.data, shape=(n,)).copy()

    # ---- operations -------------------------------------------------------
    def sort(self, nthreads=0):
        lib().ptiSparseTensorSortIndex(C.byref(self.address), 1, int(nthreads))

    def toHiCOO(self, b=7, k=7, c=7, nthreads=0, device=False):
        """Convert to HiCOO.  As in the reference, this sorts the COO tensor
        in place (src/sptensor/hicoo/convert.c:454,590).  device=True orders
        and emits the blocks on the GPU (ptiDevSparseTensorToHiCOO); the
        result is the same."""
        h = HiCOOTensor()
        mx = F.ptiNnzIndex()
        if device:
            check(lib().ptiDevSparseTensorToHiCOO(C.byref(h.address), C.byref(mx), C.byref(self.address),
                                                  int(b), int(k), int(c), None), "ptiDevSparseTensorToHiCOO")
        else:
            check(lib().ptiSparseTensorToHiCOO(C.byref(h.address), C.byref(mx), C.byref(self.address),
                                               int(b), int(k), int(c), int(nthreads)), "ptiSparseTensorToHiCOO")
        h._owned = True
        h.max_nnzb = int(mx.value)
        return h

    def MTTKRP(self, mode, mats=None, matsorder=None, type="default", testing=False):
        """GPU MTTKRP through ptiMTTKRP.  mats: nmodes+1 DenseMatrix objects,
        the last one receives the result; returns it as a numpy array.
        `type` selects among the reference's CPU/CUDA variants in HiParTIPy
        (Tensor.py:128-146); every variant here is the same GPU kernel."""
        N = self.nmodes()
        if matsorder is None:
            matsorder = _default_order(mode, N)
        check(lib().ptiMTTKRP(C.byref(self.address), _mats_array(mats), _as_index_array(matsorder), int(mode)),
              "ptiMTTKRP")
        out = mats[N]
        return out.full_view()[:self.dim_mode(mode), :out.ncols()].copy()

    def norm_squared(self):
        return float(lib().SparseTensorFrobeniusNormSquared(C.byref(self.address)))

    def cpd(self, rank, niters=50, tol=1e-5):
        """ptiCpdAls (reference src/sptensor/cpd.c:163-208) on the GPU:
        returns (lambda, factors, fit) as numpy arrays and a float."""
        k = F.ptiKruskalTensor()
        check(lib().ptiNewKruskalTensor(C.byref(k), self.nmodes(), _as_index_array(self.ndims()), int(rank)),
              "ptiNewKruskalTensor")
        try:
            check(lib().ptiCpdAls(C.byref(self.address), int(rank), int(niters), float(tol), C.byref(k)),
                  "ptiCpdAls")
            return _kruskal_out(k, self.nmodes(), int(rank))
        finally:
            lib().ptiFreeKruskalTensor(C.byref(k))

    def free(self):
        if self._owned:
            lib().ptiFreeSparseTensor(C.byref(self.address))
            self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HiCOOTensor:
    """HiCOO sparse tensor (ptiSparseTensorHiCOO)."""

    def __init__(self):
        self.address = F.ptiSparseTensorHiCOO()
        self._owned = False
        self.max_nnzb = 0

    def nmodes(self):
        return int(self.address.nmodes)

    def nnz(self):
        return int(self.address.nnz)

    def nblocks(self):
        return int(self.address.bptr.len) - 1

    def dim_mode(self, mode):
        return int(self.address.ndims[mode])

    def arrays(self):
        """Copies of the HiCOO arrays (see hicoo_struct_arrays)."""
        return hicoo_struct_arrays(self.address)

    def dump(self, filename):
        """Text dump of every HiCOO array (ptiDumpSparseTensorHiCOO,
        HiParTIPy/Tensor.py:156-158, which opens the file for reading; here it
        is opened for writing)."""
        with _CFile(filename, "w") as fp:
            check(lib().ptiDumpSparseTensorHiCOO(C.byref(self.address), fp), "ptiDumpSparseTensorHiCOO")

    def cpd(self, rank, niters=50, tol=1e-5):
        """ptiCpdAlsHiCOO (reference src/sptensor/hicoo/cpd.c:153-204) on the
        GPU, rank <= 248: returns (lambda, factors, fit)."""
        k = F.ptiRankKruskalTensor()
        nd = [self.dim_mode(m) for m in range(self.nmodes())]
        check(lib().ptiNewRankKruskalTensor(C.byref(k), self.nmodes(), _as_index_array(nd), int(rank)),
              "ptiNewRankKruskalTensor")
        try:
            check(lib().ptiCpdAlsHiCOO(C.byref(self.address), int(rank), int(niters), float(tol), C.byref(k)),
                  "ptiCpdAlsHiCOO")
            return _kruskal_out(k, self.nmodes(), int(rank))
        finally:
            lib().ptiFreeRankKruskalTensor(C.byref(k))

    def MTTKRP(self, mode, mats=None, matsorder=None, type="default", testing=False):
        N = self.nmodes()
        if matsorder is None:
            matsorder = _default_order(mode, N)
        check(lib().ptiMTTKRPHiCOO(C.byref(self.address), _mats_array(mats), _as_index_array(matsorder),
                                   int(mode)), "ptiMTTKRPHiCOO")
        out = mats[N]
        return out.full_view()[:self.dim_mode(mode), :out.ncols()].copy()

    def free(self):
        if self._owned:
            lib().ptiFreeSparseTensorHiCOO(C.byref(self.address))
            self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def hicoo_struct_arrays(a):
    """Copies of the arrays of a ptiSparseTensorHiCOO struct as numpy arrays.
    kschr-like per-row lists are flattened per mode into (ptr, data)."""
    N = int(a.nmodes)
    nnz = int(a.nnz)
    nb = int(a.bptr.len) - 1

    def vec(v, dtype):
        n = int(v.len)
        if n == 0:
            return np.empty(0, dtype)
        return np.ctypeslib.as_array(v.data, shape=(n,)).astype(dtype, copy=True)

    sk = 1 << int(a.sk_bits)
    out = {
        "bptr": vec(a.bptr, np.uint64),
        "binds": np.stack([vec(a.binds[m], np.uint32)[:nb] for m in range(N)]),
        "einds": np.stack([vec(a.einds[m], np.uint8)[:nnz] for m in range(N)]),
        "values": vec(a.values, np.float32)[:nnz],
        "kptr": vec(a.kptr, np.uint64),
        "cptr": vec(a.cptr, np.uint64),
        "knnzs": vec(a.knnzs, np.uint64),
        "nkiters": np.array([a.nkiters[m] for m in range(N)], np.uint32),
        "nkpars": np.array([a.nkpars[m] for m in range(N)], np.uint32),
    }
    for name, field in (("kschr", a.kschr), ("bal", a.kschr_balanced), ("pos", a.kschr_balanced_pos)):
        for m in range(N):
            kd = (int(a.ndims[m]) + sk - 1) // sk
            lists = [vec(field[m][i], np.uint32) for i in range(kd)]
            ptr = np.zeros(kd + 1, np.uint64)
            if kd:
                ptr[1:] = np.cumsum([len(x) for x in lists])
            out["%s_ptr%d" % (name, m)] = ptr
            out["%s_dat%d" % (name, m)] = (np.concatenate(lists).astype(np.uint32)
                                            if kd and ptr[-1] else np.empty(0, np.uint32))
    for m in range(N):
        out["rest%d" % m] = vec(a.kschr_rest[m], np.uint32)
    return out
