"""ctypes declarations of the libHiParTI.so C ABI.

The structure layouts mirror include/HiParTI.h, which in turn mirrors the
reference's include/includes/structs.h:27-312.  This module only declares
types and signatures; hiparti_amd/_lib.py loads the library.
"""
import ctypes as C

ptiIndex = C.c_uint32
ptiBlockIndex = C.c_uint32
ptiValue = C.c_float
ptiElementIndex = C.c_uint8
ptiNnzIndex = C.c_uint64

PTIERR_NO_ERROR = 0
PTIERR_UNKNOWN = 1
PTIERR_SHAPE_MISMATCH = 2
PTIERR_VALUE_ERROR = 3
PTIERR_ZERO_DIVISION = 4
PTIERR_OS_ERROR = 0x10000
PTIERR_CUDA_ERROR = 0x20000


def _vec(name, elem):
    return type(name, (C.Structure,), {
        "_fields_": [("len", ptiNnzIndex), ("cap", ptiNnzIndex), ("data", C.POINTER(elem))]})


ptiValueVector = _vec("ptiValueVector", ptiValue)
ptiIndexVector = _vec("ptiIndexVector", ptiIndex)
ptiElementIndexVector = _vec("ptiElementIndexVector", ptiElementIndex)
ptiBlockIndexVector = _vec("ptiBlockIndexVector", ptiBlockIndex)
ptiNnzIndexVector = _vec("ptiNnzIndexVector", ptiNnzIndex)


class ptiMatrix(C.Structure):
    _fields_ = [("nrows", ptiIndex), ("ncols", ptiIndex), ("cap", ptiIndex),
                ("stride", ptiIndex), ("values", C.POINTER(ptiValue))]


class ptiSparseTensor(C.Structure):
    _fields_ = [("nmodes", ptiIndex), ("sortorder", C.POINTER(ptiIndex)),
                ("ndims", C.POINTER(ptiIndex)), ("nnz", ptiNnzIndex),
                ("inds", C.POINTER(ptiIndexVector)), ("values", ptiValueVector)]


class ptiSparseTensorHiCOO(C.Structure):
    _fields_ = [("nmodes", ptiIndex), ("sortorder", C.POINTER(ptiIndex)),
                ("ndims", C.POINTER(ptiIndex)), ("nnz", ptiNnzIndex),
                ("sb_bits", ptiElementIndex), ("sk_bits", ptiElementIndex),
                ("sc_bits", ptiElementIndex),
                ("kptr", ptiNnzIndexVector),
                ("kschr", C.POINTER(C.POINTER(ptiIndexVector))),
                ("nkiters", C.POINTER(ptiIndex)),
                ("cptr", ptiNnzIndexVector),
                ("kschr_balanced", C.POINTER(C.POINTER(ptiIndexVector))),
                ("kschr_balanced_pos", C.POINTER(C.POINTER(ptiIndexVector))),
                ("nkpars", C.POINTER(ptiIndex)),
                ("kschr_rest", C.POINTER(ptiIndexVector)),
                ("knnzs", ptiNnzIndexVector),
                ("bptr", ptiNnzIndexVector),
                ("binds", C.POINTER(ptiBlockIndexVector)),
                ("einds", C.POINTER(ptiElementIndexVector)),
                ("values", ptiValueVector)]


class ptiRankMatrix(C.Structure):
    _fields_ = [("nrows", ptiIndex), ("ncols", ptiElementIndex), ("cap", ptiIndex),
                ("stride", ptiElementIndex), ("values", C.POINTER(ptiValue))]


class ptiRankKruskalTensor(C.Structure):
    _fields_ = [("nmodes", ptiIndex), ("rank", ptiElementIndex), ("ndims", C.POINTER(ptiIndex)),
                ("lambda_", C.POINTER(ptiValue)), ("fit", C.c_double),
                ("factors", C.POINTER(C.POINTER(ptiRankMatrix)))]


class ptiKruskalTensor(C.Structure):
    _fields_ = [("nmodes", ptiIndex), ("rank", ptiIndex), ("ndims", C.POINTER(ptiIndex)),
                ("lambda_", C.POINTER(ptiValue)), ("fit", C.c_double),
                ("factors", C.POINTER(C.POINTER(ptiMatrix)))]


_P = C.POINTER
_vp = C.c_void_p

# name -> (restype, argtypes).  Every symbol declared in include/*.h that a
# Python caller uses is listed here; tests check the library exports all of
# the header's symbols independently.
SIGNATURES = {
    # errors / timers
    "ptiGetLastError": (C.c_int, [_P(C.c_char_p), _P(C.c_char_p), _P(C.c_uint), _P(C.c_char_p)]),
    "ptiClearLastError": (None, []),
    # vectors / matrices
    "ptiNewMatrix": (C.c_int, [_P(ptiMatrix), ptiIndex, ptiIndex]),
    "ptiFreeMatrix": (None, [_P(ptiMatrix)]),
    "ptiConstantMatrix": (C.c_int, [_P(ptiMatrix), ptiValue]),
    "ptiRandomizeMatrix": (C.c_int, [_P(ptiMatrix)]),
    "ptiCopyMatrix": (C.c_int, [_P(ptiMatrix), _P(ptiMatrix)]),
    "ptiMatrix2Norm": (C.c_int, [_P(ptiMatrix), _P(ptiValue)]),
    "ptiMatrixMaxNorm": (C.c_int, [_P(ptiMatrix), _P(ptiValue)]),
    # COO
    "ptiNewSparseTensor": (C.c_int, [_P(ptiSparseTensor), ptiIndex, _P(ptiIndex)]),
    "ptiCopySparseTensor": (C.c_int, [_P(ptiSparseTensor), _P(ptiSparseTensor), C.c_int]),
    "ptiFreeSparseTensor": (None, [_P(ptiSparseTensor)]),
    "ptiLoadSparseTensor": (C.c_int, [_P(ptiSparseTensor), ptiIndex, C.c_char_p]),
    "ptiDumpSparseTensor": (C.c_int, [_P(ptiSparseTensor), ptiIndex, _vp]),
    "ptiDumpSparseTensorHiCOO": (C.c_int, [_P(ptiSparseTensorHiCOO), _vp]),
    "ptiDumpMatrix": (C.c_int, [_P(ptiMatrix), _vp]),
    "ptiDumpKruskalTensor": (C.c_int, [_P(ptiKruskalTensor), _vp]),
    "ptiSparseTensorSortIndex": (None, [_P(ptiSparseTensor), C.c_int, C.c_int]),
    "SparseTensorFrobeniusNormSquared": (C.c_double, [_P(ptiSparseTensor)]),
    "ptiResizeIndexVector": (C.c_int, [_P(ptiIndexVector), ptiNnzIndex]),
    "ptiResizeValueVector": (C.c_int, [_P(ptiValueVector), ptiNnzIndex]),
    # HiCOO
    "ptiSparseTensorToHiCOO": (C.c_int, [_P(ptiSparseTensorHiCOO), _P(ptiNnzIndex), _P(ptiSparseTensor),
                                         ptiElementIndex, ptiElementIndex, ptiElementIndex, C.c_int]),
    "ptiDevSparseTensorToHiCOO": (C.c_int, [_P(ptiSparseTensorHiCOO), _P(ptiNnzIndex), _P(ptiSparseTensor),
                                            ptiElementIndex, ptiElementIndex, ptiElementIndex, _vp]),
    "ptiFreeSparseTensorHiCOO": (None, [_P(ptiSparseTensorHiCOO)]),
    # MTTKRP
    "ptiMTTKRP": (C.c_int, [_P(ptiSparseTensor), _P(_P(ptiMatrix)), _P(ptiIndex), ptiIndex]),
    "ptiMTTKRPHiCOO": (C.c_int, [_P(ptiSparseTensorHiCOO), _P(_P(ptiMatrix)), _P(ptiIndex), ptiIndex]),
    # Kruskal / CPD
    "ptiNewKruskalTensor": (C.c_int, [_P(ptiKruskalTensor), ptiIndex, _P(ptiIndex), ptiIndex]),
    "ptiFreeKruskalTensor": (None, [_P(ptiKruskalTensor)]),
    "ptiCpdAls": (C.c_int, [_P(ptiSparseTensor), ptiIndex, ptiIndex, C.c_double, _P(ptiKruskalTensor)]),
    "ptiOmpCpdAls": (C.c_int, [_P(ptiSparseTensor), ptiIndex, ptiIndex, C.c_double, C.c_int, C.c_int,
                               _P(ptiKruskalTensor)]),
    "ptiCpdAlsHiCOO": (C.c_int, [_P(ptiSparseTensorHiCOO), ptiIndex, ptiIndex, C.c_double,
                                 _P(ptiRankKruskalTensor)]),
    "ptiNewRankKruskalTensor": (C.c_int, [_P(ptiRankKruskalTensor), ptiIndex, _P(ptiIndex), ptiElementIndex]),
    "ptiFreeRankKruskalTensor": (None, [_P(ptiRankKruskalTensor)]),
    "ptiNewRankMatrix": (C.c_int, [_P(ptiRankMatrix), ptiIndex, ptiElementIndex]),
    "ptiFreeRankMatrix": (None, [_P(ptiRankMatrix)]),
    "ptiMatrixSolveNormals": (C.c_int, [ptiIndex, ptiIndex, _P(_P(ptiMatrix)), _P(ptiMatrix)]),
    "ptiMatrixDotMulSeqTriangle": (C.c_int, [ptiIndex, ptiIndex, _P(_P(ptiMatrix))]),
    "KruskalTensorFit": (C.c_double, [_P(ptiSparseTensor), _P(ptiValue), _P(_P(ptiMatrix)), _P(_P(ptiMatrix))]),
    "KruskalTensorFrobeniusNormSquared": (C.c_double, [ptiIndex, _P(ptiValue), _P(_P(ptiMatrix))]),
    "SparseKruskalTensorInnerProduct": (C.c_double, [ptiIndex, _P(ptiValue), _P(_P(ptiMatrix))]),
    "GetFinalLambda": (None, [ptiIndex, ptiIndex, _P(_P(ptiMatrix)), _P(ptiValue)]),
    "ptiOmpMTTKRP": (C.c_int, [_P(ptiSparseTensor), _P(_P(ptiMatrix)), _P(ptiIndex), ptiIndex, C.c_int]),
    "ptiOmpMTTKRPHiCOO": (C.c_int, [_P(ptiSparseTensorHiCOO), _P(_P(ptiMatrix)), _P(ptiIndex), ptiIndex,
                                    C.c_int, C.c_int]),
    "ptiMTTKRPHiCOO_MatrixTiling": (C.c_int, [_P(ptiSparseTensorHiCOO), _P(_P(ptiRankMatrix)), _P(ptiIndex),
                                              ptiIndex]),
    # synthetic inputs
    "ptiSynthSparseTensor": (C.c_int, [_P(ptiSparseTensor), ptiIndex, _P(ptiIndex), ptiNnzIndex,
                                       _P(C.c_int), _P(C.c_double), _P(C.c_int), C.c_int,
                                       C.c_uint64, C.c_int]),
    "ptiSynthUniformMatrix": (C.c_int, [_P(ptiMatrix), C.c_uint64]),
    # device-resident extension
    "ptiDevCount": (C.c_int, []),
    "ptiDevHiCOOUpload": (C.c_int, [_P(_vp), _P(ptiSparseTensorHiCOO), _vp]),
    "ptiDevHiCOOFree": (None, [_vp]),
    "ptiDevHiCOOInfo": (C.c_int, [_vp, _P(C.c_uint64), _P(C.c_uint64), _P(C.c_uint32), _P(C.c_uint64)]),
    "ptiDevCOOUpload": (C.c_int, [_P(_vp), _P(ptiSparseTensor), _vp]),
    "ptiDevCOOPrepareMode": (C.c_int, [_vp, ptiIndex, _vp]),
    "ptiDevCOOFree": (None, [_vp]),
    "ptiDevMTTKRPHiCOO": (C.c_int, [_vp, _P(_vp), ptiIndex, ptiIndex, _P(ptiIndex), ptiIndex, _vp]),
    "ptiDevMTTKRPCOO": (C.c_int, [_vp, _P(_vp), ptiIndex, ptiIndex, _P(ptiIndex), ptiIndex, _vp]),
    "ptiDevMTTKRPCOOValues": (C.c_int, [_vp, _vp, _P(_vp), ptiIndex, ptiIndex, _P(ptiIndex), ptiIndex, _vp]),
    "ptiDevCOOCompletionForward": (C.c_int, [_vp, _P(_vp), ptiIndex, ptiIndex, _vp, _vp, _vp]),
    "ptiDevCOOCompletionForwardNoAlloc": (C.c_int, [_vp, _P(_vp), ptiIndex, ptiIndex, _vp, _vp, _vp]),
    "ptiDevCOOCompletionGrad": (C.c_int, [_vp, _vp, _P(_vp), ptiIndex, ptiIndex, ptiIndex, _vp]),
    "ptiDevCOOCompletionFused": (C.c_int, [_vp, ptiIndex, ptiIndex]),
    "ptiDevCOOCompletionLossGrad": (C.c_int, [_vp, _P(_vp), _P(_vp), ptiIndex, ptiIndex, _vp, _vp]),
    "ptiDevMTTKRPHiCOO_BF16": (C.c_int, [_vp, _P(_vp), _vp, ptiIndex, ptiIndex, ptiIndex,
                                         _P(ptiIndex), ptiIndex, _vp]),
    "ptiDevGram": (C.c_int, [_vp, ptiIndex, ptiIndex, ptiIndex, _vp, _vp]),
    "ptiDevSolveNormals": (C.c_int, [_vp, ptiIndex, ptiIndex, ptiIndex, _vp, _vp, ptiIndex, ptiIndex,
                                     _P(C.c_int), _vp]),
    "ptiDevNormalizeColumns": (C.c_int, [_vp, ptiIndex, ptiIndex, ptiIndex, C.c_int, _vp, _vp]),
    "ptiDevKruskalFit": (C.c_int, [_vp, ptiIndex, ptiIndex, _vp, _vp, _vp, ptiIndex, ptiIndex, C.c_double,
                                   _P(C.c_double), _vp]),
    "ptiDevCpdAlsHiCOO": (C.c_int, [_vp, _P(_vp), ptiIndex, ptiIndex, ptiIndex, C.c_double, ptiIndex, _vp,
                                    _P(C.c_double), _P(C.c_double), _P(ptiIndex), _vp]),
    "ptiDevCpdAlsCOO": (C.c_int, [_vp, _P(_vp), ptiIndex, ptiIndex, ptiIndex, C.c_double, _vp,
                                  _P(C.c_double), _P(C.c_double), _P(ptiIndex), _vp]),
    "ptiRegisterHiCOO": (C.c_int, [_P(ptiSparseTensorHiCOO)]),
    "ptiUnregisterHiCOO": (C.c_int, [_P(ptiSparseTensorHiCOO)]),
    "ptiRegisterSparseTensor": (C.c_int, [_P(ptiSparseTensor)]),
    "ptiUnregisterSparseTensor": (C.c_int, [_P(ptiSparseTensor)]),
}


def declare(lib, names=None):
    """Attach restype/argtypes to the functions of a loaded library."""
    for name, (res, args) in SIGNATURES.items():
        if names is not None and name not in names:
            continue
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib
