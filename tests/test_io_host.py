"""Tensor file formats (CPU only): ptiLoadSparseTensor on the text
coordinate format and the SPLATT binary format (reference
src/sptensor/load.c:25-272), and ptiDumpSparseTensor round trips.

Pinned by the files' own contents (the arrays written are the arrays read)
and, where oracle/_ref/libhiparti_ref.so has been built from the reference
sources in this container, by the reference's own loader on the same files.
"""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from conftest import ROOT
from hiparti_amd import COOTensor, PtiError
from hiparti_amd import _ffi as F

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libhiparti_ref.so")


def write_bin(path, ndims, inds, vals, idx_width=4, val_width=4, magic=0):
    with open(path, "wb") as f:
        f.write(struct.pack("<iQQ", magic, idx_width, val_width))
        f.write(struct.pack("<I", len(ndims)))
        f.write(np.asarray(ndims, np.uint32).tobytes())
        if idx_width == 8:
            f.write(struct.pack("<Q", inds.shape[1]))
        else:
            f.write(struct.pack("<I", inds.shape[1]))
        for m in range(len(ndims)):
            f.write(np.asarray(inds[m], np.uint32).tobytes())
        f.write(np.asarray(vals, np.float32 if val_width == 4 else np.float64).tobytes())


def random_tensor(ndims, nnz, seed):
    rng = np.random.default_rng(seed)
    inds = np.array([rng.integers(0, d, nnz) for d in ndims], np.uint32)
    vals = rng.standard_normal(nnz)
    return inds, vals


def ref_load(path, start_index=1):
    """The reference's ptiLoadSparseTensor on `path` (arrays copied out)."""
    L = C.CDLL(REF_LIB)
    L.ptiLoadSparseTensor.argtypes = [C.POINTER(F.ptiSparseTensor), F.ptiIndex, C.c_char_p]
    L.ptiFreeSparseTensor.argtypes = [C.POINTER(F.ptiSparseTensor)]
    t = F.ptiSparseTensor()
    assert L.ptiLoadSparseTensor(C.byref(t), start_index, path.encode()) == 0
    N, nnz = t.nmodes, t.nnz
    inds = np.array([np.ctypeslib.as_array(t.inds[m].data, shape=(nnz,)) for m in range(N)]) if nnz else \
        np.zeros((N, 0), np.uint32)
    vals = np.ctypeslib.as_array(t.values.data, shape=(nnz,)).copy() if nnz else np.zeros(0, np.float32)
    ndims = [t.ndims[m] for m in range(N)]
    L.ptiFreeSparseTensor(C.byref(t))
    return ndims, inds, vals


@pytest.mark.parametrize("val_width", [4, 8])
def test_binary_round_trip(tmp_path, val_width):
    ndims = [17, 40, 9, 3]
    inds, vals = random_tensor(ndims, 2000, val_width)
    path = str(tmp_path / "t.bin")
    write_bin(path, ndims, inds, vals, val_width=val_width)
    t = COOTensor().load(path)
    assert t.ndims() == ndims and t.nnz() == 2000
    assert np.array_equal(t.inds(), inds)
    assert np.array_equal(t.values(), vals.astype(np.float32))
    if os.path.exists(REF_LIB):
        rd, ri, rv = ref_load(path)
        assert rd == ndims and np.array_equal(ri, inds) and np.array_equal(rv, t.values())


def test_binary_rejects_wide_indices_and_truncation(tmp_path):
    ndims = [5, 6]
    inds, vals = random_tensor(ndims, 10, 1)
    path = str(tmp_path / "wide.bin")
    write_bin(path, ndims, inds, vals, idx_width=8)
    with pytest.raises(PtiError) as e:
        COOTensor().load(path)
    assert e.value.code == F.PTIERR_VALUE_ERROR
    path = str(tmp_path / "short.bin")
    write_bin(path, ndims, inds, vals)
    data = open(path, "rb").read()
    open(path, "wb").write(data[:-8])
    with pytest.raises(PtiError):
        COOTensor().load(path)


def test_text_round_trip_and_reference_loader(tmp_path):
    ndims = [30, 20, 10]
    inds, vals = random_tensor(ndims, 500, 2)
    path = str(tmp_path / "t.tns")
    with open(path, "w") as f:
        f.write("3\n30 20 10\n")
        for z in range(500):
            f.write("%d %d %d %.9g\n" % (inds[0, z] + 1, inds[1, z] + 1, inds[2, z] + 1, vals[z]))
    t = COOTensor().load(path)
    assert t.ndims() == ndims and t.nnz() == 500
    assert np.array_equal(t.inds(), inds)
    assert np.array_equal(t.values(), np.array(["%.9g" % v for v in vals], np.float64).astype(np.float32))
    if os.path.exists(REF_LIB):
        rd, ri, rv = ref_load(path)
        assert rd == ndims and np.array_equal(ri, t.inds()) and np.array_equal(rv, t.values())


def test_text_tokens_and_stop_at_malformed_record(tmp_path):
    # Records may span lines; reading stops at the first malformed record.
    path = str(tmp_path / "odd.tns")
    with open(path, "w") as f:
        f.write("2\n4 5\n1 1 0.5\n2\n3 -1.25e1\n4 5 7\n3 x 9\n1 1 1\n")
    t = COOTensor().load(path)
    assert t.nnz() == 3
    assert np.array_equal(t.inds(), np.array([[0, 1, 3], [0, 2, 4]], np.uint32))
    assert np.array_equal(t.values(), np.array([0.5, -12.5, 7.0], np.float32))
    if os.path.exists(REF_LIB):
        rd, ri, rv = ref_load(path)
        assert np.array_equal(ri, t.inds()) and np.array_equal(rv, t.values())


def test_dump_then_load(tmp_path):
    ndims = [12, 11, 10, 9, 8]
    inds, vals = random_tensor(ndims, 300, 3)
    vals = np.round(vals, 3)      # the dump prints values with %f
    t = COOTensor.from_arrays(ndims, inds, vals.astype(np.float32))
    path = str(tmp_path / "d.tns")
    t.dump(path)
    u = COOTensor().load(path)
    assert u.ndims() == ndims
    assert np.array_equal(u.inds(), inds)
    assert np.allclose(u.values(), vals.astype(np.float32), atol=1e-6)
