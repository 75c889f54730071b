"""The C-ABI boundary (CPU only, no compute calls on a device).

- libHiParTI.so loads and exports every function declared in include/*.h;
- the struct layouts of include/HiParTI.h equal the ctypes declarations and,
  where /root/reference is present, the reference's own headers
  (include/includes/structs.h), so existing reference objects can be passed;
- host-side argument checks of the MTTKRP entry points return the
  reference's error codes before any device work;
- without a GPU the MTTKRP entry points fail loudly (no CPU fallback).
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from hiparti_amd import COOTensor, DenseMatrix, PtiError, lib
from hiparti_amd import _ffi as F

HEADERS = [os.path.join(ROOT, "include", h) for h in ("HiParTI.h", "hiparti_dev.h", "hiparti_synth.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(", text):
            name = m.group(1)
            if name.startswith(("pti", "Sparse", "Kruskal")) and not name.startswith("ptiAssert"):
                names.add(name)
    return names - {"ptiTimer"}


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "hiparti_amd", "lib", "libHiParTI.so")],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_every_declared_function():
    lib()  # loads
    missing = declared_functions() - exported_symbols()
    assert not missing, "declared but not exported: %s" % sorted(missing)
    assert {"ptiMTTKRP", "ptiMTTKRPHiCOO", "ptiSparseTensorToHiCOO", "ptiCpdAls"} <= declared_functions()


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
%s
#define P(T, f) printf(#T "." #f " %%zu %%zu\n", offsetof(T, f), sizeof(((T *)0)->f));
#define S(T) printf(#T " %%zu\n", sizeof(T));
int main(void) {
  S(ptiMatrix) P(ptiMatrix, nrows) P(ptiMatrix, ncols) P(ptiMatrix, cap) P(ptiMatrix, stride) P(ptiMatrix, values)
  S(ptiSparseTensor) P(ptiSparseTensor, nmodes) P(ptiSparseTensor, sortorder) P(ptiSparseTensor, ndims)
  P(ptiSparseTensor, nnz) P(ptiSparseTensor, inds) P(ptiSparseTensor, values)
  S(ptiSparseTensorHiCOO) P(ptiSparseTensorHiCOO, nmodes) P(ptiSparseTensorHiCOO, nnz)
  P(ptiSparseTensorHiCOO, sb_bits) P(ptiSparseTensorHiCOO, sk_bits) P(ptiSparseTensorHiCOO, sc_bits)
  P(ptiSparseTensorHiCOO, kptr) P(ptiSparseTensorHiCOO, kschr) P(ptiSparseTensorHiCOO, nkiters)
  P(ptiSparseTensorHiCOO, cptr) P(ptiSparseTensorHiCOO, kschr_balanced) P(ptiSparseTensorHiCOO, kschr_balanced_pos)
  P(ptiSparseTensorHiCOO, nkpars) P(ptiSparseTensorHiCOO, kschr_rest) P(ptiSparseTensorHiCOO, knnzs)
  P(ptiSparseTensorHiCOO, bptr) P(ptiSparseTensorHiCOO, binds) P(ptiSparseTensorHiCOO, einds) P(ptiSparseTensorHiCOO, values)
  S(ptiKruskalTensor) P(ptiKruskalTensor, nmodes) P(ptiKruskalTensor, rank) P(ptiKruskalTensor, ndims)
  P(ptiKruskalTensor, lambda) P(ptiKruskalTensor, fit) P(ptiKruskalTensor, factors)
  S(ptiRankMatrix) P(ptiRankMatrix, nrows) P(ptiRankMatrix, ncols) P(ptiRankMatrix, cap) P(ptiRankMatrix, stride)
  P(ptiRankMatrix, values)
  S(ptiRankKruskalTensor) P(ptiRankKruskalTensor, nmodes) P(ptiRankKruskalTensor, rank)
  P(ptiRankKruskalTensor, ndims) P(ptiRankKruskalTensor, lambda) P(ptiRankKruskalTensor, fit)
  P(ptiRankKruskalTensor, factors)
  S(ptiValueVector) S(ptiIndexVector) S(ptiElementIndexVector) S(ptiBlockIndexVector) S(ptiNnzIndexVector)
  return 0;
}
"""


def run_probe(tmp_path, include_line, extra_flags, tag):
    src = tmp_path / ("probe_%s.c" % tag)
    exe = tmp_path / ("probe_%s" % tag)
    src.write_text(PROBE % include_line)
    subprocess.run(["gcc", "-std=gnu99", str(src), "-o", str(exe)] + extra_flags, check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout
    res = {}
    for line in out.splitlines():
        key, _, rest = line.partition(" ")
        res[key] = rest
    return res


def ctypes_layout():
    res = {}
    for name, T in (("ptiMatrix", F.ptiMatrix), ("ptiSparseTensor", F.ptiSparseTensor),
                    ("ptiSparseTensorHiCOO", F.ptiSparseTensorHiCOO), ("ptiKruskalTensor", F.ptiKruskalTensor),
                    ("ptiRankMatrix", F.ptiRankMatrix), ("ptiRankKruskalTensor", F.ptiRankKruskalTensor),
                    ("ptiValueVector", F.ptiValueVector), ("ptiIndexVector", F.ptiIndexVector),
                    ("ptiElementIndexVector", F.ptiElementIndexVector),
                    ("ptiBlockIndexVector", F.ptiBlockIndexVector), ("ptiNnzIndexVector", F.ptiNnzIndexVector)):
        res[name] = str(C.sizeof(T))
        for fname, _ in getattr(T, "_fields_", []):
            f = getattr(T, fname)
            cname = "lambda" if fname == "lambda_" else fname
            res["%s.%s" % (name, cname)] = "%d %d" % (f.offset, f.size)
    return res


def test_struct_layout_matches_ctypes(tmp_path):
    ours = run_probe(tmp_path, '#include "HiParTI.h"', ["-I" + os.path.join(ROOT, "include")], "ours")
    ct = ctypes_layout()
    for k, v in ours.items():
        assert ct[k] == v, k


@pytest.mark.skipif(not os.path.isdir("/root/reference/include"), reason="reference headers not present")
def test_struct_layout_matches_reference_headers(tmp_path):
    ours = run_probe(tmp_path, '#include "HiParTI.h"', ["-I" + os.path.join(ROOT, "include")], "ours")
    theirs = run_probe(tmp_path, '#include "HiParTI.h"', ["-I/root/reference/include"], "ref")
    assert ours == theirs


def _mats(ndims, R, out_rows=None):
    mats = [DenseMatrix(d, R) for d in ndims]
    for m in mats:
        m.constant(1.0)
    mats.append(DenseMatrix(out_rows or max(ndims), R))
    return mats


def test_mttkrp_shape_errors_are_reported_before_device_work():
    t = COOTensor.from_arrays([4, 5, 6], np.array([[0], [1], [2]], np.uint32), np.ones(1, np.float32))
    mats = _mats([4, 5, 6], 8)
    bad = mats[:2] + [DenseMatrix(6, 4)] + [mats[3]]   # wrong ncols
    with pytest.raises(PtiError) as e:
        t.MTTKRP(0, bad)
    assert e.value.code == F.PTIERR_SHAPE_MISMATCH
    bad = mats[:1] + [DenseMatrix(7, 8)] + mats[2:]    # wrong nrows
    with pytest.raises(PtiError) as e:
        t.MTTKRP(0, bad)
    assert e.value.code == F.PTIERR_SHAPE_MISMATCH
    h = t.toHiCOO(1, 1, 1)
    with pytest.raises(PtiError) as e:
        h.MTTKRP(1, bad)
    assert e.value.code == F.PTIERR_SHAPE_MISMATCH


def test_mttkrp_order_must_be_a_permutation_of_the_other_modes():
    t = COOTensor.from_arrays([4, 5, 6], np.array([[0], [1], [2]], np.uint32), np.ones(1, np.float32))
    h = t.toHiCOO(1, 1, 1)
    for order in ([0, 1, 1], [0, 2, 2], [0, 0, 1], [0, 1, 3]):
        for obj in (t, h):
            with pytest.raises(PtiError) as e:
                obj.MTTKRP(0, _mats([4, 5, 6], 8), order)
            assert e.value.code == F.PTIERR_VALUE_ERROR, order


def test_mttkrp_without_gpu_fails_loudly(gpu_available):
    if gpu_available:
        pytest.skip("a GPU is present")
    t = COOTensor.from_arrays([4, 5, 6], np.array([[0], [1], [2]], np.uint32), np.ones(1, np.float32))
    with pytest.raises(PtiError) as e:
        t.MTTKRP(0, _mats([4, 5, 6], 8))
    assert e.value.code & F.PTIERR_CUDA_ERROR
