"""GPU check of ptiDevCOOCompletionForwardNoAlloc, the CP-completion forward
that allocates nothing (its partial sums live in a fixed device array).

Called through the C ABI, it must give the same residuals and the same FP64
loss, bit for bit, as ptiDevCOOCompletionForward on the same tensor and
factors (both run the same kernels in the same order).  Two calls in a row,
the second on a smaller tensor, show that the fixed array's leftover contents
from a larger call do not leak into a smaller one; a third call at a rank
whose lane groups step over column blocks covers the wide-rank shape.  The
residuals are also checked against an FP64 evaluation of the formula.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import rel_fro
from hiparti_amd import COOTensor, lib
from hiparti_amd._ffi import _vp
from hiparti_amd.device import DevCOO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu(gpu_available):
    if not gpu_available:
        pytest.fail("no GPU visible: the GPU tests must run on the GPU box")


def forward(entry, dev, facs):
    R = facs[0].shape[1]
    resid = torch.full((dev.nnz,), float("nan"), dtype=torch.float32, device="cuda")
    loss = torch.full((), float("nan"), dtype=torch.float64, device="cuda")
    ptrs = (_vp * len(facs))(*[f.data_ptr() for f in facs])
    rc = getattr(lib(), entry)(dev.handle, ptrs, R, R, _vp(resid.data_ptr()), _vp(loss.data_ptr()), None)
    assert rc == 0, (entry, rc)
    torch.cuda.synchronize()
    return resid.cpu().numpy(), loss.item()


def case(ndims, nnz, R, seed):
    print("no-alloc forward: ndims %s nnz %d R %d seed %d" % (ndims, nnz, R, seed), flush=True)
    X = COOTensor.synthetic(ndims, nnz, seed=seed)
    inds, vals = X.inds(), X.values()
    dev = DevCOO(X)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    facs = [(torch.rand((d, R), generator=gen, device="cuda") - 0.5).contiguous() for d in ndims]
    want_r, want_l = forward("ptiDevCOOCompletionForward", dev, facs)
    got_r, got_l = forward("ptiDevCOOCompletionForwardNoAlloc", dev, facs)
    assert np.array_equal(got_r, want_r)
    assert got_l == want_l
    prod = np.ones((nnz, R), np.float64)
    for m, f in enumerate(facs):
        prod *= f.cpu().numpy().astype(np.float64)[inds[m]]
    ref = vals.astype(np.float64) - prod.sum(axis=1)
    assert rel_fro(got_r, ref) <= 1e-5
    assert abs(got_l - 0.5 * float(ref @ ref)) <= 1e-5 * 0.5 * float(ref @ ref)
    X.free()
    del dev


def test_no_alloc_forward_matches_allocating_forward_on_consecutive_calls():
    case([3000, 2000, 500], 1_500_000, 32, seed=5)   # 1465 partial sums
    case([400, 300, 50], 70_000, 32, seed=6)         # fewer partials than the call before
    case([200, 150, 40, 30], 20_000, 150, seed=7)    # lane groups stepping over column blocks


def test_no_alloc_forward_empty_tensor():
    X = COOTensor.from_arrays([10, 10, 10], np.zeros((3, 0), np.uint32), np.zeros(0, np.float32))
    dev = DevCOO(X)
    facs = [torch.rand((10, 16), device="cuda") for _ in range(3)]
    got_r, got_l = forward("ptiDevCOOCompletionForwardNoAlloc", dev, facs)
    assert got_r.size == 0 and got_l == 0.0
    X.free()
    del dev


def test_no_alloc_forward_rejects_bad_rank():
    X = COOTensor.synthetic([50, 40, 30], 500, seed=8)
    dev = DevCOO(X)
    facs = [torch.rand((d, 16), device="cuda") for d in (50, 40, 30)]
    ptrs = (_vp * 3)(*[f.data_ptr() for f in facs])
    resid = torch.empty(500, device="cuda")
    loss = torch.empty((), dtype=torch.float64, device="cuda")
    rc = lib().ptiDevCOOCompletionForwardNoAlloc(dev.handle, ptrs, 300, 300, _vp(resid.data_ptr()),
                                                 _vp(loss.data_ptr()), None)
    assert rc != 0
    X.free()
    del dev
