"""GPU checks of the windowed stream-form order.

A mode whose rows are few and long (here 128 rows of a uniform 2M-nonzero
tensor whose other modes have 40000 rows) gets stream forms whose entries
are ordered, inside each 128-row block, by 4096-row windows of the largest
gathered mode and then by row, with the units dealt to XCDs by their rank
inside the block.  The order only changes which workgroup adds which
contributions, so the results must match the FP64 oracle within the FP32
tolerance: the COO MTTKRP (ptiDevMTTKRPCOO), the HiCOO MTTKRP
(ptiDevMTTKRPHiCOO, sb = 7) and the CP-completion loss and gradients, whose
10k-row-mode pass on the benchmark's tensor is the case the order was built
for.  The other modes (40000 rows, 50 nonzeros per row) keep the row order,
so the same calls also cover a mix of windowed and plain forms on one
tensor.
"""
import numpy as np
import pytest
import torch

from conftest import rel_fro
from hiparti_amd import COOTensor
from hiparti_amd.completion import cp_completion_loss
from hiparti_amd.device import DevCOO, DevHiCOO

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
NDIMS = [40000, 40000, 128]
NNZ = 2_000_000


@pytest.fixture(scope="module", autouse=True)
def _need_gpu(gpu_available):
    if not gpu_available:
        pytest.fail("no GPU visible: the GPU tests must run on the GPU box")


@pytest.fixture(scope="module")
def tensor():
    X = COOTensor.synthetic(NDIMS, NNZ, seed=91)
    yield X, X.inds(), X.values()
    X.free()


@pytest.mark.parametrize("R", [16, 32, 48])
def test_windowed_coo_and_hicoo_mttkrp_match_fp64(tensor, R):
    X, inds, vals = tensor
    rng = np.random.default_rng(R)
    facs = [rng.random((d, R)).astype(np.float32) for d in NDIMS]
    dfacs = [torch.from_numpy(f).cuda() for f in facs]
    dc = DevCOO(X)
    H = X.toHiCOO(7, 14, 14)
    dh = DevHiCOO(H)
    for mode in range(3):
        ref = O.mttkrp(inds, vals, facs, mode, kind="f64")
        ec = rel_fro(dc.mttkrp(dfacs, mode).cpu().numpy()[:, :R], ref)
        eh = rel_fro(dh.mttkrp(dfacs, mode).cpu().numpy()[:, :R], ref)
        print("windowed order R=%d mode %d: COO %.3e HiCOO %.3e" % (R, mode, ec, eh), flush=True)
        assert ec <= TOL and eh <= TOL
    H.free()


def test_windowed_completion_loss_and_gradients_match_fp64(tensor):
    X, inds, vals = tensor
    R = 32
    dc = DevCOO(X)
    gen = torch.Generator(device="cuda").manual_seed(5)
    facs = [(0.5 * torch.rand((d, R), generator=gen, device="cuda")).contiguous() for d in NDIMS]
    facs_np = [f.cpu().numpy() for f in facs]
    params = [f.clone().requires_grad_(True) for f in facs]
    loss = cp_completion_loss(dc, params)
    loss.backward()
    for mode in range(3):
        rows = np.arange(NDIMS[mode], dtype=np.uint32) if NDIMS[mode] <= 128 else \
            np.random.default_rng(mode).choice(NDIMS[mode], 512, replace=False).astype(np.uint32)
        ref_loss, ref_rows = O.completion_f64(inds, vals, facs_np, mode, rows)
        lerr = abs(loss.item() - ref_loss) / ref_loss
        gerr = rel_fro(-params[mode].grad.cpu().numpy()[rows], ref_rows)
        print("windowed completion mode %d: loss %.3e gradient rows %.3e" % (mode, lerr, gerr), flush=True)
        assert lerr <= TOL and gerr <= TOL
