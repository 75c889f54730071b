"""GPU checks of the stream kernel with LDS-resident factors.

For orders 4 and 5 at four ranks per lane (FP32 and BF16), the last
gathered modes in ascending mode order whose whole rank-tile copies fit
beside the LDS tile in 80 KB are copied into LDS by every workgroup (at most
one for order 4, two for order 5, and only where a unit holds at least four
nonzeros per resident row), and the walk reads them there.  The tensors
below are shaped so that this happens: short trailing modes and few, long
block-rows.  For order 5 at R = 16 and mode 0 the two trailing modes (300
and 24 rows) are resident, for mode 4 only the 300-row one (the 1000-row
mode would not fit beside it); for order 4 at R = 64 the 100-row mode is
resident for modes 0-2 with FP32 factors (BF16 factors take eight ranks per
lane there and no copies).  R = 96 runs two column tiles (64 and 32 ranks)
that each copy their own columns of the resident factors.  Every output is
compared with the FP64 oracle (BF16: on the bf16-rounded factors), and the
same tensors at R = 18 (two ranks per lane, no resident factors) cover the
plain walk on the same data.
"""
import numpy as np
import pytest
import torch

from conftest import rel_fro
from hiparti_amd import COOTensor
from hiparti_amd.device import DevHiCOO

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def _need_gpu(gpu_available):
    if not gpu_available:
        pytest.fail("no GPU visible: the GPU tests must run on the GPU box")


CASES = {
    "order5": ([2000, 1500, 1000, 300, 24], 2_000_000, [1, 1, 1, 0, 0]),
    "order4": ([3000, 2000, 500, 100], 1_500_000, [1, 1, 0, 0]),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def tensor(request):
    ndims, nnz, dist = CASES[request.param]
    X = COOTensor.synthetic(ndims, nnz, dist=dist, zipf_s=[1.1] * len(ndims), permute=[0] * len(ndims), seed=17)
    inds, vals = X.inds(), X.values()
    H = X.toHiCOO(7, 14, 14)
    X.free()
    yield request.param, ndims, inds, vals, DevHiCOO(H)
    H.free()


@pytest.mark.parametrize("R", [16, 20, 64, 96, 18])
def test_resident_factors_fp32_and_bf16_match_fp64(tensor, R):
    name, ndims, inds, vals, dev = tensor
    rng = np.random.default_rng(R)
    facs = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in ndims]
    dfacs = [torch.from_numpy(f).cuda() for f in facs]
    bf = [f.to(torch.bfloat16) for f in dfacs]
    rounded = [b.float().cpu().numpy() for b in bf]
    for mode in range(len(ndims)):
        e32 = rel_fro(dev.mttkrp(dfacs, mode).cpu().numpy()[:, :R], O.mttkrp(inds, vals, facs, mode, kind="f64"))
        e16 = rel_fro(dev.mttkrp_bf16(bf, mode).cpu().numpy()[:, :R], O.mttkrp(inds, vals, rounded, mode, kind="f64"))
        print("%s R=%d mode %d: FP32 %.3e BF16 %.3e" % (name, R, mode, e32, e16), flush=True)
        assert e32 <= TOL and e16 <= TOL
