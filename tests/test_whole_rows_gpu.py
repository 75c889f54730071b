"""GPU checks of the HiCOO output write-out with whole rows.

When every rank tile runs the stream kernel and the last tile's lanes reach
the end of the output row, a unit that owns its block-row alone writes all of
that block-row's output rows and columns (zeros and padding included), and
the zero fill before the kernel covers only block-rows that several units
share or that hold no nonzero.  The tensor below has all three kinds in mode
0: a Zipf(1.3) row distribution puts about half of the nonzeros into block-row
0 (split over many units), the other used block-rows have one unit each, and
every other block-row of 128 rows is empty.  Each output is prefilled with
NaN, so any element that neither the fill nor the kernel writes shows up;
columns past R must be exactly zero.  Rank and output widths cover one and
two rank tiles, padding inside the last tile's lanes (whole rows) and an
output wider than the lanes reach (BF16, out width 64 at R = 20: full fill).
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


@pytest.fixture(scope="module")
def tensor():
    rng = np.random.default_rng(12)
    n = 1_500_000
    r = (rng.zipf(1.3, n) - 1) % 2000
    rows = (r // 128) * 256 + r % 128          # block-rows 1, 3, 5, ... stay empty
    ndims = [4096, 2500, 700]
    inds = np.stack([rows, rng.integers(0, ndims[1], n), rng.integers(0, ndims[2], n)]).astype(np.uint32)
    key = (inds[0].astype(np.uint64) * ndims[1] + inds[1]) * ndims[2] + inds[2]
    inds = inds[:, np.unique(key, return_index=True)[1]]
    vals = rng.random(inds.shape[1]).astype(np.float32)
    X = COOTensor.from_arrays(ndims, inds, vals)
    H = X.toHiCOO(7, 10, 10)
    yield ndims, inds, vals, DevHiCOO(H)
    H.free()


def _check(got, ref, R, what):
    assert not torch.isnan(got).any().item(), what
    g = got.cpu().numpy()
    assert rel_fro(g[:, :R], ref) <= TOL, what
    assert np.all(g[:, R:] == 0.0), what


@pytest.mark.parametrize("R,stride", [(32, 32), (20, 24), (96, 104), (16, 16)])
def test_fp32_whole_rows_and_padding(tensor, R, stride):
    ndims, inds, vals, dev = tensor
    rng = np.random.default_rng(R)
    facs = [np.zeros((d, stride), np.float32) for d in ndims]
    for f in facs:
        f[:, :R] = rng.uniform(-1, 1, (f.shape[0], R))
    dfacs = [torch.from_numpy(f).cuda() for f in facs]
    for mode in range(3):
        ref = O.mttkrp(inds, vals, [np.ascontiguousarray(f[:, :R]) for f in facs], mode, kind="f64")
        for rep in range(2):
            out = torch.full((ndims[mode], stride), float("nan"), device="cuda")
            dev.mttkrp(dfacs, mode, R=R, out=out)
            _check(out, ref, R, ("fp32", R, stride, mode, rep))


@pytest.mark.parametrize("R,width", [(32, 32), (20, 24), (20, 64), (64, 64)])
def test_bf16_whole_rows_and_wide_outputs(tensor, R, width):
    ndims, inds, vals, dev = tensor
    rng = np.random.default_rng(R + width)
    stride = (R + 7) // 8 * 8
    bf = []
    for d in ndims:
        f = torch.zeros((d, stride), dtype=torch.float32)
        f[:, :R] = torch.from_numpy(rng.uniform(-1, 1, (d, R)).astype(np.float32))
        bf.append(f.to(torch.bfloat16).cuda())
    rounded = [np.ascontiguousarray(b.float().cpu().numpy()[:, :R]) for b in bf]
    for mode in range(3):
        ref = O.mttkrp(inds, vals, rounded, mode, kind="f64")
        out = torch.full((ndims[mode], width), float("nan"), device="cuda")
        dev.mttkrp_bf16(bf, mode, R=R, out=out)
        _check(out, ref, R, ("bf16", R, width, mode))
