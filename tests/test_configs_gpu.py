"""GPU parity at the benchmark configurations' shapes.

- Plain COO MTTKRP at full size: order-3 1000^3 tensor, 1M unique uniform
  nonzeros, values and rank-16 factors U(0,1), all modes, against the FP64
  oracle and against the oracle's restatement of the reference's sequential
  FP32 loop (src/sptensor/mttkrp.c:58-91), each <= 1e-5.
- Precision form at its real dimensions (10M x 50k x 2k x 100, Zipf(1.2) on
  modes 0-1, uniform on modes 2-3, lognormal values, R = 64, HiCOO sb=7
  sk=14) with 2M nonzeros instead of 200M so that the FP64 oracle finishes in
  seconds: FP32 HiCOO MTTKRP and BF16-factor MTTKRP per mode, errors against
  the FP64 oracle (the BF16 form against the oracle on the same bf16-rounded
  factors, and its rounding error against the unrounded FP64 result stated).
- Differentiation form at its real dimensions (500k x 500k x 10k, uniform
  indices, rank-8 planted values + N(0, 0.01) noise, R = 32) with 1M observed
  entries: FP64 central finite differences on 1000 sampled factor entries.
  The loss is quadratic in any single factor entry, so the central difference
  is exact up to rounding; each sample only needs the nonzeros in its slice.
"""
import numpy as np
import pytest
import torch

from conftest import rel_fro
from hiparti_amd import COOTensor, DenseMatrix
from hiparti_amd.completion import completion_residuals, cp_completion_loss
from hiparti_amd.device import DevCOO, DevHiCOO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu(gpu_available):
    if not gpu_available:
        pytest.fail("no GPU visible: the GPU tests must run on the GPU box")


def uniform_factor(rows, R, seed):
    m = DenseMatrix(rows, R)
    m.uniform(seed)
    return m.numpy()


def test_plain_coo_mttkrp_full_size(oracle):
    ndims = [1000, 1000, 1000]
    X = COOTensor.synthetic(ndims, 1_000_000, seed=11)
    inds, vals = X.inds(), X.values()
    assert X.nnz() == 1_000_000
    facs = [uniform_factor(d, 16, 100 + m) for m, d in enumerate(ndims)]
    for mode in range(3):
        mats = [DenseMatrix.from_numpy(f) for f in facs] + [DenseMatrix(1000, 16)]
        out = X.MTTKRP(mode, mats)
        f64 = oracle.mttkrp(inds, vals, facs, mode, kind="f64")
        seq = oracle.mttkrp(inds, vals, facs, mode, kind="seq")
        err_gpu = rel_fro(out, f64)
        err_ref = rel_fro(seq, f64)
        print("mode %d: |gpu - fp64| %.2e, |ref fp32 - fp64| %.2e" % (mode, err_gpu, err_ref))
        assert err_gpu <= 1e-5
        assert rel_fro(out, seq) <= 1e-5


def bf16_round(a):
    return torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()


def test_precision_form_dims_fp32_and_bf16(oracle):
    ndims = [10_000_000, 50_000, 2_000, 100]
    R = 64
    X = COOTensor.synthetic(ndims, 2_000_000, dist=[1, 1, 0, 0], zipf_s=[1.2, 1.2, 1.0, 1.0], lognormal=True,
                            seed=21)
    H = X.toHiCOO(7, 14, 14)
    inds, vals = X.inds(), X.values()     # the COO was sorted in place by the conversion
    D = DevHiCOO(H)
    facs = [uniform_factor(d, R, 200 + m) for m, d in enumerate(ndims)]
    tf = [torch.from_numpy(f).cuda() for f in facs]
    tb = [t.to(torch.bfloat16) for t in tf]
    rounded = [bf16_round(f) for f in facs]
    for mode in range(4):
        f64 = oracle.mttkrp(inds, vals, facs, mode, kind="f64")
        out = D.mttkrp(tf, mode).cpu().numpy()
        e32 = rel_fro(out, f64)
        outb = D.mttkrp_bf16([t if m != mode else None for m, t in enumerate(tb)], mode).cpu().numpy()
        f64r = oracle.mttkrp(inds, vals, rounded, mode, kind="f64")
        eb_same = rel_fro(outb, f64r)
        eb_round = rel_fro(outb, f64)
        print("mode %d: fp32 %.2e | bf16 vs fp64(rounded factors) %.2e | bf16 vs fp64 %.2e" %
              (mode, e32, eb_same, eb_round))
        assert e32 <= 1e-5
        assert eb_same <= 1e-5
        assert eb_round <= 1e-2      # BF16 keeps 8 significant bits per factor entry


def test_differentiation_form_dims_fd_1000_entries():
    ndims = [500_000, 500_000, 10_000]
    nnz = 1_000_000
    R = 32
    rng = np.random.default_rng(31)
    X0 = COOTensor.synthetic(ndims, nnz, seed=31)
    inds = X0.inds().astype(np.int64)
    P = [rng.uniform(0, 1, (d, 8)) for d in ndims]
    planted = np.ones((nnz, 8))
    for m in range(3):
        planted *= P[m][inds[m]]
    vals = (planted.sum(axis=1) + rng.normal(0, 0.01, nnz)).astype(np.float32)
    X = COOTensor.from_arrays(ndims, inds.astype(np.uint32), vals)
    dev = DevCOO(X)
    facs = [(0.5 * rng.uniform(0, 1, (d, R))).astype(np.float32) for d in ndims]
    params = [torch.from_numpy(f).cuda().requires_grad_(True) for f in facs]
    loss = cp_completion_loss(dev, params)
    loss.backward()
    grads = [p.grad.cpu().numpy().astype(np.float64) for p in params]
    f64 = [f.astype(np.float64) for f in facs]
    x64 = vals.astype(np.float64)
    # Loss and residuals against FP64.
    pred = np.ones((nnz, R))
    for m in range(3):
        pred *= f64[m][inds[m]]
    e64 = x64 - pred.sum(axis=1)
    assert abs(loss.item() - 0.5 * float(e64 @ e64)) <= 1e-5 * 0.5 * float(e64 @ e64)
    _, resid = completion_residuals(dev, [p.detach() for p in params])
    assert rel_fro(resid.cpu().numpy(), e64) <= 1e-5
    # 1000 sampled entries: (mode, row with observed entries, rank column).
    order = [np.argsort(inds[m], kind="stable") for m in range(3)]
    sorted_rows = [inds[m][order[m]] for m in range(3)]
    scale = np.sqrt(np.mean(np.concatenate([g.ravel() ** 2 for g in grads])))
    h = 1e-3
    worst = 0.0
    for s in range(1000):
        n = int(rng.integers(0, 3))
        z = order[n][int(rng.integers(0, nnz))]
        i = int(inds[n][z])
        r = int(rng.integers(0, R))
        lo = np.searchsorted(sorted_rows[n], i, "left")
        hi = np.searchsorted(sorted_rows[n], i, "right")
        sub = order[n][lo:hi]
        others = np.ones((len(sub), R))
        for m in range(3):
            if m != n:
                others *= f64[m][inds[m][sub]]
        base = (others * f64[n][i]).sum(axis=1)
        def slice_loss(delta):
            e = x64[sub] - (base + delta * others[:, r])
            return 0.5 * float(e @ e)
        fd = (slice_loss(h) - slice_loss(-h)) / (2 * h)
        err = abs(fd - grads[n][i, r])
        worst = max(worst, err / scale)
        assert err <= 1e-4 * max(scale, abs(fd)), (s, n, i, r, fd, grads[n][i, r])
    print("worst |fd - grad| / rms(grad) over 1000 entries: %.2e" % worst)
