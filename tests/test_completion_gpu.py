"""GPU checks of the CP-completion loss (hiparti_amd.completion).

The loss L = 1/2 sum_z (x_z - sum_r prod_m A_m[i_mz, r])^2 over the observed
entries and its gradient dL/dA_n = -MTTKRP(residuals, n) are checked:
  - loss and residuals against an FP64 numpy evaluation of the same formula
    (relative error <= 1e-5, the FP32 tolerance);
  - the gradient against FP64 central finite differences of that numpy loss,
    along random directions and at single entries (a wrong sign, a dropped
    1/2 or a transposed index fails these);
  - the gradient against the oracle's FP64 COO MTTKRP applied to the GPU
    residuals, i.e. ptiMTTKRP (reference src/sptensor/mttkrp.c:58-91) on the
    residual tensor;
  - ptiDevMTTKRPCOOValues fed the tensor's own values against
    ptiDevMTTKRPCOO and the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from conftest import rel_fro
from hiparti_amd import COOTensor
from hiparti_amd.completion import CPCompletionLoss, completion_residuals, cp_completion_loss, mttkrp_values
from hiparti_amd.device import DevCOO

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def _need_gpu(gpu_available):
    if not gpu_available:
        pytest.fail("no GPU visible: the GPU tests must run on the GPU box")


def random_tensor(ndims, nnz, seed):
    rng = np.random.default_rng(seed)
    keys = set()
    inds = []
    while len(inds) < nnz:
        t = tuple(int(rng.integers(0, d)) for d in ndims)
        if t not in keys:
            keys.add(t)
            inds.append(t)
    inds = np.array(inds, np.uint32).T.copy() if nnz else np.zeros((len(ndims), 0), np.uint32)
    vals = rng.uniform(-1, 1, nnz).astype(np.float32)
    return inds, vals


def random_factors(ndims, R, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return [(scale * rng.uniform(-1, 1, (d, R))).astype(np.float32) for d in ndims]


def np_resid(inds, vals, facs):
    """FP64 residuals x_z - sum_r prod_m A_m[i_mz, r]."""
    prod = np.ones((inds.shape[1], facs[0].shape[1]), np.float64)
    for m, f in enumerate(facs):
        prod *= f.astype(np.float64)[inds[m]]
    return vals.astype(np.float64) - prod.sum(axis=1)


def np_loss(inds, vals, facs):
    e = np_resid(inds, vals, facs)
    return 0.5 * float(e @ e)


def setup(ndims, nnz, R, seed, scale=1.0):
    inds, vals = random_tensor(ndims, nnz, seed)
    facs = random_factors(ndims, R, seed + 1, scale)
    coo = COOTensor.from_arrays(ndims, inds, vals)
    dev = DevCOO(coo)
    tfacs = [torch.from_numpy(f).cuda() for f in facs]
    return coo, dev, inds, vals, facs, tfacs


SHAPES = [
    ((40, 50, 60), 3000, 16),       # rank a multiple of 4: float4 lanes
    ((30, 20, 25, 15), 2000, 10),   # rank 2 mod 4: float2 lanes
    ((17, 300), 1500, 7),           # order 2, odd rank: scalar lanes
    ((9, 8, 7, 6, 5), 4000, 33),    # order 5
    ((200, 150, 100), 70000, 64),   # many chunks, ragged last chunk
    ((12, 11, 10), 500, 256),       # largest rank the forward kernel takes
    ((60, 50, 40), 5000, 65),       # odd ranks above 64: lane groups step over column blocks
    ((60, 50, 40), 5000, 75),
    ((60, 50, 40), 5000, 130),      # 2 mod 4 above 128
]


@pytest.mark.parametrize("ndims,nnz,R", SHAPES)
def test_loss_and_residuals_match_fp64(ndims, nnz, R):
    coo, dev, inds, vals, facs, tfacs = setup(ndims, nnz, R, seed=7, scale=0.6)
    loss, resid = completion_residuals(dev, tfacs)
    e64 = np_resid(inds, vals, facs)
    assert rel_fro(resid.cpu().numpy(), e64) <= TOL
    ref = 0.5 * float(e64 @ e64)
    assert abs(loss.item() - ref) <= TOL * ref


@pytest.mark.parametrize("ndims,nnz,R", SHAPES[:4])
def test_gradient_matches_fp64_finite_differences(ndims, nnz, R):
    coo, dev, inds, vals, facs, tfacs = setup(ndims, nnz, R, seed=11, scale=0.6)
    params = [t.clone().requires_grad_(True) for t in tfacs]
    loss = cp_completion_loss(dev, params)
    loss.backward()
    grads = [p.grad.cpu().numpy().astype(np.float64) for p in params]
    f64 = [f.astype(np.float64) for f in facs]
    rng = np.random.default_rng(3)
    h = 1e-4
    gscale = max(np.abs(g).max() for g in grads)
    # Directional derivatives, one random direction per factor and one over all.
    dirs = []
    for n in range(len(facs)):
        d = [np.zeros_like(f) for f in f64]
        d[n] = rng.standard_normal(f64[n].shape)
        dirs.append(d)
    dirs.append([rng.standard_normal(f.shape) for f in f64])
    for d in dirs:
        plus = np_loss(inds, vals, [f + h * x for f, x in zip(f64, d)])
        minus = np_loss(inds, vals, [f - h * x for f, x in zip(f64, d)])
        fd = (plus - minus) / (2 * h)
        an = sum(float((g * x).sum()) for g, x in zip(grads, d))
        assert abs(fd - an) <= 1e-4 * max(abs(fd), 1.0), (fd, an)
    # Single entries: the touched row/column must be right, not just the sum.
    for n in range(len(facs)):
        rows = np.unique(inds[n])[:3]
        for i in rows:
            for r in (0, R - 1):
                d = [np.zeros_like(f) for f in f64]
                d[n][i, r] = 1.0
                plus = np_loss(inds, vals, [f + h * x for f, x in zip(f64, d)])
                minus = np_loss(inds, vals, [f - h * x for f, x in zip(f64, d)])
                fd = (plus - minus) / (2 * h)
                assert abs(fd - grads[n][i, r]) <= 1e-4 * max(gscale, 1.0), (n, i, r, fd, grads[n][i, r])


@pytest.mark.parametrize("ndims,nnz,R", SHAPES)
def test_gradient_is_negative_mttkrp_of_residuals(ndims, nnz, R, oracle):
    coo, dev, inds, vals, facs, tfacs = setup(ndims, nnz, R, seed=5, scale=0.6)
    params = [t.clone().requires_grad_(True) for t in tfacs]
    loss = CPCompletionLoss.apply(dev, *params)
    (2.5 * loss).backward()   # the incoming gradient scales the result
    _, resid = completion_residuals(dev, tfacs)
    e = resid.cpu().numpy()
    for n in range(len(facs)):
        ref = -2.5 * oracle.mttkrp(inds, e, facs, n, kind="f64")
        assert rel_fro(params[n].grad.cpu().numpy(), ref) <= TOL, n


def test_values_override_with_own_values_matches_plain_mttkrp(oracle):
    # Rows that cross a chunk boundary are accumulated with atomics, so the
    # two calls agree to rounding, not bit for bit.
    coo, dev, inds, vals, facs, tfacs = setup((120, 90, 70), 20000, 32, seed=9)
    tv = torch.from_numpy(vals).cuda()
    for n in range(3):
        a = mttkrp_values(dev, tv, tfacs, n).cpu().numpy()
        b = dev.mttkrp(tfacs, n).cpu().numpy()
        ref = oracle.mttkrp(inds, vals, facs, n, kind="f64")
        assert rel_fro(a, b) <= 1e-6, n
        assert rel_fro(a, ref) <= TOL, n


def test_replaced_and_own_values_alternate(oracle):
    """Replaced-value MTTKRPs (the segmented-scan COO kernel reading the
    values through the sort permutation) alternate with own-value ones (the
    stream kernel on the same tensor): random, then doubled own values, on
    every mode, each against the FP64 oracle."""
    coo, dev, inds, vals, facs, tfacs = setup((300, 2000, 1500), 40000, 32, seed=13)
    rng = np.random.default_rng(14)
    other = rng.standard_normal(len(vals)).astype(np.float32)
    for n in range(3):
        ref_own = oracle.mttkrp(inds, vals, facs, n, kind="f64")
        got = mttkrp_values(dev, torch.from_numpy(other).cuda(), tfacs, n).cpu().numpy()
        assert rel_fro(got, oracle.mttkrp(inds, other, facs, n, kind="f64")) <= TOL, n
        assert rel_fro(dev.mttkrp(tfacs, n).cpu().numpy(), ref_own) <= TOL, n
        got = mttkrp_values(dev, torch.from_numpy(2 * vals).cuda(), tfacs, n).cpu().numpy()
        assert rel_fro(got, 2 * ref_own) <= TOL, n
        assert rel_fro(dev.mttkrp(tfacs, n).cpu().numpy(), ref_own) <= TOL, n


@pytest.mark.parametrize("ndims,nnz,R", [((200, 150, 100), 70000, 64), ((40, 50, 60), 3000, 16),
                                         ((30, 20, 25, 15), 2000, 10), ((9, 8, 7, 6, 5), 4000, 33)])
def test_fused_and_residual_paths_agree(ndims, nnz, R):
    """With every factor requiring a gradient the loss and gradients come
    from the fused passes (ranks <= 64 on orders <= 5; 8-, 4- and 16-lane
    groups here), with one factor frozen from the forward kernel plus one
    gradient MTTKRP per mode; both against the FP64 loss, and the gradients
    of the shared factors against each other.  Order 5 at rank 33 has no
    fused form and takes the residual path both times."""
    coo, dev, inds, vals, facs, tfacs = setup(ndims, nnz, R, seed=17, scale=0.6)
    ref = np_loss(inds, vals, facs)
    full = [t.clone().requires_grad_(True) for t in tfacs]
    loss_f = cp_completion_loss(dev, full)
    loss_f.backward()
    part = [t.clone().requires_grad_(i != 0) for i, t in enumerate(tfacs)]
    loss_p = cp_completion_loss(dev, part)
    loss_p.backward()
    assert abs(loss_f.item() - ref) <= TOL * ref
    assert abs(loss_p.item() - ref) <= TOL * ref
    for n in range(1, len(ndims)):
        assert rel_fro(full[n].grad.cpu().numpy(), part[n].grad.cpu().numpy()) <= TOL, n


def test_partial_gradients_and_descent():
    coo, dev, inds, vals, facs, tfacs = setup((50, 40, 30), 6000, 8, seed=21, scale=0.5)
    params = [t.clone().requires_grad_(i != 1) for i, t in enumerate(tfacs)]
    loss = cp_completion_loss(dev, params)
    loss.backward()
    assert params[1].grad is None
    assert params[0].grad is not None and params[2].grad is not None
    # Plain gradient descent on all factors lowers the loss monotonically
    # for a small enough step.
    params = [t.clone().requires_grad_(True) for t in tfacs]
    opt = torch.optim.SGD(params, lr=2e-3)
    losses = []
    for _ in range(20):
        opt.zero_grad()
        loss = cp_completion_loss(dev, params)
        loss.backward()
        opt.step()
        losses.append(loss.item())
    assert all(b < a for a, b in zip(losses, losses[1:])), losses


def test_empty_tensor_has_zero_loss_and_gradient():
    coo, dev, inds, vals, facs, tfacs = setup((5, 6, 7), 0, 4, seed=1)
    params = [t.clone().requires_grad_(True) for t in tfacs]
    loss = cp_completion_loss(dev, params)
    loss.backward()
    assert loss.item() == 0.0
    for p in params:
        assert torch.count_nonzero(p.grad).item() == 0


def test_rank_above_256_is_rejected():
    coo, dev, inds, vals, facs, tfacs = setup((5, 6, 7), 20, 260, seed=1)
    with pytest.raises(Exception):
        completion_residuals(dev, tfacs)
