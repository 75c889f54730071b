"""CP-completion loss as a torch.autograd.Function.

Model: a rank-R CP tensor [[A_0, ..., A_{N-1}]] fitted to the observed
entries of a sparse tensor X (its nonzeros).  The loss is

    L = 1/2 * sum_z (x_z - sum_r prod_m A_m[i_{m,z}, r])^2

computed in FP64 from FP32 residuals.  Both directions run in libHiParTI's
HIP kernels:

  forward   ptiDevCOOCompletionForward: one pass over the nonzeros that
            gathers the N factor rows of each entry, forms the rank sum, and
            writes the residual e_z = x_z - y_z and the fixed-order FP64 sum
            of e_z^2 / 2.
  backward  dL/dA_n = -MTTKRP(E, n), where E is X's sparsity pattern carrying
            the residuals as values (reference ptiMTTKRP,
            src/sptensor/mttkrp.c:58-91, on that tensor):
            ptiDevCOOCompletionGrad, which for ranks <= 64 on orders <= 5
            recomputes each residual inside the MTTKRP's walk (the output
            row's own factor row staged in LDS) instead of reading the saved
            residuals through the sort permutation, and otherwise runs
            ptiDevMTTKRPCOOValues on the saved residuals.

The gradient is exact for the loss above (the sampled reconstruction is
multilinear in the factors), so gradient checks against FP64 finite
differences only see FP32 rounding.
"""
import ctypes as C

import torch

from ._lib import check, lib
from .device import _check_factor, _check_factors, _check_mode, _check_out, _order, _stream_ptr

_vp = C.c_void_p


class CPCompletionLoss(torch.autograd.Function):
    """loss = CPCompletionLoss.apply(dev_coo, A_0, ..., A_{N-1}).

    dev_coo: a DevCOO holding the observed entries.  A_m: contiguous float32
    CUDA tensors of shape (ndims[m], R), 1 <= R <= 256.  Returns a 0-d
    float64 tensor.

    When every factor needs a gradient and the shape allows it
    (ptiDevCOOCompletionFused: ranks <= 64 on orders <= 5), the forward runs
    ptiDevCOOCompletionLossGrad, which returns the loss and all N gradient
    MTTKRPs from N fused passes, and the backward only scales them by the
    incoming gradient.  Otherwise the forward keeps the residuals and the
    backward runs one ptiDevCOOCompletionGrad per needed mode."""

    @staticmethod
    def forward(ctx, dev, *facs):
        N = dev.nmodes
        if len(facs) != N:
            raise ValueError("expected %d factor matrices, got %d" % (N, len(facs)))
        R = facs[0].shape[1]
        for m, f in enumerate(facs):
            _check_factor(f, torch.float32, "factor %d" % m)
            if f.shape[1] != R or f.shape[0] < dev.ndims[m]:
                raise ValueError("factor %d must have shape (>= %d, %d)" % (m, dev.ndims[m], R))
        device = facs[0].device
        loss = torch.empty((), dtype=torch.float64, device=device)
        ptrs = (_vp * N)(*[f.data_ptr() for f in facs])
        ctx.dev = dev
        ctx.grads = None
        if all(ctx.needs_input_grad[1:]) and lib().ptiDevCOOCompletionFused(dev.handle, R, R) == 1:
            grads = [_grad_buffer(f, dev.ndims[m]) for m, f in enumerate(facs)]
            gptrs = (_vp * N)(*[g.data_ptr() for g in grads])
            check(lib().ptiDevCOOCompletionLossGrad(dev.handle, ptrs, gptrs, R, R, _vp(loss.data_ptr()),
                                                    _stream_ptr()), "ptiDevCOOCompletionLossGrad")
            ctx.grads = grads
            return loss
        resid = torch.empty(dev.nnz, dtype=torch.float32, device=device)
        check(lib().ptiDevCOOCompletionForward(dev.handle, ptrs, R, R, _vp(resid.data_ptr()),
                                               _vp(loss.data_ptr()), _stream_ptr()),
              "ptiDevCOOCompletionForward")
        ctx.save_for_backward(resid, *facs)
        return loss

    @staticmethod
    def backward(ctx, gloss):
        scale = -gloss.to(torch.float32)
        if ctx.grads is not None:
            return (None,) + tuple(g * scale for g in ctx.grads)
        dev = ctx.dev
        resid, *facs = ctx.saved_tensors
        N = dev.nmodes
        R = facs[0].shape[1]
        grads = [None]
        for n in range(N):
            if not ctx.needs_input_grad[n + 1]:
                grads.append(None)
                continue
            out = _grad_buffer(facs[n], dev.ndims[n])
            ptrs = (_vp * (N + 1))(*[f.data_ptr() for f in facs], out.data_ptr())
            check(lib().ptiDevCOOCompletionGrad(dev.handle, _vp(resid.data_ptr()), ptrs, R, R, n, _stream_ptr()),
                  "ptiDevCOOCompletionGrad")
            grads.append(out.mul_(scale))
        return tuple(grads)


def _grad_buffer(f, nrows):
    """Gradient storage shaped like factor f; the kernels write its first
    nrows rows, rows past the tensor's extent are zero."""
    if f.shape[0] > nrows:
        return torch.zeros(f.shape, dtype=torch.float32, device=f.device)
    return torch.empty(f.shape, dtype=torch.float32, device=f.device)


def cp_completion_loss(dev, facs):
    """Functional form: the FP64 CP-completion loss of factors `facs` on the
    observed entries held by `dev` (a DevCOO), differentiable in `facs`."""
    return CPCompletionLoss.apply(dev, *facs)


def completion_residuals(dev, facs):
    """Forward only: (loss, residuals) with residuals in the tensor's storage
    order.  No autograd graph is recorded."""
    N = dev.nmodes
    R = _check_factors(facs, dev.ndims, torch.float32, None)
    resid = torch.empty(dev.nnz, dtype=torch.float32, device=facs[0].device)
    loss = torch.empty((), dtype=torch.float64, device=facs[0].device)
    ptrs = (_vp * N)(*[f.data_ptr() for f in facs])
    check(lib().ptiDevCOOCompletionForward(dev.handle, ptrs, R, R, _vp(resid.data_ptr()), _vp(loss.data_ptr()),
                                           _stream_ptr()), "ptiDevCOOCompletionForward")
    return loss, resid


def mttkrp_values(dev, values, facs, mode, out=None):
    """MTTKRP of X's sparsity pattern carrying `values` (nnz float32, storage
    order) instead of X's own values."""
    N = dev.nmodes
    _check_mode(mode, N)
    R = _check_factors(facs, dev.ndims, torch.float32, None, skip=mode)
    if not (values.is_cuda and values.dtype == torch.float32 and values.is_contiguous()
            and values.numel() == dev.nnz):
        raise ValueError("values must be a contiguous float32 CUDA tensor of nnz elements")
    if out is None:
        out = torch.empty((dev.ndims[mode], R), dtype=torch.float32, device=facs[0].device)
    _check_out(out, dev.ndims[mode], R)
    ptrs = (_vp * (N + 1))(*[f.data_ptr() for f in facs], out.data_ptr())
    check(lib().ptiDevMTTKRPCOOValues(dev.handle, _vp(values.data_ptr()), ptrs, R, R, _order(mode, N, None),
                                      mode, _stream_ptr()), "ptiDevMTTKRPCOOValues")
    return out


__all__ = ["CPCompletionLoss", "cp_completion_loss", "completion_residuals", "mttkrp_values"]
