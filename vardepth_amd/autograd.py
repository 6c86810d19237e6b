"""PyTorch autograd over the HIP forward and adjoint kernels.

:class:`PairHMMLogLik` maps (gap, ext, err[64]) to the ln likelihood of every
pair of a :class:`~vardepth_amd.engine.DeviceBatch`.  Its forward is the
batched FP32 + FP64-rescue forward; its backward is the adjoint pass
(``vd_grad_device``) with the incoming gradient as per-pair weights, so one
backward call gives d(sum_p w_p ln L_p)/d(gap, ext, err) for any scalar loss
built on the likelihoods.  PyTorch here is only the autograd frame: the
arithmetic runs in the HIP kernels.

:class:`HmmParameters` holds the three parameter groups as unconstrained
FP32 tensors (gap = 0.5 sigmoid(a), ext = sigmoid(b), err = sigmoid(c)) so a
plain optimiser keeps them inside the model's ranges.
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import DeviceBatch


class PairHMMLogLik(torch.autograd.Function):
    """ln L per pair (float64, device) as a differentiable function of gap, ext, err."""

    @staticmethod
    def forward(ctx, gap, ext, err, dev_batch: DeviceBatch, mode: int = 0, fused: bool = False):
        if err.shape != (64,):
            raise ValueError("err must have 64 quality bins")
        # the parameter tables are built on the host: one copy back for all 66 values
        host = torch.cat([gap.detach().reshape(1).to(torch.float64), ext.detach().reshape(1).to(torch.float64),
                          err.detach().to(torch.float64)]).cpu().numpy()
        g, x, e = float(host[0]), float(host[1]), host[2:].copy()
        ctx.unit_grad = None
        if fused:
            # one adjoint call returns ln L per pair and the unit-weight gradient
            ctx.unit_grad = dev_batch.grad(gap=g, ext=x, err=e, mode=mode)
        else:
            dev_batch.forward(mode=mode, precision="mixed", err=e, gap=g, ext=x)
        out = dev_batch.out[:dev_batch.npairs].clone()
        ctx.dev_batch, ctx.mode, ctx.params = dev_batch, mode, (g, x, e)
        ctx.meta = (gap.dtype, gap.device, ext.dtype, ext.device, err.dtype, err.device)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        g, x, e = ctx.params
        gvec = None
        if ctx.unit_grad is not None and grad_out.numel() > 0:
            # d(sum_p c ln L_p) = c d(sum_p ln L_p): a uniform incoming gradient
            # scales the unit-weight gradient of the forward call
            c = grad_out.reshape(-1)[:1]
            if bool(torch.all(grad_out == c)):
                gvec = ctx.unit_grad * c.to(ctx.unit_grad.dtype)
        if gvec is None:
            gvec = ctx.dev_batch.grad(gap=g, ext=x, err=e, mode=ctx.mode, weights=grad_out)
        gd, gdev, xd, xdev, ed, edev = ctx.meta
        gvec = gvec.to(device=gdev)  # one copy when the parameters live on one device (the host, typically)
        dgap = gvec[0].to(dtype=gd)
        dext = gvec[1].to(dtype=xd, device=xdev)
        derr = gvec[2:].to(dtype=ed, device=edev)
        return dgap, dext, derr, None, None, None


def loglik(gap, ext, err, dev_batch: DeviceBatch, mode: int = 0, fused: bool = False):
    """ln L per pair.  fused: the forward runs the adjoint with unit weights
    (it also returns ln L), so that a backward whose incoming gradient is the
    same for every pair (a loss c * sum ln L, e.g. -mean ln L) costs no second
    pass; any other incoming gradient runs the weighted adjoint as usual."""
    return PairHMMLogLik.apply(gap, ext, err, dev_batch, mode, fused)


def _logit(p):
    p = np.clip(np.asarray(p, dtype=np.float64), 1e-12, 1 - 1e-12)
    return np.log(p) - np.log1p(-p)


class HmmParameters(torch.nn.Module):
    """gap open, gap continuation and the 64-bin error table as FP32 parameters."""

    def __init__(self, gap: float, ext: float, err):
        super().__init__()
        self.a = torch.nn.Parameter(torch.tensor(_logit(2 * gap), dtype=torch.float32))
        self.b = torch.nn.Parameter(torch.tensor(_logit(ext), dtype=torch.float32))
        self.c = torch.nn.Parameter(torch.tensor(_logit(err), dtype=torch.float32))

    def values(self):
        return 0.5 * torch.sigmoid(self.a), torch.sigmoid(self.b), torch.sigmoid(self.c)

    def forward(self, dev_batch: DeviceBatch, mode: int = 0, fused: bool = False):
        gap, ext, err = self.values()
        return loglik(gap, ext, err, dev_batch, mode, fused)
