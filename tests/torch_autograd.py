"""torch.autograd.Function over the engine's C ABI (test wrapper only).

The forward pass runs a batched FP64 system built from (e_site, coupling
scale) and returns the per-trajectory losses; the backward pass returns the
engine's own reverse-sweep gradients (dq_adjoint), contracted with the
incoming loss gradient.  torch only carries the tensors: nothing here
differentiates by itself."""
import numpy as np
import torch


class EngineLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, e_site, g, build, nsteps, ckpt, loss_kw):
        sysobj = build(e_site.detach().cpu().numpy().astype(np.float64), float(g))
        try:
            loss, ge, gg = sysobj.adjoint(nsteps, ckpt=ckpt, **loss_kw)
        finally:
            sysobj.close()
        ctx.save_for_backward(torch.from_numpy(ge), torch.from_numpy(gg))
        return torch.from_numpy(loss)

    @staticmethod
    def backward(ctx, grad_loss):
        ge, gg = ctx.saved_tensors
        grad_e = (grad_loss[:, None] * ge).sum(0)
        grad_g = (grad_loss * gg).sum()
        return grad_e, grad_g, None, None, None, None


def engine_loss(e_site, g, build, nsteps, ckpt, **loss_kw):
    """per-trajectory losses as a differentiable function of e_site[n] and g"""
    return EngineLoss.apply(e_site, g, build, nsteps, ckpt, loss_kw)
