"""torch.autograd binding of the GPU objective (nomad_obj_eval_host).

    obj = lib.Objective(train, k, storage="fp64")
    W = torch.rand(m, k, dtype=torch.float64, requires_grad=True)   # any device
    H = torch.rand(n, k, dtype=torch.float64, requires_grad=True)
    L = factor_objective(W, H, obj, reg)
    L.backward()                                   # W.grad, H.grad from the same kernels

The forward pass computes the loss and both gradients with one call of the
hand-written kernels; backward scales the saved gradients.  torch wheels
bundle their own HIP runtime, distinct from the one this library links, so
tensors are handed over through host memory rather than by device address.
"""
from __future__ import annotations

import torch

from . import lib


class _FactorObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, W, H, obj, reg):
        loss, gW, gH = obj.eval_host(W.detach().double().cpu().numpy(), H.detach().double().cpu().numpy(),
                                     float(reg))
        ctx.save_for_backward(torch.from_numpy(gW).to(W.device, W.dtype), torch.from_numpy(gH).to(H.device, H.dtype))
        return torch.tensor(loss, dtype=W.dtype, device=W.device)

    @staticmethod
    def backward(ctx, g):
        gW, gH = ctx.saved_tensors
        return g * gW, g * gH, None, None


def factor_objective(W: torch.Tensor, H: torch.Tensor, obj: lib.Objective, reg: float) -> torch.Tensor:
    """L = sum (w_i.h_j - r)^2 + reg (sum deg(i)|w_i|^2 + sum deg(j)|h_j|^2) for
    W (m x k), H (n x k)."""
    if W.shape != (obj.m, obj.k) or H.shape != (obj.n, obj.k):
        raise ValueError(f"W must be ({obj.m}, {obj.k}) and H ({obj.n}, {obj.k})")
    return _FactorObjective.apply(W, H, obj, reg)
