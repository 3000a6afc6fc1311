"""torch.autograd binding of the GPU objective (nomad_obj_eval).

    obj = lib.Objective(train, k, storage="fp64")
    W = torch.randn(m, obj.ldf, dtype=torch.float64, device="cuda", requires_grad=True)
    H = ...
    L = factor_objective(W, H, obj, reg)      # 0-d tensor on the device
    L.backward()                               # W.grad, H.grad from the same kernels

The forward pass computes the loss and both gradients in one call on the
tensors' current stream; backward scales the saved gradients.
"""
from __future__ import annotations

import torch

from . import lib

_DTYPES = {"fp64": torch.float64, "fp32": torch.float32}


class _FactorObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, W, H, obj, reg):
        gW = torch.empty_like(W)
        gH = torch.empty_like(H)
        stream = torch.cuda.current_stream(W.device).cuda_stream
        loss = obj.eval(W.data_ptr(), H.data_ptr(), float(reg), gW.data_ptr(), gH.data_ptr(), stream)
        ctx.save_for_backward(gW, gH)
        return torch.tensor(loss, dtype=W.dtype, device=W.device)

    @staticmethod
    def backward(ctx, g):
        gW, gH = ctx.saved_tensors
        return g * gW, g * gH, None, None


def factor_objective(W: torch.Tensor, H: torch.Tensor, obj: lib.Objective, reg: float) -> torch.Tensor:
    """L = sum (w_i.h_j - r)^2 + reg (sum deg(i)|w_i|^2 + sum deg(j)|h_j|^2) with
    W (m x ldf), H (n x ldf) contiguous device tensors of the objective's
    storage type (padding columns zero)."""
    for name, t, rows in (("W", W, obj.m), ("H", H, obj.n)):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous device tensor")
        if t.shape != (rows, obj.ldf):
            raise ValueError(f"{name} must have shape ({rows}, {obj.ldf}), got {tuple(t.shape)}")
        if t.dtype != _DTYPES[obj.storage]:
            raise ValueError(f"{name} must be {_DTYPES[obj.storage]}")
    return _FactorObjective.apply(W, H, obj, reg)
