"""Fused loss head of the training step (reference: train.py:61-62 with
utils.py:35-36).

``power_cross_entropy(y, labels)`` equals
``torch.nn.functional.cross_entropy(normalize_power(y.sum(dim=1)), labels)``
-- the reference's criterion on the probe series -- computed by three short
HIP launches (``wt_loss_power_ce``: chunked partial sums over time, the
per-sample softmax and gradient, the batch mean) instead of a chain of torch
reductions.
Its backward hands the time loop dJ/dy as a (B,T,P) view of the (B,P)
gradient of the integrated powers, expanded over T.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


class _PowerCE(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, labels):
        if y.dim() != 3:
            raise ValueError(f"power_cross_entropy: y must be (B, T, P), got {tuple(y.shape)}")
        if not y.is_cuda or y.dtype != torch.float32:
            raise RuntimeError("power_cross_entropy: y must be a float32 tensor on the HIP device")
        B, T, P = y.shape
        if tuple(labels.shape) != (B,):
            raise ValueError(f"power_cross_entropy: labels must be ({B},), got "
                             f"{tuple(labels.shape)}")
        y = y.contiguous()
        lab = labels.to(device=y.device, dtype=torch.int64).contiguous()
        loss = torch.empty((), dtype=torch.float32, device=y.device)
        gI = torch.empty((B, P), dtype=torch.float32, device=y.device)
        L = _lib.lib()
        nbytes = ctypes.c_int64()
        _lib.check(L.wt_loss_power_ce_workspace_bytes(B, T, P, ctypes.byref(nbytes)),
                   "wt_loss_power_ce_workspace_bytes")
        work = torch.empty((nbytes.value + 7) // 8, dtype=torch.float64, device=y.device)
        _lib.check(L.wt_loss_power_ce(
            y.data_ptr(), lab.data_ptr(), B, T, P, loss.data_ptr(), gI.data_ptr(),
            work.data_ptr(), _lib.stream_ptr(y.device)), "wt_loss_power_ce")
        ctx.save_for_backward(gI)
        ctx.T = T
        return loss

    @staticmethod
    def backward(ctx, g):
        gI, = ctx.saved_tensors
        return (g * gI).unsqueeze(1).expand(-1, ctx.T, -1), None


def power_cross_entropy(y: torch.Tensor, labels: torch.Tensor) -> torch.Tensor:
    """Cross-entropy of the normalised integrated probe powers (train.py:61-62,
    ``criterion = torch.nn.CrossEntropyLoss()``): y (B, T, P) probe series,
    labels (B,) class indices.  A label outside [0, P) gives a NaN loss (the
    reference raises; the check would need a device-to-host read)."""
    return _PowerCE.apply(y, labels)


def is_default_cross_entropy(criterion) -> bool:
    """True when `criterion` is a torch.nn.CrossEntropyLoss with the settings
    power_cross_entropy reproduces (no weights, mean reduction, no label
    smoothing)."""
    return (isinstance(criterion, torch.nn.CrossEntropyLoss) and criterion.weight is None
            and criterion.reduction == "mean" and criterion.label_smoothing == 0.0)
