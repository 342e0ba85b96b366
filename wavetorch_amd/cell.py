"""The wave-equation cell (reference: wavetorch/cell.py).

``TimeStep`` keeps the reference's autograd-Function interface
(``TimeStep.apply(b, c, y1, y2, dt, h)``, cell.py:20-44) with its forward and
backward running as gfx950 kernels (``wt_step_fwd`` / ``wt_step_bwd``).
``WaveCell`` keeps the reference's buffers, CFL check and single-step
``forward``; ``WaveRNN`` does not call it per step but runs the whole loop
in fused kernels (see rnn.py).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .utils import to_tensor


def saturable_damping(u, uth, b0):
    """cell.py:8-9."""
    return b0 / (1 + torch.abs(u / uth).pow(2))


def _check_dev(*ts):
    for t in ts:
        if not t.is_cuda:
            raise RuntimeError("wavetorch_amd: the time step runs on the HIP device only; "
                               "move the tensors/model with .to('cuda')")
        if t.dtype not in (torch.float32, torch.float64):
            raise NotImplementedError(f"wavetorch_amd: the time step computes in float32 or "
                                      f"float64 (got {t.dtype})")
    if len({t.dtype for t in ts}) != 1:
        raise TypeError("wavetorch_amd: b, c, y1 and y2 of a time step must share one dtype")


class TimeStep(torch.autograd.Function):
    """One leapfrog step y = _time_step(b, c, y1, y2, dt, h) (cell.py:12-44),
    in float32 (wt_step_fwd / wt_step_bwd) or float64 (the _f64 entry
    points)."""

    @staticmethod
    def forward(ctx, b, c, y1, y2, dt, h):
        _check_dev(b, c, y1, y2)
        y1c, y2c = y1.contiguous(), y2.contiguous()
        B = y1c.shape[0]
        Nx, Ny = y1c.shape[-2:]
        shapes = (tuple(b.shape), tuple(c.shape))
        per_sample = 1 if (b.dim() == 3 or c.dim() == 3) else 0
        if per_sample:
            b = b.expand(B, Nx, Ny)
            c = c.expand(B, Nx, Ny)
        bc, cc = b.contiguous(), c.contiguous()
        y = torch.empty_like(y1c)
        dtf, hf = float(dt), float(h)
        name = "wt_step_fwd_f64" if y.dtype == torch.float64 else "wt_step_fwd"
        _lib.check(getattr(_lib.lib(), name)(bc.data_ptr(), cc.data_ptr(), per_sample,
                                             y1c.data_ptr(), y2c.data_ptr(), B, Nx, Ny, dtf, hf,
                                             y.data_ptr(), _lib.stream_ptr(y.device)), name)
        ctx.save_for_backward(bc, cc, y1c, y2c)
        ctx.meta = (per_sample, dtf, hf, shapes)
        return y

    @staticmethod
    def backward(ctx, grad_output):
        b, c, y1, y2 = ctx.saved_tensors
        per_sample, dt, h, shapes = ctx.meta
        g = grad_output.to(y1.dtype).contiguous()
        B = y1.shape[0]
        Nx, Ny = y1.shape[-2:]
        need = ctx.needs_input_grad
        gb = torch.empty_like(b) if need[0] else None
        gc = torch.empty_like(c) if need[1] else None
        gy1 = torch.empty_like(y1) if need[2] else None
        gy2 = torch.empty_like(y2) if need[3] else None
        name = "wt_step_bwd_f64" if y1.dtype == torch.float64 else "wt_step_bwd"
        _lib.check(getattr(_lib.lib(), name)(b.data_ptr(), c.data_ptr(), per_sample,
                                             y1.data_ptr(), y2.data_ptr(), g.data_ptr(), B, Nx,
                                             Ny, dt, h, _lib.ptr(gb), _lib.ptr(gc),
                                             _lib.ptr(gy1), _lib.ptr(gy2),
                                             _lib.stream_ptr(g.device)), name)
        # a shared (Nx,Ny) field broadcast against a per-sample one: reduce
        # over the batch like autograd's sum_to_size
        if gb is not None and gb.dim() > len(shapes[0]):
            gb = gb.sum(0)
        if gc is not None and gc.dim() > len(shapes[1]):
            gc = gc.sum(0)
        return gb, gc, gy1, gy2, None, None


class WaveCell(torch.nn.Module):
    """The recurrent cell implementing the scalar wave equation (cell.py:47-107)."""

    def __init__(self, dt: float, geometry, satdamp_b0: float = 0.0, satdamp_uth: float = 0.0,
                 c_nl: float = 0.0, c_clamp=None):
        super().__init__()
        self.register_buffer("dt", to_tensor(dt))
        self.geom = geometry
        self.register_buffer("satdamp_b0", to_tensor(satdamp_b0))
        self.register_buffer("satdamp_uth", to_tensor(satdamp_uth))
        self.register_buffer("c_nl", to_tensor(c_nl))
        # extension (not in the reference): clamp the nonlinear speed to
        # [cmin, cmax]; None = reference behaviour
        self.c_clamp = None if c_clamp is None else (float(c_clamp[0]), float(c_clamp[1]))
        cmax = self.geom.cmax
        h = self.geom.h
        if dt > 1 / cmax * h / np.sqrt(2):
            raise ValueError(
                'The spatial discretization defined by the geometry `h = %f` and the temporal '
                'discretization defined by the model `dt = %f` do not satisfy the CFL stability '
                'criteria' % (h, dt))

    def parameters(self, recursive=True):
        for param in self.geom.parameters():
            yield param

    def forward(self, h1, h2, c_linear, rho):
        """One step: (h1, h2) = (u_n, u_{n-1}) -> (u_{n+1}, u_n) (cell.py:79-107)."""
        if self.satdamp_b0 > 0:
            b = self.geom.b + rho * saturable_damping(h1, uth=self.satdamp_uth, b0=self.satdamp_b0)
        else:
            b = self.geom.b
        if self.c_nl != 0:
            c = c_linear + rho * self.c_nl * h1.pow(2)
        else:
            c = c_linear
        if self.c_clamp is not None:
            c = torch.clamp(c, self.c_clamp[0], self.c_clamp[1])
        y = TimeStep.apply(b, c, h1, h2, self.dt, self.geom.h)
        return y, h1
