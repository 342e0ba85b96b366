"""The 5-point Laplacian of operators.py:5-11.

``_laplacian(y, h)`` keeps the reference's name and signature: ``y`` is
(B, Nx, Ny), ``h`` the grid step (a number or a 0-dim tensor); the result is
h^-2 times the 5-point stencil with zero padding, on the HIP device:
``wt_laplacian`` (float32), which accumulates in conv2d's order and so equals
the reference's fp32 result bit for bit, or ``wt_laplacian_f64`` (float64,
after set_dtype('float64')).  The operator is its own adjoint (the
stencil is symmetric and the padding zero), so the backward is the same
kernel on the output gradient.  As in the reference's use of it (``h`` is a
geometry buffer), no gradient flows to ``h``.
"""
import torch

from . import _lib


def _apply(y: torch.Tensor, h: float) -> torch.Tensor:
    y = y.contiguous()
    B, Nx, Ny = y.shape
    out = torch.empty_like(y)
    name = "wt_laplacian_f64" if y.dtype == torch.float64 else "wt_laplacian"
    _lib.check(getattr(_lib.lib(), name)(y.data_ptr(), out.data_ptr(), B, Nx, Ny, h,
                                         _lib.stream_ptr(y.device)), name)
    return out


class _Laplacian(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, h):
        ctx.h = h
        ctx.dtype = y.dtype
        return _apply(y, h)

    @staticmethod
    def backward(ctx, g):
        return _apply(g.to(ctx.dtype), ctx.h), None


def _laplacian(y: torch.Tensor, h) -> torch.Tensor:
    """operators.py:5-11."""
    if y.dim() != 3:
        raise ValueError(f"wavetorch_amd: _laplacian takes (B, Nx, Ny) fields, got {tuple(y.shape)}")
    if not y.is_cuda:
        raise RuntimeError("wavetorch_amd: _laplacian runs on the HIP device only; move the "
                           "tensor with .to('cuda')")
    if y.dtype not in (torch.float32, torch.float64):
        raise NotImplementedError(f"wavetorch_amd: _laplacian computes in float32 or float64 "
                                  f"(got {y.dtype})")
    return _Laplacian.apply(y, float(h))
