"""GPU test of wt_loss_power_ce_static, the loss head that takes no workspace.

It runs the kernels of wt_loss_power_ce on a static device array of the
library instead of a caller-provided workspace.  Called through the C ABI on
two shapes in a row (the second smaller, so stale partial sums of the first
call would show), its loss and dJ/dI equal wt_loss_power_ce's bit for bit
(same kernels, same order of summation).  A shape larger than the static array
is rejected with WT_ENOMEM before any launch.
"""
import ctypes

import pytest
import torch

from wavetorch_amd import _lib

pytestmark = pytest.mark.gpu

WT_ENOMEM = -5


def _series(B, T, P, seed, dev):
    g = torch.Generator().manual_seed(seed)
    y = (torch.rand((B, T, P), generator=g) * 0.02 + 1e-4).to(dev)
    labels = torch.randint(0, P, (B,), generator=g).to(dev)
    return y.contiguous(), labels.contiguous()


def _with_workspace(L, y, labels, stream):
    B, T, P = y.shape
    nbytes = ctypes.c_int64(0)
    _lib.check(L.wt_loss_power_ce_workspace_bytes(B, T, P, ctypes.byref(nbytes)),
               "wt_loss_power_ce_workspace_bytes")
    work = torch.empty(nbytes.value, dtype=torch.uint8, device=y.device)
    loss = torch.empty(1, dtype=torch.float32, device=y.device)
    gI = torch.empty((B, P), dtype=torch.float32, device=y.device)
    _lib.check(L.wt_loss_power_ce(y.data_ptr(), labels.data_ptr(), B, T, P, loss.data_ptr(),
                                  gI.data_ptr(), work.data_ptr(), stream), "wt_loss_power_ce")
    return loss, gI


def _static(L, y, labels, stream):
    B, T, P = y.shape
    loss = torch.full((1,), float("nan"), dtype=torch.float32, device=y.device)
    gI = torch.full((B, P), float("nan"), dtype=torch.float32, device=y.device)
    _lib.check(L.wt_loss_power_ce_static(y.data_ptr(), labels.data_ptr(), B, T, P,
                                         loss.data_ptr(), gI.data_ptr(), stream),
               "wt_loss_power_ce_static")
    return loss, gI


def test_static_workspace_loss_equals_workspace_loss(cuda):
    L = _lib.lib()
    stream = _lib.stream_ptr(cuda)
    # the training step's shape, then a smaller one with more probes
    for (B, T, P), seed in (((64, 3000, 3), 1), ((5, 37, 7), 2)):
        y, labels = _series(B, T, P, seed, cuda)
        want_loss, want_g = _with_workspace(L, y, labels, stream)
        got_loss, got_g = _static(L, y, labels, stream)
        torch.cuda.synchronize()
        print(f"static loss head B {B} T {T} P {P}: {got_loss.item():.9g}", flush=True)
        assert torch.equal(got_loss, want_loss)
        assert torch.equal(got_g, want_g)
        assert torch.isfinite(got_g).all()


def test_static_workspace_rejects_oversized_shape(cuda):
    L = _lib.lib()
    stream = _lib.stream_ptr(cuda)
    B, T, P = 4096, 2000, 256  # 4096 * 1 * 256 + 4096 doubles > 131072
    y = torch.zeros(1, device=cuda)
    st = L.wt_loss_power_ce_static(y.data_ptr(), y.data_ptr(), B, T, P, y.data_ptr(),
                                   y.data_ptr(), stream)
    assert st == WT_ENOMEM
    buf = ctypes.create_string_buffer(512)
    L.wt_last_error(buf, 512)
    assert b"static workspace" in buf.value
