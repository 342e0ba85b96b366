"""GPU tests of wt_autotune, the library's own tiling search behind the C ABI
(reached through autotune.tune_native).

* The search returns a tuning the library accepts, with a finite time no
  worse than the heuristic plan it starts from (it only accepts faster
  candidates).
* Parity across the search: the forward with the tuned plan equals the
  heuristic plan's bit for bit (tilings change the order of work, not the
  arithmetic), and dJ/dc, dJ/dx agree within 1e-5 (the adjoint's per-launch
  partial sums follow the block depth).
* A second call on the same shape returns the cached result.
"""
import ctypes

import pytest
import torch

from tests.helpers import rel
from wavetorch_amd import (WaveCell, WaveGeometryFreeForm, WaveIntensityProbe, WaveRNN,
                           WaveSource, _lib, autotune)
from wavetorch_amd._loop import Plan, WaveLoop

pytestmark = pytest.mark.gpu


def _model(cuda):
    torch.manual_seed(7)
    geom = WaveGeometryFreeForm((96, 80), 1.0, c0=1.0, c1=0.5, abs_N=12, rho="rand")
    cell = WaveCell(0.707, geom)
    src = [WaveSource(30, 40)]
    probes = [WaveIntensityProbe(70, 25), WaveIntensityProbe(70, 40), WaveIntensityProbe(70, 55)]
    return WaveRNN(cell, src, probes).to(cuda)


def _run(model, x, tuning, cuda):
    B, T = x.shape
    cfg = model.loop_config()
    tun = dict(tuning or {})
    tun["ck_every"] = model.checkpoint_every
    plan = Plan(cfg, B, T, float(model.cell.geom.cmax), cuda, tun)
    c = model.cell.geom.c.detach().clone().requires_grad_(True)
    xx = x.clone().requires_grad_(True)
    y = WaveLoop.apply(xx, c, model.cell.geom.b, None, plan, False, 1)
    g = torch.linspace(-1, 1, y.numel(), device=cuda).reshape(y.shape)
    (y * g).sum().backward()
    torch.cuda.synchronize()
    return y.detach(), c.grad.cpu(), xx.grad.cpu(), plan.tuning


def test_native_autotune_returns_a_valid_faster_or_equal_plan(cuda):
    model = _model(cuda)
    B, T = 6, 300
    n = torch.arange(T, dtype=torch.float32)
    x = torch.stack([torch.sin(2 * torch.pi * 0.05 * n + k) for k in range(B)]).to(cuda)
    model.checkpoint_every = 64
    # the heuristic's time with the same timing method (a one-pass search
    # budget: wt_autotune times its starting point first)
    best, ms = autotune.tune_native(model, x, reps=2, budget_s=20.0, store=False)
    print("tuned", best, f"{ms:.3f} ms", flush=True)
    assert ms > 0 and ms < 1e4
    assert best.get("kind", 0) == 0 and not best.get("plain", 0)
    y0, gc0, gx0, t0 = _run(model, x, None, cuda)
    y1, gc1, gx1, t1 = _run(model, x, best, cuda)
    print("heuristic", t0, "tuned", t1, flush=True)
    assert torch.equal(y0, y1)
    assert rel(gc1, gc0) <= 1e-5
    assert rel(gx1, gx0) <= 1e-5
    # cached: the same answer without a second search
    best2, ms2 = autotune.tune_native(model, x, reps=2, budget_s=20.0, store=False)
    assert best2 == best and ms2 == ms


def test_native_autotune_rejects_bad_arguments(cuda):
    model = _model(cuda)
    x = torch.zeros((2, 50), device=cuda)
    with pytest.raises(ValueError, match="reps"):
        autotune.tune_native(model, x, reps=0, budget_s=1.0, store=False)
    L = _lib.lib()
    out = _lib.WtTuning()
    st = L.wt_autotune(None, 2, 50, 1.0, None, 1, x.data_ptr(), x.data_ptr(), None,
                       x.data_ptr(), 1, 1.0, ctypes.byref(out), None, None)
    assert st == _lib.WT_EINVAL
