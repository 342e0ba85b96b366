"""GPU parity at the full sizes of the benchmark configurations.

* 100x100, batch 4, 1000 steps (the CPU-runnable configuration): forward
  probes bitwise equal to the C oracle, dJ/dc within 1e-4 of the fp64 oracle.
* 1024x1024, 4000 steps, snapshots every 500 (forward-only): snapshots of the
  temporally blocked kernels bitwise equal to the one-launch-per-step kernels
  at the full length, and to the C oracle over the first 300 steps.
* 150x150, 3000 steps (training geometry): batch 4 at the full length against
  the fp64 oracle; batch 64 (the benchmark batch) blocked vs plain, forward
  bitwise and gradients within 1e-5 (the two differ only in the order of
  the fp32 gradient sums).
* nonlinear 150x150 batch 32: blocked vs plain over 1500 steps.  (Over the
  full 3000 steps the adjoint of this configuration grows without bound for
  some samples -- the fp64 oracle gives |dJ/dc| ~ 1e24 for sample 26 -- because
  saturable damping falls with intensity, so the linearised dynamics have
  gain; fp32 then overflows in any implementation, the reference's included.)
"""
import numpy as np
import pytest
import torch

import oracle.c_oracle as co
from tests.helpers import rel
from wavetorch_amd import workloads
from wavetorch_amd._loop import Plan, WaveLoop

pytestmark = pytest.mark.gpu


def _loop_inputs(model, x, cuda):
    model = model.to(cuda)
    cfg = model.loop_config()
    g = model.cell.geom
    with torch.no_grad():
        c = g.c.detach().clone()
    return model, cfg, c, g.b.detach().clone(), x.to(cuda)


def _run(cfg, B, T, cuda, tun, c, b, x, rho=None, fields=False, stride=1, grad=False, K=64):
    plan = Plan(cfg, B, T, 1.0, cuda, {**tun, "ck_every": K})
    ct = c.clone().requires_grad_(grad)
    xt = x.clone().requires_grad_(grad)
    rt = rho.clone().requires_grad_(grad) if rho is not None else None
    y = WaveLoop.apply(xt, ct, b, rt, plan, fields, stride)
    return plan, y, ct, xt, rt


def test_propagate_100_b4_full_vs_oracle(cuda):
    model, x, _ = workloads.propagate_100_b4()
    model, cfg, c, b, x = _loop_inputs(model, x, cuda)
    B, T = x.shape
    src = [list(p) for p in cfg.src_px]
    taps = [list(t) for t in cfg.taps]
    yo, _ = co.forward(b.cpu().numpy(), c.cpu().numpy(), x.cpu().numpy(), cfg.dt, 1.0, src, taps)
    w = np.random.default_rng(3).standard_normal(yo.shape).astype(np.float32)
    g = co.adjoint(b.cpu().numpy(), c.cpu().numpy(), x.cpu().numpy(), cfg.dt, 1.0, src, taps,
                   ybar=w)
    for tun in ({}, {"plain": 1}):
        print("case", tun, flush=True)
        plan, y, ct, xt, _ = _run(cfg, B, T, cuda, tun, c, b, x, grad=True)
        assert np.array_equal(y.detach().cpu().numpy(), yo), plan.tuning
        (y * torch.from_numpy(w).to(cuda)).sum().backward()
        assert rel(ct.grad.cpu().numpy(), g["gc"]) < 1e-4, plan.tuning
        assert rel(xt.grad.cpu().numpy(), g["gx"]) < 1e-4, plan.tuning


def test_scatter_1024_snapshots(cuda):
    model, x, _ = workloads.scatter_1024()
    model, cfg, c, b, x = _loop_inputs(model, x, cuda)
    T = x.shape[1]
    outs = {}
    for tun in ({}, {"plain": 1}):
        print("case full", tun, flush=True)
        plan, f, _, _, _ = _run(cfg, 1, T, cuda, tun, c, b, x, fields=True, stride=500)
        outs[plan.tuning["plain"]] = f.cpu().numpy()
    assert outs[0].shape == (1, 8, 1024, 1024)
    assert np.abs(outs[0][0, -1]).max() > 0
    assert np.array_equal(outs[0], outs[1])
    # the first 300 steps against the oracle, snapshots every 100
    Ts = 300
    print("case oracle", flush=True)
    plan, f, _, _, _ = _run(cfg, 1, Ts, cuda, {}, c, b, x[:, :Ts].contiguous(), fields=True,
                            stride=100)
    _, fo = co.forward(b.cpu().numpy(), c.cpu().numpy(), x[:, :Ts].cpu().numpy(), cfg.dt, 1.0,
                       [list(p) for p in cfg.src_px], [], fields=True)
    assert np.array_equal(f.cpu().numpy(), fo[:, 99::100])


def test_train_150_full_length_b4_vs_oracle(cuda):
    model, x, _ = workloads.train_150_b64(B=4)
    model, cfg, c, b, x = _loop_inputs(model, x, cuda)
    B, T = x.shape
    src = [list(p) for p in cfg.src_px]
    taps = [list(t) for t in cfg.taps]
    bn, cn, xn = b.cpu().numpy(), c.cpu().numpy(), x.cpu().numpy()
    yo, _ = co.forward(bn, cn, xn, cfg.dt, 1.0, src, taps)
    w = np.random.default_rng(4).standard_normal(yo.shape).astype(np.float32)
    g = co.adjoint(bn, cn, xn, cfg.dt, 1.0, src, taps, ybar=w)
    print("case b4 full", flush=True)
    plan, y, ct, xt, _ = _run(cfg, B, T, cuda, {}, c, b, x, grad=True)
    assert np.array_equal(y.detach().cpu().numpy(), yo), plan.tuning
    (y * torch.from_numpy(w).to(cuda)).sum().backward()
    assert rel(ct.grad.cpu().numpy(), g["gc"]) < 1e-4
    assert rel(xt.grad.cpu().numpy(), g["gx"]) < 1e-4


@pytest.mark.parametrize("wl", ["train_150_b64", "nonlinear_150_b32"])
def test_full_batch_blocked_equals_plain(wl, cuda):
    model, x, _ = getattr(workloads, wl)(**({"T": 1500} if wl.startswith("nonlinear") else {}))
    model, cfg, c, b, x = _loop_inputs(model, x, cuda)
    B, T = x.shape
    rho = model.cell.geom.rho.detach().clone() if cfg.nonlinear else None
    w = torch.from_numpy(np.random.default_rng(6).standard_normal((B, T, 3)).astype(np.float32)
                         ).to(cuda)
    res = {}
    for tun in ({}, {"plain": 1}):
        print("case", wl, tun, flush=True)
        plan, y, ct, xt, rt = _run(cfg, B, T, cuda, tun, c, b, x, rho=rho, grad=True)
        (y * w).sum().backward()
        res[plan.tuning["plain"]] = (y.detach().cpu().numpy(), ct.grad.cpu().numpy(),
                                     xt.grad.cpu().numpy(),
                                     rt.grad.cpu().numpy() if rt is not None else None)
        del plan, y, ct, xt, rt
    (y0, gc0, gx0, gr0), (y1, gc1, gx1, gr1) = res[0], res[1]
    assert np.isfinite(y0).all() and np.abs(y0).max() > 0
    assert np.isfinite(gc0).all() and np.isfinite(gx0).all()
    assert np.array_equal(y0, y1)
    assert rel(gc0, gc1) < 1e-5 and rel(gx0, gx1) < 1e-5
    if gr0 is not None:
        assert rel(gr0, gr1) < 1e-5


def test_nonlinear_150_vs_oracle(cuda):
    """The nonlinear 150x150 configuration (saturable damping + clamped Kerr
    speed, dt 0.64) at batch 2 over 600 steps: forward bitwise, dJ/dc,
    dJ/drho and dJ/dx within 1e-4 of the fp64 oracle.  (Past ~900 steps the
    exact gradient of this configuration moves by more than 1e-4 under a 1e-7
    relative change of its inputs -- test_oracle.py pins that -- so no fp32
    computation can match it there.)"""
    import math
    model, x, _ = workloads.nonlinear_150_b32(B=2, T=600)
    model, cfg, c, b, x = _loop_inputs(model, x, cuda)
    B, T = x.shape
    rho = model.cell.geom.rho.detach().clone()
    src = [list(p) for p in cfg.src_px]
    taps = [list(t) for t in cfg.taps]
    nl = dict(rho=rho.cpu().numpy(), sat_b0=0.5, sat_uth=math.sqrt(0.1), c_nl=0.05,
              clamp=(0.4, 1.1))
    bn, cn, xn = b.cpu().numpy(), c.cpu().numpy(), x.cpu().numpy()
    yo, _ = co.forward(bn, cn, xn, cfg.dt, 1.0, src, taps, nl=nl)
    w = np.random.default_rng(8).standard_normal(yo.shape).astype(np.float32)
    g = co.adjoint(bn, cn, xn, cfg.dt, 1.0, src, taps, ybar=w, nl=nl)
    print("case nonlinear b2", flush=True)
    plan, y, ct, xt, rt = _run(cfg, B, T, cuda, {}, c, b, x, rho=rho, grad=True)
    assert np.array_equal(y.detach().cpu().numpy(), yo), plan.tuning
    (y * torch.from_numpy(w).to(cuda)).sum().backward()
    assert rel(ct.grad.cpu().numpy(), g["gc"]) < 1e-4
    assert rel(rt.grad.cpu().numpy(), g["grho"]) < 1e-4
    assert rel(xt.grad.cpu().numpy(), g["gx"]) < 1e-4
