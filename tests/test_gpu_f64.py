"""GPU parity of the float64 time loop (set_dtype('float64'), utils.py:14-20;
the reference's own example configurations train in float64).

* Forward: the probe series and fields of wt_loop_fwd_f64 equal the C
  oracle's fp64 loop bit for bit (the same operation order in double, no
  contraction), for the linear cell and the nonlinear cells (saturable
  damping, Kerr speed, both, with the clamp extension), at h = 1 and at the
  reference example's h = 1.4283556979968262, dt = 1.0.
* Adjoint: dJ/dc, dJ/db (the linear cell's too), dJ/drho and dJ/dx of
  wt_loop_adj_f64 within 1e-10 relative of the oracle's fp64 adjoint (pinned
  by fp64 central differences and by the dot-product identity in
  test_oracle.py); the sums run in another order, so agreement is to
  rounding, not bitwise.
* WaveRNN end to end in float64 (FreeForm geometry: blur and projection by
  wt_geom_freeform_c_f64 and its backward, the double-precision loop, the
  reference's CE loss head): probe series within 1e-12 and dJ/drho within
  1e-9 of the reference's op sequence in float64 with its autograd tape
  (oracle/ref_torch.py, on the CPU).
* TimeStep.apply in float64 (the single-step API, wt_step_fwd_f64 /
  wt_step_bwd_f64): output within 1e-14 and gradients within 1e-12 of the
  restated reference step.
* WaveGeometryHoley in float64: c and its hole-parameter gradients against
  the restated reference in float64.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle.c_oracle as co
from oracle import ref_torch
from tests.helpers import rel, small_problem
from wavetorch_amd import (WaveCell, WaveGeometryFreeForm, WaveIntensityProbe, WaveRNN,
                           WaveSource, _lib)
from wavetorch_amd._loop import LoopConfig, Plan, WaveLoop
from wavetorch_amd.cell import TimeStep
from wavetorch_amd.utils import normalize_power, set_dtype

pytestmark = pytest.mark.gpu

NL_CASES = {
    "linear": {},
    "sat": dict(sat_b0=0.5, sat_uth=0.1 ** 0.5),
    "kerr": dict(c_nl=0.05),
    "both+clamp": dict(sat_b0=0.5, sat_uth=0.1 ** 0.5, c_nl=0.05, clamp=(0.55, 0.9)),
}


def _problem(h, dt, seed):
    p = small_problem(amp=0.3, seed=seed)
    p["h"], p["dt"] = h, dt
    rng = np.random.default_rng(200 + seed)
    p["rho"] = rng.uniform(0, 1, (p["Nx"], p["Ny"]))
    for k in ("b", "c", "x"):
        p[k] = p[k].astype(np.float64)
    return p


def _loop64(p, nl, cuda, ybar):
    cfg = LoopConfig(Nx=p["Nx"], Ny=p["Ny"], dt=p["dt"], h=p["h"],
                     src_px=tuple(map(tuple, p["src"])), taps=tuple(map(tuple, p["taps"])),
                     sat_b0=nl.get("sat_b0", 0.0), sat_uth=nl.get("sat_uth", 0.0),
                     c_nl=nl.get("c_nl", 0.0), c_clamp=nl.get("clamp"))
    plan = Plan(cfg, p["B"], p["T"], 1.0, cuda, {"plain": 1})
    t = lambda a, g=False: torch.from_numpy(np.asarray(a, dtype=np.float64)).to(cuda) \
        .requires_grad_(g)
    x, c, b = t(p["x"], True), t(p["c"], True), t(p["b"], True)
    rho = t(p["rho"], True) if nl else None
    y = WaveLoop.apply(x, c, b, rho, plan, False, 1)
    (y * t(ybar)).sum().backward()
    torch.cuda.synchronize()
    out = dict(y=y.detach().cpu().numpy(), gc=c.grad.cpu().numpy(), gx=x.grad.cpu().numpy(),
               gb=b.grad.cpu().numpy())
    if nl:
        out["grho"] = rho.grad.cpu().numpy()
    return out


@pytest.mark.parametrize("case", sorted(NL_CASES))
@pytest.mark.parametrize("h,dt", [(1.0, 0.707), (1.4283556979968262, 1.0)])
def test_loop_f64_matches_oracle(cuda, case, h, dt):
    nl = NL_CASES[case]
    p = _problem(h, dt, seed=3)
    if nl:  # the nonlinear speed exceeds c0 by up to c_nl u^2: stay inside CFL
        p["dt"] = 0.6 if h == 1.0 else 0.85
    ybar = np.random.default_rng(5).standard_normal((p["B"], p["T"], len(p["taps"])))
    print(f"float64 loop {case} h={h} dt={p['dt']}", flush=True)
    got = _loop64(p, nl, cuda, ybar)
    nlarg = dict(nl, rho=p["rho"]) if nl else None
    y64, _ = co.forward(p["b"], p["c"], p["x"], p["dt"], p["h"], p["src"], p["taps"], nl=nlarg,
                        dtype=np.float64)
    assert np.array_equal(got["y"], y64), f"forward differs from the fp64 oracle ({case})"
    g = co.adjoint(p["b"], p["c"], p["x"], p["dt"], p["h"], p["src"], p["taps"], ybar=ybar,
                   nl=nlarg)
    assert rel(got["gc"], g["gc"]) < 1e-10, rel(got["gc"], g["gc"])
    assert rel(got["gx"], g["gx"]) < 1e-10, rel(got["gx"], g["gx"])
    assert rel(got["gb"], g["gb"]) < 1e-10, rel(got["gb"], g["gb"])  # linear cell too
    if nl:
        assert rel(got["grho"], g["grho"]) < 1e-10, rel(got["grho"], g["grho"])


def test_loop_f64_fields_and_no_grad(cuda):
    """Field snapshots (stride 5) without gradients (three rotating fields)
    equal the oracle's bit for bit."""
    p = _problem(1.0, 0.707, seed=4)
    cfg = LoopConfig(Nx=p["Nx"], Ny=p["Ny"], dt=p["dt"], h=p["h"],
                     src_px=tuple(map(tuple, p["src"])), taps=tuple(map(tuple, p["taps"])))
    plan = Plan(cfg, p["B"], p["T"], 1.0, cuda, {"plain": 1})
    t = lambda a: torch.from_numpy(np.asarray(a, dtype=np.float64)).to(cuda)
    with torch.no_grad():
        f = WaveLoop.apply(t(p["x"]), t(p["c"]), t(p["b"]), None, plan, True, 5)
    _, f64 = co.forward(p["b"], p["c"], p["x"], p["dt"], p["h"], p["src"], p["taps"],
                        fields=True, dtype=np.float64, field_stride=5)
    assert f.dtype == torch.float64
    assert np.array_equal(f.cpu().numpy(), f64)


def test_wavernn_float64_training_step(cuda):
    """The reference example's shape class in float64: FreeForm geometry
    (blur, projection), WaveSource, three intensity probes, CE loss."""
    set_dtype("float64")
    try:
        torch.manual_seed(11)
        geom = WaveGeometryFreeForm((60, 50), 1.4283556979968262, c0=1.0, c1=0.5, abs_N=10,
                                    abs_sig=3.0, abs_p=4.0, rho="rand")
        cell = WaveCell(1.0, geom)
        src = [WaveSource(15, 25)]
        probes = [WaveIntensityProbe(45, 15), WaveIntensityProbe(45, 25),
                  WaveIntensityProbe(45, 35)]
        model = WaveRNN(cell, src, probes)
        B, T = 3, 400
        n = torch.arange(T, dtype=torch.float64)
        x = torch.stack([torch.sin(2 * np.pi * 0.05 * n + k) * torch.exp(-((n - 150) / 60) ** 2)
                         for k in range(B)])
        labels = torch.tensor([0, 2, 1])
        # the reference's op sequence in float64 on the CPU, with its tape
        rho_r = geom.rho.detach().clone().requires_grad_(True)
        c_r = ref_torch.freeform_c(rho_r, 1.0, 0.5, geom.blur_kernel, int(geom.blur_N.item()),
                                   geom.eta.item(), geom.beta.item())
        y_r = ref_torch.rnn_forward(x, c=c_r, rho=rho_r, b_geom=geom.b, dt=cell.dt, h=geom.h,
                                    sources=[(s.x.clone(), s.y.clone()) for s in src],
                                    probes=[(p.x.clone(), p.y.clone(), True) for p in probes])
        J_r = torch.nn.functional.cross_entropy(ref_torch.normalize_power(y_r.sum(dim=1)),
                                                labels)
        J_r.backward()
        # the HIP float64 path
        model = model.to(cuda)
        y = model(x.to(cuda))
        J = torch.nn.functional.cross_entropy(normalize_power(y.sum(dim=1)), labels.to(cuda))
        J.backward()
        torch.cuda.synchronize()
        assert y.dtype == torch.float64
        e_y = rel(y.detach().cpu().numpy(), y_r.detach().numpy())
        e_g = rel(model.cell.geom.rho.grad.cpu().numpy(), rho_r.grad.numpy())
        print(f"float64 WaveRNN: forward {e_y:.2e}, dJ/drho {e_g:.2e}", flush=True)
        assert e_y < 1e-12 and e_g < 1e-9
    finally:
        set_dtype("float32")


def test_timestep_float64(cuda):
    g = torch.Generator().manual_seed(9)
    B, Nx, Ny = 2, 20, 17
    mk = lambda *s: torch.rand(*s, generator=g, dtype=torch.float64)
    b, c = mk(Nx, Ny) * 0.3, 0.5 + 0.5 * mk(Nx, Ny)
    y1, y2, w = mk(B, Nx, Ny), mk(B, Nx, Ny), mk(B, Nx, Ny)
    dt, h = torch.tensor(0.6, dtype=torch.float64), torch.tensor(1.3, dtype=torch.float64)
    ins = [t.clone().requires_grad_(True) for t in (b, c, y1, y2)]
    ref = ref_torch.TimeStepRef.apply(*ins, dt, h)
    (ref * w).sum().backward()
    dins = [t.clone().to(cuda).requires_grad_(True) for t in (b, c, y1, y2)]
    out = TimeStep.apply(*dins, dt.to(cuda), h.to(cuda))
    (out * w.to(cuda)).sum().backward()
    assert out.dtype == torch.float64
    assert rel(out.detach().cpu().numpy(), ref.detach().numpy()) < 1e-14
    for a, r in zip(dins, ins):
        assert rel(a.grad.cpu().numpy(), r.grad.numpy()) < 1e-12


def test_holey_geometry_float64(cuda):
    """WaveGeometryHoley in float64 (geom.py:88-135): c and dJ/d(x, y, r) from
    wt_geom_holey_c_f64 / _bwd_f64 equal the restated reference
    (oracle/ref_torch.holey_c) in float64 on the CPU, off-grid hole centres."""
    set_dtype("float64")
    try:
        from wavetorch_amd import WaveGeometryHoley
        rng = np.random.default_rng(21)
        shape = (48, 40)
        x = rng.uniform(8, 40, 6) + 0.37
        y = rng.uniform(8, 32, 6) + 0.29
        r = rng.uniform(2.0, 5.0, 6)
        g = WaveGeometryHoley(shape, 1.0, c0=1.0, c1=0.5, abs_N=5, eta=0.5, beta=100.0,
                              x=x, y=y, r=r).to(cuda)
        w = torch.from_numpy(rng.standard_normal(shape))
        c = g.c
        (c * w.to(cuda)).sum().backward()
        torch.cuda.synchronize()
        hx, hy, hr = (torch.from_numpy(v).requires_grad_(True) for v in (x, y, r))
        c_r = ref_torch.holey_c(hx, hy, hr, shape, 0.5, 100.0, 1.0, 0.5)
        (c_r * w).sum().backward()
        assert c.dtype == torch.float64
        assert rel(c.detach().cpu().numpy(), c_r.detach().numpy()) < 1e-13
        for p, q in ((g.x, hx), (g.y, hy), (g.r, hr)):
            assert rel(p.grad.cpu().numpy(), q.grad.numpy()) < 1e-10
    finally:
        set_dtype("float32")


def test_nonfinite_flags_float64(cuda):
    """check_finite in float64 (wt_check_finite_f64): flags stay 0 for a
    finite run; a NaN input sample sets the output flag, an infinite dJ/dy
    the gradient flag."""
    set_dtype("float64")
    try:
        from wavetorch_amd import workloads
        from wavetorch_amd.utils import normalize_power
        model, x, labels = workloads.propagate_100_b4(T=48)
        model = model.to(cuda)
        model.check_finite = True
        x, labels = x.double().to(cuda), labels.to(cuda)
        y = model(x)
        torch.nn.functional.cross_entropy(normalize_power(y.sum(dim=1)), labels).backward()
        assert model.nonfinite.tolist() == [0, 0]
        xb = x.clone()
        xb[1, 10] = float("nan")
        model(xb)
        assert model.nonfinite.tolist() == [1, 0]
        model.nonfinite.zero_()
        y = model(x)
        g = torch.zeros_like(y)
        g[0, 5, 1] = float("inf")
        y.backward(g)
        assert model.nonfinite.tolist() == [0, 1]
    finally:
        set_dtype("float32")
