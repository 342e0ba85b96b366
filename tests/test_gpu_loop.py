"""GPU parity of the fused time loop against the oracle.

Forward: the HIP kernels follow the reference's fp32 operation order with no
contraction, so probe series and fields must equal the C oracle's fp32 loop
(wave_oracle.c, itself bitwise equal to the torch restatement of the
reference) bit for bit, for the plain and every tiled variant.
Adjoint: gradients wrt c and x must match the fp64 oracle adjoint within
rel 1e-4 (BASELINE.json tolerance for gradients).
"""
import numpy as np
import pytest
import torch

import oracle.c_oracle as co
from tests.helpers import rel, small_problem
from wavetorch_amd._loop import LoopConfig, Plan, WaveLoop

pytestmark = pytest.mark.gpu

TUNINGS = [
    {"plain": 1},
    {},                                              # library heuristic
    {"strips": 1, "tblock": 64, "nt": 1024, "cpt": 8},
    {"strips": 3, "tblock": 4, "nt": 1024, "cpt": 8},
    {"strips": 2, "tblock": 16, "nt": 512, "cpt": 16},
    {"strips": 5, "tblock": 7, "nt": 256, "cpt": 32},
]


def _cfg(p):
    return LoopConfig(Nx=p["Nx"], Ny=p["Ny"], dt=p["dt"], h=p["h"],
                      src_px=tuple(map(tuple, p["src"])),
                      taps=tuple((t[0], t[1], t[2]) for t in p["taps"]))


def _run(p, tun, cuda, grad=False, K=32):
    cfg = _cfg(p)
    tun = dict(tun)
    tun["ck_every"] = K
    plan = Plan(cfg, p["B"], p["T"], 1.0, cuda, tun)
    x = torch.from_numpy(p["x"]).to(cuda).requires_grad_(grad)
    c = torch.from_numpy(p["c"]).to(cuda).requires_grad_(grad)
    b = torch.from_numpy(p["b"]).to(cuda)
    y = WaveLoop.apply(x, c, b, None, plan, False, 1)
    return plan, x, c, y


@pytest.mark.parametrize("tun", TUNINGS, ids=lambda t: "-".join(f"{k}{v}" for k, v in t.items()) or "auto")
def test_forward_bitwise(tun, cuda):
    p = small_problem()
    print("case", tun, flush=True)
    plan, x, c, y = _run(p, tun, cuda)
    yo, _ = co.forward(p["b"], p["c"], p["x"], p["dt"], p["h"], p["src"], p["taps"])
    yg = y.cpu().numpy()
    assert np.isfinite(yg).all()
    assert np.array_equal(yg, yo), f"max |diff| {np.abs(yg - yo).max()} (tuning {plan.tuning})"


@pytest.mark.parametrize("tun", TUNINGS, ids=lambda t: "-".join(f"{k}{v}" for k, v in t.items()) or "auto")
def test_adjoint_vs_oracle(tun, cuda):
    p = small_problem()
    print("case", tun, flush=True)
    plan, x, c, y = _run(p, tun, cuda, grad=True)
    rng = np.random.default_rng(5)
    w = rng.standard_normal(y.shape).astype(np.float32)
    (y * torch.from_numpy(w).to(cuda)).sum().backward()
    g = co.adjoint(p["b"], p["c"], p["x"], p["dt"], p["h"], p["src"], p["taps"], ybar=w)
    assert rel(c.grad.cpu().numpy(), g["gc"]) < 1e-4
    assert rel(x.grad.cpu().numpy(), g["gx"]) < 1e-4
