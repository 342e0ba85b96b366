"""GPU parity of the nonlinear cell (cell.py:8-9, 94-102): saturable damping
b = b_geom + rho*b0/(1+(u/uth)^2) and Kerr-like speed c = c_lin + rho*c_nl*u^2,
optionally clamped (an extension).

Forward: probe series equal the C oracle's fp32 loop bit for bit for every
kernel family.  Adjoint: dJ/dc, dJ/db, dJ/drho and dJ/dx within 1e-4 relative
of the fp64 oracle adjoint (itself pinned by central finite differences and
by torch autograd through the restated reference cell in test_oracle.py), over
several checkpoint intervals so that segment boundaries fall mid-run."""
import os

import numpy as np
import pytest
import torch

import oracle.c_oracle as co
from tests.helpers import rel, small_problem
from wavetorch_amd._loop import LoopConfig, Plan, WaveLoop

pytestmark = pytest.mark.gpu

CASES = {
    "sat": dict(sat_b0=0.5, sat_uth=0.1 ** 0.5),
    "kerr": dict(c_nl=0.05),
    "both": dict(sat_b0=0.5, sat_uth=0.1 ** 0.5, c_nl=0.05),
    "both+clamp": dict(sat_b0=0.5, sat_uth=0.1 ** 0.5, c_nl=0.05, clamp=(0.55, 0.9)),
}
TUNINGS = [{"plain": 1}, {}, {"kind": 1},
           {"kind": 0, "strips": 2, "tcols": 2, "tblock": 8, "nt": 1024, "cpt": 8}]


def _problem(seed=0):
    p = small_problem(amp=0.3, seed=seed)
    p["dt"] = 0.6
    rng = np.random.default_rng(100 + seed)
    p["rho"] = rng.uniform(0, 1, (p["Nx"], p["Ny"])).astype(np.float32)
    return p


def _run(p, case, tun, cuda, K, grad):
    nl = CASES[case]
    cfg = LoopConfig(Nx=p["Nx"], Ny=p["Ny"], dt=p["dt"], h=1.0,
                     src_px=tuple(map(tuple, p["src"])), taps=tuple(map(tuple, p["taps"])),
                     sat_b0=nl.get("sat_b0", 0.0), sat_uth=nl.get("sat_uth", 0.0),
                     c_nl=nl.get("c_nl", 0.0), c_clamp=nl.get("clamp"))
    plan = Plan(cfg, p["B"], p["T"], 1.0, cuda, {**tun, "ck_every": K})
    t = {k: torch.from_numpy(p[k]).to(cuda).requires_grad_(grad) for k in ("x", "c", "b", "rho")}
    y = WaveLoop.apply(t["x"], t["c"], t["b"], t["rho"], plan, False, 1)
    return plan, t, y


def _oracle_nl(p, case):
    nl = dict(CASES[case])
    nl["rho"] = p["rho"]
    return nl


GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", ["nlsat", "nlkerr", "nlboth"])
@pytest.mark.parametrize("tun,K", [({"plain": 1}, 16), ({}, 20), ({"kind": 1}, 48),
                                   ({"kind": 0, "strips": 2, "tcols": 2, "tblock": 8,
                                     "nt": 1024, "cpt": 8}, 13)],
                         ids=["plain-K16", "auto-K20", "flat-K48", "tiles2x2-K13"])
def test_nonlinear_reference_golden(name, tun, K, gold, cuda):
    """Against a 48-step loop of the reference's own WaveCell.forward
    (cell.py:79-107, tests/golden/make_golden.py): probe series and the last
    two fields bit for bit, and the reference's autograd gradients with
    respect to c_linear, b, raw rho and x within 1e-4."""
    sb0, suth, cnl = (float(v) for v in gold[f"{name}_par"])
    B, T = gold["nl_x"].shape
    Nx, Ny = gold["nl_b"].shape
    cfg = LoopConfig(Nx=Nx, Ny=Ny, dt=float(gold["nl_dt"]), h=1.0,
                     src_px=tuple(map(tuple, gold["nl_src"].tolist())),
                     taps=tuple(map(tuple, gold["nl_taps"].tolist())),
                     sat_b0=sb0, sat_uth=suth, c_nl=cnl)
    print("case", name, tun, K, flush=True)
    plan = Plan(cfg, B, T, 1.0, cuda, {**tun, "ck_every": K})
    t = {k: torch.from_numpy(gold[f"nl_{k}"]).to(cuda).requires_grad_(True)
         for k in ("x", "c", "b", "rho")}
    y = WaveLoop.apply(t["x"], t["c"], t["b"], t["rho"], plan, False, 1)
    assert np.array_equal(y.detach().cpu().numpy(), gold[f"{name}_y"]), plan.tuning
    (y * torch.from_numpy(gold["nl_w"]).to(cuda)).sum().backward()
    for key, arg in (("gc", "c"), ("gb", "b"), ("grho", "rho"), ("gx", "x")):
        ref = gold[f"{name}_{key}"]
        got = t[arg].grad.cpu().numpy()
        if not np.any(ref):  # the Kerr-free cell has no rho dependence
            assert np.abs(got).max() == 0.0, key
            continue
        assert rel(got, ref) < 1e-4, (key, rel(got, ref), plan.tuning)
    with torch.no_grad():
        f = WaveLoop.apply(t["x"], t["c"], t["b"], t["rho"], plan, True, 1)
    f = f.cpu().numpy()
    assert np.array_equal(f[:, -1], gold[f"{name}_u1"]), plan.tuning
    assert np.array_equal(f[:, -2], gold[f"{name}_u2"]), plan.tuning


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("tun", TUNINGS, ids=lambda t: "-".join(f"{k}{v}" for k, v in t.items()) or "auto")
def test_nonlinear_forward_bitwise(case, tun, cuda):
    p = _problem()
    print("case", case, tun, flush=True)
    plan, _, y = _run(p, case, tun, cuda, 32, False)
    yo, _ = co.forward(p["b"], p["c"], p["x"], p["dt"], 1.0, p["src"], p["taps"],
                       nl=_oracle_nl(p, case))
    yg = y.cpu().numpy()
    assert np.isfinite(yg).all() and np.abs(yo).max() > 0
    assert np.array_equal(yg, yo), f"max |diff| {np.abs(yg - yo).max()} ({plan.tuning})"


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("tun,K", [({"plain": 1}, 32), ({}, 16), ({"kind": 1}, 50),
                                   ({"kind": 0, "strips": 2, "tcols": 2, "tblock": 8,
                                     "nt": 1024, "cpt": 8}, 120)],
                         ids=["plain-K32", "auto-K16", "flat-K50", "tiles2x2-K120"])
def test_nonlinear_adjoint_vs_oracle(case, tun, K, cuda):
    p = _problem(seed=1)
    print("case", case, tun, K, flush=True)
    plan, t, y = _run(p, case, tun, cuda, K, True)
    w = np.random.default_rng(9).standard_normal(tuple(y.shape)).astype(np.float32)
    (y * torch.from_numpy(w).to(cuda)).sum().backward()
    g = co.adjoint(p["b"], p["c"], p["x"], p["dt"], 1.0, p["src"], p["taps"], ybar=w,
                   nl=_oracle_nl(p, case))
    for key, name in (("gc", "c"), ("gb", "b"), ("grho", "rho"), ("gx", "x")):
        r = rel(t[name].grad.cpu().numpy(), g[key])
        assert r < 1e-4, (key, r, plan.tuning)
