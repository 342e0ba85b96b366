"""GPU parity of the fused time loop against the oracle.

Forward: the HIP kernels follow the reference's fp32 operation order with no
contraction, so probe series and fields must equal the C oracle's fp32 loop
(wave_oracle.c, itself bitwise equal to the torch restatement of the
reference) bit for bit, for the plain and every tiled variant.
Adjoint: gradients wrt c and x must match the fp64 oracle adjoint within
rel 1e-4 (the gradient tolerance of the path).
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
    # flat cell-list family
    {"kind": 1, "strips": 1, "tblock": 64, "nt": 1024, "cpt": 8},
    {"kind": 1, "strips": 3, "tblock": 4, "nt": 1024, "cpt": 8},
    {"kind": 1, "strips": 2, "tblock": 16, "nt": 512, "cpt": 16},
    {"kind": 1, "strips": 5, "tblock": 7, "nt": 256, "cpt": 32},
    # column-segment family (cpt = rows per thread)
    {"kind": 0, "strips": 1, "tblock": 64, "nt": 1024, "cpt": 4},
    {"kind": 0, "strips": 2, "tblock": 9, "nt": 1024, "cpt": 6},
    {"kind": 0, "strips": 3, "tblock": 5, "nt": 1024, "cpt": 12},
    {"kind": 0, "strips": 4, "tblock": 16, "nt": 1024, "cpt": 10},
    {"kind": 0, "strips": 6, "tblock": 3, "nt": 512, "cpt": 16},
    {"kind": 0, "strips": 2, "tcols": 2, "tblock": 8, "nt": 1024, "cpt": 8},
    {"kind": 0, "strips": 3, "tcols": 3, "tblock": 6, "nt": 1024, "cpt": 4},
    {"kind": 0, "strips": 2, "tcols": 2, "tblock": 12, "nt": 1024, "cpt": 8,
     "a_tblock": 10, "a_nt": 512, "a_cpt": 16},
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


def test_train_shape_bitwise_and_grad(cuda):
    """The 150x150 training geometry (20-cell PML, source (40,75), probes at
    row 110) at batch 4 and 200 steps, library-chosen tiling."""
    from oracle import ref_torch
    rng = np.random.default_rng(21)
    N, B, T = 150, 4, 200
    b = ref_torch.pml_b((N, N), 20, 1.0, 2).numpy()
    c = np.ones((N, N), np.float32)
    c[20:130, 20:130] = rng.uniform(0.5, 1.0, (110, 110)).astype(np.float32)
    n = np.arange(T)
    x = np.stack([np.sin(2 * np.pi * (0.03 + 0.01 * k) * n) for k in range(B)]).astype(np.float32)
    p = dict(b=b, c=c, x=x, src=[[40, 75, 0]], taps=[[110, 55, 1], [110, 75, 1], [110, 95, 1]],
             dt=0.707, h=1.0, Nx=N, Ny=N, B=B, T=T)
    for tun in ({}, {"kind": 1}, {"plain": 1},
                {"strips": 2, "tcols": 2, "tblock": 12, "nt": 1024, "cpt": 8},
                {"strips": 2, "tcols": 2, "tblock": 12, "nt": 1024, "cpt": 8,
                 "a_tblock": 10, "a_nt": 512, "a_cpt": 16}):
        print("case", tun, flush=True)
        plan, xt, ct, y = _run(p, tun, cuda, grad=True, K=64)
        yo, _ = co.forward(b, c, x, 0.707, 1.0, p["src"], p["taps"])
        assert np.array_equal(y.detach().cpu().numpy(), yo), plan.tuning
        w = rng.standard_normal(y.shape).astype(np.float32)
        (y * torch.from_numpy(w).to(cuda)).sum().backward()
        g = co.adjoint(b, c, x, 0.707, 1.0, p["src"], p["taps"], ybar=w)
        assert rel(ct.grad.cpu().numpy(), g["gc"]) < 1e-4, plan.tuning
        assert rel(xt.grad.cpu().numpy(), g["gx"]) < 1e-4, plan.tuning


EDGE = [
    # (Nx, Ny, B, T, K, tuning)
    (40, 36, 1, 1, 32, {}),                                   # one step
    (40, 36, 2, 2, 32, {}),
    (40, 36, 2, 3, 32, {"strips": 2, "tcols": 2, "tblock": 8, "nt": 1024, "cpt": 8}),
    (40, 36, 2, 20, 64, {}),                                  # K > T
    (40, 36, 3, 97, 32, {}),                                  # T = 3K + 1
    (40, 36, 3, 97, 32, {"plain": 1}),
    (40, 36, 2, 10, 1, {}),                                   # a checkpoint every step
    (23, 19, 2, 60, 16, {"strips": 3, "tcols": 3, "tblock": 12, "nt": 1024, "cpt": 4}),
]


@pytest.mark.parametrize("Nx,Ny,B,T,K,tun", EDGE,
                         ids=[f"{e[0]}x{e[1]}-B{e[2]}-T{e[3]}-K{e[4]}-{'-'.join(map(str, e[5].values())) or 'auto'}"
                              for e in EDGE])
def test_edge_sizes(Nx, Ny, B, T, K, tun, cuda):
    """One to three steps, K larger than T, T not a multiple of K, a
    checkpoint every step, batch 1, and a small odd grid whose tiles are
    narrower than the time-block halo: forward bitwise, gradients within 1e-4
    of the fp64 oracle."""
    p = small_problem(Nx=Nx, Ny=Ny, B=B, T=T)
    print("case", Nx, Ny, B, T, K, tun, flush=True)
    plan, x, c, y = _run(p, tun, cuda, grad=True, K=K)
    yo, _ = co.forward(p["b"], p["c"], p["x"], p["dt"], p["h"], p["src"], p["taps"])
    assert y.shape == (B, T, 3)
    assert np.array_equal(y.detach().cpu().numpy(), yo), plan.tuning
    w = np.random.default_rng(T).standard_normal(y.shape).astype(np.float32)
    (y * torch.from_numpy(w).to(cuda)).sum().backward()
    g = co.adjoint(p["b"], p["c"], p["x"], p["dt"], p["h"], p["src"], p["taps"], ybar=w)
    assert rel(c.grad.cpu().numpy(), g["gc"]) < 1e-4, plan.tuning
    assert rel(x.grad.cpu().numpy(), g["gx"]) < 1e-4, plan.tuning


def test_loop_graph_replay(cuda):
    """wt_loop_fwd + wt_loop_adj captured once (wt_graph_begin/end on a side
    stream) and replayed three times with new c and x written into the same
    buffers, eager runs in between: every replay equals the eager result
    bit for bit (the loop records no memset nodes: buffers are cleared by
    fill kernels)."""
    import ctypes
    from wavetorch_amd import _lib
    L = _lib.lib()
    p = small_problem(B=2, T=90)
    cfg = _cfg(p)
    plan = Plan(cfg, p["B"], p["T"], 1.0, cuda, {"ck_every": 32})
    B, T, ntap = p["B"], p["T"], len(cfg.taps)
    b = torch.from_numpy(p["b"]).to(cuda)
    c = torch.from_numpy(p["c"]).to(cuda)
    x = torch.from_numpy(p["x"]).to(cuda)
    ybar = torch.from_numpy(np.random.default_rng(3).standard_normal((B, T, ntap))
                            .astype(np.float32)).to(cuda)
    y = torch.empty((B, T, ntap), device=cuda)
    saved = torch.empty(plan.saved_bytes, dtype=torch.uint8, device=cuda)
    temp = torch.empty(plan.temp_bytes, dtype=torch.uint8, device=cuda)
    gc = torch.empty_like(c)
    gx = torch.empty_like(x)
    s = torch.cuda.Stream(device=cuda)

    def run(stream):
        sp = stream.cuda_stream
        _lib.check(L.wt_loop_fwd(plan.handle, b.data_ptr(), c.data_ptr(), None, x.data_ptr(),
                                 y.data_ptr(), None, 1, 1, saved.data_ptr(), temp.data_ptr(),
                                 sp), "wt_loop_fwd")
        _lib.check(L.wt_loop_adj(plan.handle, b.data_ptr(), c.data_ptr(), None, x.data_ptr(),
                                 ybar.data_ptr(), None, saved.data_ptr(), temp.data_ptr(),
                                 gc.data_ptr(), None, None, gx.data_ptr(), sp), "wt_loop_adj")

    graph = None
    try:
        with torch.cuda.stream(s):
            for rep in range(3):
                c.copy_(torch.from_numpy(p["c"]).to(cuda) * (1.0 - 0.04 * rep))
                x.copy_(torch.from_numpy(p["x"]).to(cuda) * (1.0 + 0.5 * rep))
                run(s)  # eager
                s.synchronize()
                ref = (y.clone(), gc.clone(), gx.clone())
                if graph is None:
                    _lib.check(L.wt_graph_begin(s.cuda_stream), "wt_graph_begin")
                    run(s)
                    g = ctypes.c_void_p()
                    _lib.check(L.wt_graph_end(s.cuda_stream, ctypes.byref(g)), "wt_graph_end")
                    graph = g
                for t in (y, gc, gx):
                    t.fill_(float("nan"))
                _lib.check(L.wt_graph_launch(graph, s.cuda_stream), "wt_graph_launch")
                s.synchronize()
                assert torch.equal(y, ref[0]) and torch.equal(gc, ref[1]) and \
                    torch.equal(gx, ref[2]), rep
    finally:
        if graph is not None:
            L.wt_graph_destroy(graph)
