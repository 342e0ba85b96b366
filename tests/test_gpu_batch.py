"""GPU parity of batched instances (wavetorch_amd.batch / wt_batch_*): many
independent problems of uneven grid size and length in one launch sequence.

Each instance's probe series must equal the C oracle's fp32 loop for that
instance alone, bit for bit, over its own length; its dJ/dc and dJ/dx must be
within 1e-4 of the fp64 oracle adjoint.  A captured graph replayed with new
inputs (and eager work between the replays) must reproduce the eager run."""
import numpy as np
import pytest
import torch

import oracle.c_oracle as co
from oracle import ref_torch
from tests.helpers import rel
from wavetorch_amd._loop import LoopConfig
from wavetorch_amd.batch import Instance, InstanceBatch

pytestmark = pytest.mark.gpu


def _instances(seed=0):
    """Two grid sizes x three lengths, a line source and a focal probe each
    (the lens-style problem shape at small size)."""
    rng = np.random.default_rng(seed)
    out = []
    for i, (N, T) in enumerate([(40, 150), (56, 260), (40, 211), (56, 97), (40, 260),
                                (56, 180)]):
        pml = 6
        r = pml + 3
        src = tuple((r, j, 0) for j in range(N // 4, 3 * N // 4))
        taps = ((N - pml - N // 4, N // 2, 1),)
        cfg = LoopConfig(Nx=N, Ny=N, dt=0.707, h=1.0, src_px=src, taps=taps)
        b = ref_torch.pml_b((N, N), pml, 1.0, 2)
        c = np.ones((N, N), np.float32)
        c[pml + 1:N - pml - 1, pml + 1:N - pml - 1] = rng.uniform(
            0.5, 1.0, (N - 2 * pml - 2, N - 2 * pml - 2))
        x = np.sin(2 * np.pi * 0.1 * np.arange(T)).astype(np.float32) * rng.uniform(0.5, 1.5)
        out.append((Instance(cfg, b, T, 1.0), c, x))
    return out


def _oracle(inst, c, x, ybar):
    src = [list(p) for p in inst.cfg.src_px]
    taps = [list(t) for t in inst.cfg.taps]
    b = inst.b.numpy()
    yo, _ = co.forward(b, c, x[None], inst.cfg.dt, 1.0, src, taps)
    g = co.adjoint(b, c, x[None], inst.cfg.dt, 1.0, src, taps, ybar=ybar[None])
    return yo[0], g["gc"], g["gx"][0]


@pytest.mark.parametrize("tuning,K", [(None, 64), (None, 32), (None, 300),
                                      ({"tblock": 8, "nt": 1024, "cpt": 8, "a_tblock": 6,
                                        "a_nt": 1024, "a_cpt": 6, "strips": 2, "tcols": 2}, 40)],
                         ids=["auto-K64", "auto-K32", "auto-K300-one-segment", "tiles2x2-K40"])
def test_batch_matches_oracle(tuning, K, cuda):
    items = _instances()
    bt = InstanceBatch([it[0] for it in items], cuda, tuning=tuning, checkpoint_every=K)
    print("case", bt.tuning, flush=True)
    cs = [torch.from_numpy(it[1]).to(cuda) for it in items]
    xs = [torch.from_numpy(it[2]).to(cuda) for it in items]
    # focal-intensity objective J = -sum_n y[n]  ->  dJ/dy = -1 over each length
    ybars = [-torch.ones((it[0].length, 1), device=cuda) for it in items]
    bt.set_inputs(cs, xs, ybars)
    bt.fwd_adj()
    torch.cuda.synchronize()
    for (inst, c, x), (y, gc, gx) in zip(items, bt.outputs()):
        yo, gco, gxo = _oracle(inst, c, x, -np.ones((inst.length, 1)))
        assert np.array_equal(y.cpu().numpy(), yo), inst.length
        assert rel(gc.cpu().numpy(), gco) < 1e-4, inst.length
        assert rel(gx.cpu().numpy(), gxo) < 1e-4, inst.length


def test_batch_graph_replay(cuda):
    items = _instances(seed=1)
    bt = InstanceBatch([it[0] for it in items], cuda)
    xs = [torch.from_numpy(it[2]).to(cuda) for it in items]
    ybars = [-torch.ones((it[0].length, 1), device=cuda) for it in items]
    s = torch.cuda.Stream(device=cuda)
    results = []
    with torch.cuda.stream(s):
        for rep in range(3):
            scale = 1.0 - 0.05 * rep  # c <= 1 keeps dt = 0.707 inside the CFL bound
            cs = [torch.from_numpy(it[1]).to(cuda) * scale if rep else
                  torch.from_numpy(it[1]).to(cuda) for it in items]
            bt.set_inputs(cs, xs, ybars)
            bt.fwd_adj()  # eager reference for these inputs
            s.synchronize()
            eager = [(y.clone(), gc.clone()) for (y, gc, _) in bt.outputs()]
            if rep == 0:
                bt.capture()
            for g in bt.buf:  # poison the outputs so the replay must rewrite them
                g["y"].fill_(float("nan"))
                g["grad_c"].fill_(float("nan"))
            bt.replay()
            s.synchronize()
            for (y0, gc0), (y, gc, _) in zip(eager, bt.outputs()):
                assert torch.equal(y0, y)
                assert torch.equal(gc0, gc)
            results.append(eager[0][1].sum().item())
    assert len(set(results)) == 3  # the three inputs differ, so do the gradients


@pytest.mark.parametrize("parts", [2, 3, 4])
def test_batch_parts_match_single(parts, cuda):
    """The batch split into concurrent parts on the plan's own streams gives
    what the single launch sequence gives: probe series bit for bit, gradients
    to the oracle's tolerance.  Four parts over groups of three instances
    leave some parts without a group's instances."""
    items = _instances(seed=2)
    cs = [torch.from_numpy(it[1]).to(cuda) for it in items]
    xs = [torch.from_numpy(it[2]).to(cuda) for it in items]
    ybars = [-torch.ones((it[0].length, 1), device=cuda) for it in items]
    runs = []
    for st in (1, parts):
        bt = InstanceBatch([it[0] for it in items], cuda, tuning={"streams": st})
        print("case parts", st, bt.tuning[0], flush=True)
        assert bt.tuning[0]["streams"] <= st
        bt.set_inputs(cs, xs, ybars)
        bt.fwd_adj()
        torch.cuda.synchronize()
        runs.append([(y.clone(), gc.clone(), gx.clone()) for (y, gc, gx) in bt.outputs()])
        del bt
    assert len(runs[0]) == len(runs[1]) == len(items)
    for (inst, c, x), (y1, gc1, gx1), (yp, gcp, gxp) in zip(items, runs[0], runs[1]):
        assert torch.equal(y1, yp), inst.length
        _, gco, gxo = _oracle(inst, c, x, -np.ones((inst.length, 1)))
        assert rel(gcp.cpu().numpy(), gco) < 1e-4, inst.length
        assert rel(gxp.cpu().numpy(), gxo) < 1e-4, inst.length


@pytest.mark.parametrize("K", [64, 300])
def test_batch_parts_graph_replay(K, cuda):
    """A captured graph of a three-part batch (fork and join events, part
    streams) replays the eager result, with eager work between replays; K =
    300 is one segment (the forward's packed record, no recompute)."""
    items = _instances(seed=3)
    bt = InstanceBatch([it[0] for it in items], cuda, tuning={"streams": 3},
                       checkpoint_every=K)
    assert bt.tuning[0]["streams"] == 3
    xs = [torch.from_numpy(it[2]).to(cuda) for it in items]
    ybars = [-torch.ones((it[0].length, 1), device=cuda) for it in items]
    s = torch.cuda.Stream(device=cuda)
    with torch.cuda.stream(s):
        for rep in range(3):
            scale = 1.0 - 0.05 * rep
            cs = [torch.from_numpy(it[1]).to(cuda) * scale for it in items]
            bt.set_inputs(cs, xs, ybars)
            bt.fwd_adj()
            s.synchronize()
            eager = [(y.clone(), gc.clone(), gx.clone()) for (y, gc, gx) in bt.outputs()]
            if rep == 0:
                bt.capture()
            for g in bt.buf:
                g["y"].fill_(float("nan"))
                g["grad_c"].fill_(float("nan"))
            bt.replay()
            s.synchronize()
            for (y0, gc0, gx0), (y, gc, gx) in zip(eager, bt.outputs()):
                assert torch.equal(y0, y)
                assert torch.equal(gc0, gc)
                assert torch.equal(gx0, gx)


@pytest.mark.parametrize("parts,K", [(1, 64), (3, 64), (1, 300), (3, 300)])
def test_batch_instance_order(parts, K, cuda):
    """Instances longest first (launches cover only the instances with steps
    left) and in the given order (launches cover every instance of a group
    while its longest one runs) give the same results bit for bit; K = 300 is
    one segment, whose record the forward writes packed by instance length."""
    items = _instances(seed=4)
    cs = [torch.from_numpy(it[1]).to(cuda) for it in items]
    xs = [torch.from_numpy(it[2]).to(cuda) for it in items]
    ybars = [-torch.ones((it[0].length, 1), device=cuda) for it in items]
    runs = []
    for lf in (True, False):
        bt = InstanceBatch([it[0] for it in items], cuda, tuning={"streams": parts},
                           checkpoint_every=K, longest_first=lf)
        print("case order", lf, parts, K, [len(g) for g in bt.groups], flush=True)
        lens = [[bt.instances[i].length for i in idx] for idx in bt.groups]
        if lf:
            assert all(ln == sorted(ln, reverse=True) for ln in lens)
        else:
            assert any(ln != sorted(ln, reverse=True) for ln in lens)
        bt.set_inputs(cs, xs, ybars)
        bt.fwd_adj()
        torch.cuda.synchronize()
        runs.append([tuple(t.clone() for t in o) for o in bt.outputs()])
        del bt
    for a, b in zip(runs[0], runs[1]):
        for ta, tb in zip(a, b):
            assert torch.equal(ta, tb)
    for (inst, c, x), (y, gc, gx) in zip(items, runs[0]):
        yo, gco, gxo = _oracle(inst, c, x, -np.ones((inst.length, 1)))
        assert np.array_equal(y.cpu().numpy(), yo), inst.length
        assert rel(gc.cpu().numpy(), gco) < 1e-4, inst.length
        assert rel(gx.cpu().numpy(), gxo) < 1e-4, inst.length


def test_batch_forward_only_without_saved_workspace(cuda):
    """InstanceBatch(grad=False) runs the forward with no saved workspace and
    gives the probe series of the gradient-carrying batch bit for bit; the
    reverse sweep is refused."""
    items = _instances(seed=5)
    cs = [torch.from_numpy(it[1]).to(cuda) for it in items]
    xs = [torch.from_numpy(it[2]).to(cuda) for it in items]
    ys = []
    for grad in (True, False):
        bt = InstanceBatch([it[0] for it in items], cuda, checkpoint_every=300, grad=grad)
        print("case forward-only", grad, flush=True)
        assert (bt.saved is None) == (not grad)
        bt.set_inputs(cs, xs)
        bt.forward(with_grad=grad)
        torch.cuda.synchronize()
        ys.append([y.clone() for (y, _, _) in bt.outputs()])
        if not grad:
            with pytest.raises(RuntimeError):
                bt.fwd_adj()
        del bt
    for a, b in zip(*ys):
        assert torch.equal(a, b)
