"""GPU tests of the training-step caller of the time loop (reference:
train.py:13-133, utils.py:35-36).

* The fused loss head (wt_loss_power_ce through loss.power_cross_entropy)
  against the reference's criterion composition
  cross_entropy(normalize_power(y.sum(dim=1)), labels) in float64 on the
  CPU: loss and dJ/dy within 1e-6 relative (fp32 inputs, fp64 sums in the
  kernel), over batch/probe/length edge cases, with the oracle's numpy
  restatement of the same loss as a second check.
* GraphedTrainStep: three replays with new batches, each equal bit for bit
  to the eager step (same kernels in the same order) run between them, for
  the linear and the nonlinear cell and the Holey geometry.
* train(): the reference's epoch loop and history layout on a small model,
  epoch 0 without an optimizer step, the fused criterion's epoch-0 loss equal
  to the torch composition, and a saved file that loads back.
"""
import numpy as np
import pytest
import torch

from tests.helpers import rel
from tests.parity_cpu import ce_loss
from wavetorch_amd import (GraphedTrainStep, WaveCell, WaveGeometryFreeForm,
                           WaveGeometryHoley, WaveIntensityProbe, WaveRNN, WaveSource, train)
from wavetorch_amd.loss import power_cross_entropy
from wavetorch_amd.utils import accuracy_onehot, normalize_power

pytestmark = pytest.mark.gpu


def _composition64(y, labels):
    y64 = y.detach().cpu().double().requires_grad_(True)
    loss = torch.nn.functional.cross_entropy(normalize_power(y64.sum(dim=1)), labels.cpu())
    loss.backward()
    return loss.detach(), y64.grad


@pytest.mark.parametrize("B,T,P", [(8, 500, 3), (64, 3000, 3), (5, 37, 7), (1, 1, 2),
                                   (3, 11, 1), (2, 40, 256), (300, 20, 4)])
def test_power_cross_entropy_matches_reference_criterion(cuda, B, T, P):
    g = torch.Generator().manual_seed(B * 1000 + T + P)
    y = (torch.rand((B, T, P), generator=g) * 0.02 + 1e-4).to(cuda).requires_grad_(True)
    labels = torch.randint(0, P, (B,), generator=g).to(cuda)
    loss = power_cross_entropy(y, labels)
    loss.backward()
    want, gwant = _composition64(y, labels)
    assert abs(loss.item() - want.item()) <= 1e-6 * max(abs(want.item()), 1e-6)
    assert rel(y.grad.cpu().double(), gwant) <= 1e-6
    sp = {"labels": labels.cpu().numpy()}
    assert abs(loss.item() - ce_loss(sp, y.detach().cpu().double().numpy())) <= \
        1e-6 * max(abs(want.item()), 1e-6)


def test_power_cross_entropy_rejects_and_flags(cuda):
    y = torch.rand((4, 10, 3), device=cuda)
    bad = torch.tensor([0, 3, 1, 2], device=cuda)  # 3 is not a class of 3 probes
    assert torch.isnan(power_cross_entropy(y, bad)).item()
    with pytest.raises(RuntimeError, match="at most 256"):
        power_cross_entropy(torch.rand((2, 5, 257), device=cuda), torch.zeros(2, dtype=torch.int64,
                                                                                device=cuda))
    with pytest.raises(ValueError):
        power_cross_entropy(torch.rand((2, 5), device=cuda), torch.zeros(2, device=cuda))


def _model(cuda, seed, nonlinear=False, holey=False):
    torch.manual_seed(seed)
    if holey:
        geom = WaveGeometryHoley((60, 50), 1.0, c0=1.0, c1=0.5, abs_N=10, abs_sig=1.0,
                                 abs_p=2.0, eta=0.5, beta=4.0, x=[25.3, 38.6], y=[17.4, 31.1],
                                 r=[3.0, 2.5])
    else:
        geom = WaveGeometryFreeForm((60, 50), 1.0, c0=1.0, c1=0.5, abs_N=10, abs_sig=1.0,
                                    abs_p=2.0, rho="rand")
    nl = dict(satdamp_b0=0.5, satdamp_uth=0.3, c_nl=0.05, c_clamp=(0.4, 1.1)) if nonlinear \
        else {}
    m = WaveRNN(WaveCell(0.5, geom, **nl), [WaveSource(20, 25)],
                [WaveIntensityProbe(40, 10), WaveIntensityProbe(40, 25),
                 WaveIntensityProbe(40, 40)])
    return m.to(cuda)


def _eager(m, opt, x, labels):
    opt.zero_grad(set_to_none=True)
    loss = power_cross_entropy(m(x), labels)
    loss.backward()
    opt.step()
    m.cell.geom.constrain_to_design_region()
    return loss


@pytest.mark.parametrize("nonlinear,holey", [(False, False), (True, False), (False, True)])
def test_graphed_train_step_equals_eager(cuda, nonlinear, holey):
    B, T = 4, 200
    ma, mb = _model(cuda, 7, nonlinear, holey), _model(cuda, 7, nonlinear, holey)
    pa = next(iter(ma.parameters()))
    pb = next(iter(mb.parameters()))
    oa = torch.optim.Adam(ma.parameters(), lr=1e-2, capturable=True)
    ob = torch.optim.Adam(mb.parameters(), lr=1e-2, capturable=True)
    g = torch.Generator().manual_seed(3)
    x0 = torch.randn((B, T), generator=g).to(cuda)
    l0 = torch.randint(0, 3, (B,), generator=g).to(cuda)
    p0 = pb.detach().clone()
    step = GraphedTrainStep(mb, ob, x0, l0, warmup=2)
    # building the step leaves the model as it was
    assert torch.equal(pb.detach(), p0)
    for i in range(3):
        x = torch.randn((B, T), generator=g).to(cuda)
        labels = torch.randint(0, 3, (B,), generator=g).to(cuda)
        la = _eager(ma, oa, x, labels)
        lb = step(x, labels)
        torch.cuda.synchronize()
        assert la.item() == lb.item(), (i, la.item(), lb.item())
        for qa, qb in zip(ma.parameters(), mb.parameters()):
            assert torch.equal(qa.detach(), qb.detach()), i
        sa, sb = oa.state[pa], ob.state[pb]
        assert torch.equal(sa["exp_avg"], sb["exp_avg"]) and \
            torch.equal(sa["exp_avg_sq"], sb["exp_avg_sq"]), i
    assert not torch.equal(pb.detach(), p0)
    with pytest.raises(ValueError, match="captured for"):
        step(torch.zeros((B + 1, T), device=cuda), torch.zeros(B + 1, device=cuda))


def test_graphed_train_step_needs_capturable_adam(cuda):
    m = _model(cuda, 1)
    x = torch.randn((2, 50), device=cuda)
    l = torch.zeros(2, dtype=torch.int64, device=cuda)
    with pytest.raises(ValueError, match="capturable"):
        GraphedTrainStep(m, torch.optim.Adam(m.parameters(), lr=1e-3), x, l)


def test_train_loop_history_and_save(cuda, tmp_path):
    from torch.utils.data import DataLoader, TensorDataset
    from wavetorch_amd.io import load_model
    m = _model(cuda, 11)
    g = torch.Generator().manual_seed(5)
    X = (torch.randn((8, 150), generator=g) * 0.5).to(cuda)
    Y = torch.nn.functional.one_hot(torch.randint(0, 3, (8,), generator=g), 3).float().to(cuda)
    dl = DataLoader(TensorDataset(X, Y), batch_size=4)
    with torch.no_grad():
        want0 = np.mean([torch.nn.functional.cross_entropy(
            normalize_power(m(xb).sum(dim=1)), yb.argmax(dim=1)).item() for xb, yb in dl])
    rho0 = m.cell.geom.rho.detach().clone()
    opt = torch.optim.Adam(m.parameters(), lr=1e-2)
    hist, hms = train(m, opt, torch.nn.CrossEntropyLoss(), dl, dl, N_epochs=1, batch_size=4,
                      accuracy=accuracy_onehot, name="run", savedir=str(tmp_path) + "/",
                      cfg={"dtype": "float32"})
    assert list(hist.columns) == ['time', 'epoch', 'fold', 'loss_train', 'loss_test',
                                  'acc_train', 'acc_test', 'cm_train', 'cm_test']
    assert list(hist["epoch"]) == [0, 1] and len(hms) == 2
    assert abs(hist["loss_train"][0] - want0) <= 1e-6 * want0
    assert torch.equal(hms[0]["rho"].to(cuda), rho0)        # epoch 0 takes no step
    assert not torch.equal(hms[1]["rho"].to(cuda), rho0)
    assert hist["cm_train"][1].sum() == 8 and hist["cm_test"][1].sum() == 8
    m2, h2, hgs, cfg = load_model(str(tmp_path / "run.pt"), weights_only=False)
    assert len(h2) == 2 and cfg == {"dtype": "float32"}
    assert torch.equal(m2.cell.geom.rho.detach(), m.cell.geom.rho.detach().cpu())
