"""GPU parity of the Holey geometry kernels (wt_geom_holey_c / _bwd,
geom.py:109-135) against the oracle's restatement (oracle/ref_torch.py
holey_rho / holey_c, pinned on the CPU by closed forms and central
differences): c and rho within rel-L2 1e-6 of the fp32 restatement and 1e-5
of fp64; dJ/d(x, y, r) within 1e-5 of fp64 autograd at a soft projection
(beta 4) and 1e-4 at the reference's default beta 100; NaN x/y gradients for
a centre on a grid point, as the reference gives; 0 and 64 holes."""
import numpy as np
import pytest
import torch

from oracle import ref_torch as rt
from tests.helpers import rel
from wavetorch_amd import WaveGeometryHoley

pytestmark = pytest.mark.gpu


def _holes(H, shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(3, shape[0] - 3, H) + 0.37
    y = rng.uniform(3, shape[1] - 3, H) + 0.29
    r = rng.uniform(1.5, 6.0, H)
    return x, y, r


def _rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("H,beta,gtol", [(3, 4.0, 1e-5), (16, 4.0, 1e-5), (64, 4.0, 1e-5),
                                         (5, 100.0, 1e-4)])
def test_holey_c_and_gradients(cuda, H, beta, gtol):
    shape = (70, 53)
    x, y, r = _holes(H, shape, H)
    kw = dict(c0=1.0, c1=0.5, abs_N=6, abs_sig=1.0, abs_p=2.0, eta=0.5, beta=beta)
    g = WaveGeometryHoley(shape, 1.0, x=x.astype(np.float32), y=y.astype(np.float32),
                          r=r.astype(np.float32), **kw).to(cuda)
    w = np.random.default_rng(7).standard_normal(shape)
    c = g.c
    (c * torch.from_numpy(w).float().to(cuda)).sum().backward()
    p32 = [torch.tensor(v.astype(np.float32)) for v in (x, y, r)]
    c32 = rt.holey_c(*p32, shape, 0.5, beta, 1.0, 0.5)
    assert _rel_l2(c.detach().cpu(), c32) <= 1e-6
    p64 = [torch.tensor(v.astype(np.float32).astype(np.float64), requires_grad=True)
           for v in (x, y, r)]
    c64 = rt.holey_c(*p64, shape, 0.5, beta, 1.0, 0.5)
    (c64 * torch.from_numpy(w)).sum().backward()
    assert _rel_l2(c.detach().cpu(), c64.detach()) <= 1e-5
    for k, p in zip(("x", "y", "r"), p64):
        assert rel(getattr(g, k).grad.cpu().numpy(), p.grad.numpy()) <= gtol, k
    # rho: the projected density, differentiable the same way
    rho = g.rho
    assert _rel_l2(rho.detach().cpu(), rt.holey_rho(*p32, shape, 0.5, beta)) <= 1e-6


def test_holey_on_grid_centre_and_no_holes(cuda):
    shape = (40, 30)
    g = WaveGeometryHoley(shape, 1.0, c0=1.0, c1=0.5, abs_N=5, eta=0.5, beta=4.0,
                          x=[12.0, 20.6], y=[9.0, 17.1], r=[3.0, 2.0]).to(cuda)
    g.c.sum().backward()
    assert torch.isnan(g.x.grad[0]) and torch.isnan(g.y.grad[0])
    assert torch.isfinite(g.x.grad[1]) and torch.isfinite(g.r.grad).all()
    g0 = WaveGeometryHoley(shape, 1.0, c0=1.0, c1=0.5, abs_N=5, eta=0.5, beta=4.0,
                           x=np.zeros(0, np.float32), y=np.zeros(0, np.float32),
                           r=np.zeros(0, np.float32)).to(cuda)
    c0 = g0.c.detach()  # P(0) = 0 up to rounding: c = c0
    assert float((c0.cpu() - 1.0).abs().max()) <= 1e-6
