"""Host-side checks of tests/parity_report.py's CPU legs on a small problem:
the numpy rho -> c map and loss equal the reference's torch ops in fp64, and
the fp64 oracle chain rho -> c -> loop -> loss -> adjoint -> dJ/drho agrees
with central differences of the fp64 loss and with the reference's fp32 op
sequence differentiated by autograd, so the report's reference values are
themselves pinned."""
import numpy as np
import torch

from oracle import ref_torch as rt
from tests import parity_cpu as pc
from tests import parity_report as pr


def _problem(nonlinear=False):
    from wavetorch_amd import (WaveCell, WaveGeometryFreeForm, WaveIntensityProbe, WaveRNN,
                               WaveSource)
    g = torch.Generator().manual_seed(5)
    N, pml, T, B = 28, 6, 90, 3
    geom = WaveGeometryFreeForm((N, N), 1.0, c0=1.0, c1=0.5, abs_N=pml, abs_sig=1.0, abs_p=2.0,
                                eta=0.5, beta=4.0, rho=torch.rand(N, N, generator=g))
    kw = dict(satdamp_b0=0.5, satdamp_uth=0.3, c_nl=0.05, c_clamp=(0.4, 1.1)) if nonlinear \
        else {}
    model = WaveRNN(WaveCell(0.6, geom, **kw), WaveSource(pml + 3, N // 2),
                    [WaveIntensityProbe(N - pml - 4, 9), WaveIntensityProbe(N - pml - 4, 18)])
    n = np.arange(T)
    x = torch.from_numpy(np.stack([np.sin(0.3 * n + k) * np.exp(-((n - 40) / 15.0) ** 2)
                                   for k in range(B)]).astype(np.float32))
    labels = torch.tensor([0, 1, 1])
    return model, x, labels


def _check(nonlinear):
    model, x, labels = _problem(nonlinear)
    sp = pr.spec_of(model, x, labels)
    rho64 = model.cell.geom.rho.detach().numpy().astype(np.float64)
    pool = pr.make_pool(2)
    try:
        y64, J64, grho64 = pr.grad64_chain(sp, rho64, pool, 2)
        pixels = pr.fd_pixels(sp, n=8)
        fd = pr.fd_collect(pr.fd_submit(sp, rho64, pixels, pool))
    finally:
        pool.terminate()
        pool.join()
    opx = np.array([grho64[i, j] for (i, j) in pixels])
    assert pr.rel(opx, fd) < 1e-6
    yr, grho_r = pr.ref32_grad(model, x, labels)
    assert pr.rel(yr, y64) < 1e-5
    assert pr.rel(grho_r, grho64) < 1e-3  # fp32 autograd over 90 steps


def test_chain_vs_fd_and_reference_linear():
    _check(False)


def test_chain_vs_fd_and_reference_nonlinear():
    _check(True)


def test_numpy_c_map_and_loss_match_reference_ops():
    model, x, labels = _problem()
    sp = pr.spec_of(model, x, labels)
    g = model.cell.geom
    for blur_N in (0, 1, 2):
        sp["geom"]["blur_N"] = blur_N
        rho = np.random.default_rng(blur_N).uniform(0, 1, sp["b"].shape)
        ct = rt.freeform_c(torch.from_numpy(rho), g.c0.item(), g.c1.item(),
                           g.blur_kernel.double(), blur_N, g.eta.item(), g.beta.item())
        assert np.allclose(pc.c_map(sp, rho), ct.numpy(), rtol=0, atol=1e-14)
    y = np.random.default_rng(7).uniform(0.1, 1.0, (3, 40, 2))
    assert abs(pc.ce_loss(sp, y) - pr.loss64(sp, y)[0]) < 1e-14


def test_lens_instance_reference_vs_oracle():
    """The report's per-instance legs on a small lens problem: the reference's
    fp32 gradient of J = -(focal intensity summed) against the fp64 oracle
    given the same c, and the fp32 oracle series against the reference's."""
    from wavetorch_amd import workloads
    m, x = workloads.lens_instance(40, 160, seed=9)
    g = m.cell.geom
    with torch.no_grad():  # the product's c runs on the GPU only
        c = rt.freeform_c(g.rho, g.c0.item(), g.c1.item(), g.blur_kernel, int(g.blur_N.item()),
                          g.eta.item(), g.beta.item())
    sp = pr.spec_of(m, x[None], None)
    y32, y64, gc64, gx64 = pc.task_lens(sp, c.numpy(), -np.ones((1, 160, 1)))
    gcr, gxr = pr.ref32_instance(m, x, c)
    assert pr.rel(gcr, gc64) < 1e-4 and pr.rel(gxr, gx64) < 1e-4
    assert pr.rel(y32, y64) < 1e-5
