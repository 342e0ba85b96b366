"""Writes tests/golden/reference_vectors.npz from the reference's own code.

Run in the build container only (the reference is not present on the GPU
machines):   python tests/golden/make_golden.py [/root/reference]

It loads the reference's ``wavetorch/operators.py``, ``utils.py``,
``cell.py`` and ``probe.py`` as modules of the ``wavetorch`` package without
running the package ``__init__`` (which imports scikit-image, librosa and
seaborn, absent here), and records input/output vectors of:

* ``_laplacian``                         operators.py:5-11
* ``_time_step``                         cell.py:12-17
* ``TimeStep.apply`` + its backward      cell.py:20-44
* ``saturable_damping``                  cell.py:8-9
* a 48-step loop in the order of rnn.py:50-70, calling the reference's
  ``TimeStep.apply`` and ``WaveIntensityProbe``/``WaveProbe`` forward and
  adding the source sample exactly as source.py:19-22 does, with the
  reference's autograd gradient of a weighted probe sum wrt c and x.

All vectors are fp32 as the reference computes them with
torch.get_default_dtype() == float32.
"""
from __future__ import annotations

import importlib
import os
import sys
import types

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "reference_vectors.npz")


def load_reference(root):
    pkg = types.ModuleType("wavetorch")
    pkg.__path__ = [os.path.join(root, "wavetorch")]
    sys.modules["wavetorch"] = pkg
    mods = {}
    for name in ("utils", "operators", "cell", "probe"):
        mods[name] = importlib.import_module(f"wavetorch.{name}")
    return mods


def pml(Nx, Ny, N, sig, p):
    # geom.py:63-85 restated (geom.py itself imports scikit-image)
    b_vals = sig * torch.linspace(0.0, 1.0, N + 1) ** p
    bx = torch.zeros(Nx, Ny)
    by = torch.zeros(Nx, Ny)
    bx[0:N + 1, :] = torch.flip(b_vals, [0]).repeat(Ny, 1).transpose(0, 1)
    bx[(Nx - N - 1):Nx, :] = b_vals.repeat(Ny, 1).transpose(0, 1)
    by[:, 0:N + 1] = torch.flip(b_vals, [0]).repeat(Nx, 1)
    by[:, (Ny - N - 1):Ny] = b_vals.repeat(Nx, 1)
    return torch.sqrt(bx ** 2 + by ** 2)


def main(root="/root/reference"):
    torch.set_default_dtype(torch.float32)
    m = load_reference(root)
    cellm, opm, probem = m["cell"], m["operators"], m["probe"]
    g = torch.Generator().manual_seed(20240601)
    out = {}

    # _laplacian
    u = torch.randn(2, 13, 17, generator=g)
    out["lap_in"] = u.numpy()
    out["lap_out"] = opm._laplacian(u, torch.tensor(1.0)).numpy()

    # single step, PML b, random c, dt = 0.707
    Nx, Ny, B = 24, 21, 3
    b = pml(Nx, Ny, 5, 1.0, 2)
    c = 0.5 + 0.5 * torch.rand(Nx, Ny, generator=g)
    y1 = torch.randn(B, Nx, Ny, generator=g)
    y2 = torch.randn(B, Nx, Ny, generator=g)
    dt = torch.tensor(0.707)
    h = torch.tensor(1.0)
    out.update(step_b=b.numpy(), step_c=c.numpy(), step_y1=y1.numpy(), step_y2=y2.numpy(),
               step_dt=np.float32(0.707))
    out["step_y"] = cellm._time_step(b, c, y1, y2, dt, h).numpy()
    bb = b.clone().requires_grad_(True)
    cc = c.clone().requires_grad_(True)
    yy1 = y1.clone().requires_grad_(True)
    yy2 = y2.clone().requires_grad_(True)
    gout = torch.randn(B, Nx, Ny, generator=g)
    y = cellm.TimeStep.apply(bb, cc, yy1, yy2, dt, h)
    y.backward(gout)
    out.update(step_g=gout.numpy(), step_gb=bb.grad.numpy(), step_gc=cc.grad.numpy(),
               step_gy1=yy1.grad.numpy(), step_gy2=yy2.grad.numpy())

    # saturable damping
    uu = torch.randn(B, Nx, Ny, generator=g)
    out["sat_u"] = uu.numpy()
    out["sat_out"] = cellm.saturable_damping(uu, torch.tensor(0.1 ** 0.5),
                                             torch.tensor(0.5)).numpy()

    # a 48-step loop in the order of rnn.py:50-70
    T = 48
    Nx, Ny, B = 30, 26, 2
    b = pml(Nx, Ny, 5, 1.0, 2)
    c = torch.ones(Nx, Ny)
    c[6:24, 6:20] = 0.5 + 0.5 * torch.rand(18, 14, generator=g)
    n = torch.arange(T, dtype=torch.float32)
    x = torch.stack([torch.sin(2 * np.pi * 0.05 * n + 0.4 * k) for k in range(B)])
    src = (9, 13)
    probes = [probem.WaveIntensityProbe(22, 8), probem.WaveIntensityProbe(22, 13),
              probem.WaveProbe(21, 18)]
    cg = c.clone().requires_grad_(True)
    xg = x.clone().requires_grad_(True)
    h1 = torch.zeros(B, Nx, Ny)
    h2 = torch.zeros(B, Nx, Ny)
    ys = []
    for i, xi in enumerate(xg.chunk(T, dim=1)):
        y = cellm.TimeStep.apply(b, cg, h1, h2, dt, h)
        h1, h2 = y, h1
        X_expanded = torch.zeros(h1.size()).detach()
        X_expanded[:, torch.tensor(src[0]), torch.tensor(src[1])] = xi.squeeze(-1)
        h1 = h1 + 1.0 ** 2 * X_expanded
        ys.append(torch.stack([p(h1) for p in probes], dim=-1))
    yall = torch.stack(ys, dim=1)
    w = torch.randn(yall.shape, generator=g)
    (yall * w).sum().backward()
    out.update(loop_b=b.numpy(), loop_c=c.numpy(), loop_x=x.numpy(), loop_w=w.numpy(),
               loop_y=yall.detach().numpy(), loop_gc=cg.grad.numpy(), loop_gx=xg.grad.numpy(),
               loop_src=np.array([[src[0], src[1], 0]], dtype=np.int32),
               loop_taps=np.array([[22, 8, 1], [22, 13, 1], [21, 18, 0]], dtype=np.int32),
               loop_dt=np.float32(0.707))
    out.update(nonlinear_loops(cellm, probem))
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


class _Geometry:
    """The three attributes WaveCell reads from its geometry (cell.py:66-67,
    94-105): b, h and cmax.  geom.py itself needs scikit-image, absent here."""

    def __init__(self, b, h, cmax):
        self.b, self.h, self.cmax = b, h, cmax

    def parameters(self):
        return iter(())


# (name, satdamp_b0, satdamp_uth, c_nl): the saturable-damping and Kerr-speed
# branches of cell.py:94-102 alone and together, at the nonlinear training
# configuration's values (b_nl = 0.5, I_sat = 0.1 -> uth = sqrt(0.1), c_nl = 0.05)
NL_CASES = (("nlsat", 0.5, 0.1 ** 0.5, 0.0), ("nlkerr", 0.0, 0.0, 0.05),
            ("nlboth", 0.5, 0.1 ** 0.5, 0.05))


def nonlinear_loops(cellm, probem):
    """A 48-step loop of the reference's own WaveCell.forward (cell.py:79-107)
    with the nonlinear terms on, driven in the order of rnn.py:39-70: u_0 =
    u_-1 = 0; per step (h1, h2) = cell(h1, h2, c_linear, rho), then the source
    sample added as source.py:19-22 does, then the probes.  rho is the raw
    density (rnn.py:47 passes geom.rho).  Recorded per case: the probe series,
    the last two fields, and the reference's autograd gradients of a weighted
    probe sum with respect to c_linear, the PML field b, rho and x."""
    g = torch.Generator().manual_seed(20240602)
    T, Nx, Ny, B = 48, 30, 26, 2
    dt = 0.6
    b0 = pml(Nx, Ny, 5, 1.0, 2)
    c = torch.ones(Nx, Ny)
    c[6:24, 6:20] = 0.5 + 0.5 * torch.rand(18, 14, generator=g)
    rho = torch.zeros(Nx, Ny)
    rho[6:24, 6:20] = torch.rand(18, 14, generator=g)
    n = torch.arange(T, dtype=torch.float32)
    x = torch.stack([1.5 * torch.sin(2 * np.pi * 0.05 * n + 0.4 * k) for k in range(B)])
    src = (9, 13)
    probes = [probem.WaveIntensityProbe(22, 8), probem.WaveIntensityProbe(22, 13),
              probem.WaveProbe(21, 18)]
    w = torch.randn(B, T, len(probes), generator=g)
    out = dict(nl_b=b0.numpy(), nl_c=c.numpy(), nl_rho=rho.numpy(), nl_x=x.numpy(),
               nl_w=w.numpy(), nl_dt=np.float32(dt),
               nl_src=np.array([[src[0], src[1], 0]], dtype=np.int32),
               nl_taps=np.array([[22, 8, 1], [22, 13, 1], [21, 18, 0]], dtype=np.int32))
    for name, sb0, suth, cnl in NL_CASES:
        bg = b0.clone().requires_grad_(True)
        cg = c.clone().requires_grad_(True)
        rg = rho.clone().requires_grad_(True)
        xg = x.clone().requires_grad_(True)
        cell = cellm.WaveCell(dt, _Geometry(bg, torch.tensor(1.0), 1.0), satdamp_b0=sb0,
                              satdamp_uth=suth, c_nl=cnl)
        h1 = torch.zeros(B, Nx, Ny)
        h2 = torch.zeros(B, Nx, Ny)
        ys = []
        for xi in xg.chunk(T, dim=1):
            h1, h2 = cell(h1, h2, cg, rg)
            X_expanded = torch.zeros(h1.size()).detach()
            X_expanded[:, torch.tensor(src[0]), torch.tensor(src[1])] = xi.squeeze(-1)
            h1 = h1 + 1.0 ** 2 * X_expanded
            ys.append(torch.stack([p(h1) for p in probes], dim=-1))
        yall = torch.stack(ys, dim=1)
        (yall * w).sum().backward()
        out.update({f"{name}_y": yall.detach().numpy(), f"{name}_u1": h1.detach().numpy(),
                    f"{name}_u2": h2.detach().numpy(), f"{name}_gb": bg.grad.numpy(),
                    f"{name}_gc": cg.grad.numpy(), f"{name}_grho": rg.grad.numpy(),
                    f"{name}_gx": xg.grad.numpy(),
                    f"{name}_par": np.array([sb0, suth, cnl], dtype=np.float64)})
    return out


if __name__ == "__main__":
    main(*sys.argv[1:])
