"""Torch-CPU restatement of the reference's WaveRNN hot path, op for op.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``.  The product package never imports
it.

Every function follows the reference (fancompute/wavetorch v0.2.1) operation
by operation, so that on the same torch build it produces what the reference
produces, including its rounding, its autograd tape and its per-step
allocations:

* ``laplacian``         operators.py:5-11   conv2d, padding=1
* ``time_step``         cell.py:12-17
* ``TimeStepRef``       cell.py:20-44       hand-written backward
* ``saturable_damping`` cell.py:8-9
* ``cell_forward``      cell.py:79-107
* ``source_forward``    source.py:15-22     fresh zeros + index_put per step
* ``probe_forward``     probe.py:14-15, 26-27
* ``rnn_forward``       rnn.py:21-72
* ``pml_b``             geom.py:63-85
* ``blur_kernel``, ``apply_blur``, ``apply_projection``, ``freeform_c``
                        geom.py:149-154, 207-233
* ``holey_rho``, ``holey_c``  geom.py:109-135
* ``normalize_power``   utils.py:35-36

The dtype follows ``torch.get_default_dtype()`` exactly as the reference
does (utils.py:6-20); callers pass ``dtype`` explicitly instead.
"""
from __future__ import annotations

import numpy as np
import torch
from torch.nn.functional import conv2d


def laplacian(y, h):
    """operators.py:5-11 -- kernel tensor rebuilt on every call."""
    operator = h ** (-2) * torch.tensor([[[[0.0, 1.0, 0.0], [1.0, -4.0, 1.0], [0.0, 1.0, 0.0]]]],
                                        dtype=y.dtype)
    y = y.unsqueeze(1)
    return conv2d(y, operator, padding=1).squeeze(1)


def time_step(b, c, y1, y2, dt, h):
    """cell.py:12-17."""
    y = torch.mul((dt ** -2 + b * dt ** -1).pow(-1),
                  (2 / dt ** 2 * y1 - torch.mul((dt ** -2 - b * dt ** -1), y2)
                   + torch.mul(c.pow(2), laplacian(y1, h))))
    return y


class TimeStepRef(torch.autograd.Function):
    """cell.py:20-44 (forward saves inputs; backward is the hand-derived one)."""

    @staticmethod
    def forward(ctx, b, c, y1, y2, dt, h):
        ctx.save_for_backward(b, c, y1, y2, dt, h)
        return time_step(b, c, y1, y2, dt, h)

    @staticmethod
    def backward(ctx, grad_output):
        b, c, y1, y2, dt, h = ctx.saved_tensors
        grad_b = grad_c = grad_y1 = grad_y2 = None
        if ctx.needs_input_grad[0]:
            grad_b = - (dt * b + 1).pow(-2) * dt * (
                c.pow(2) * dt ** (2) * laplacian(y1, h) + 2 * y1 - 2 * y2) * grad_output
        if ctx.needs_input_grad[1]:
            grad_c = (b * dt + 1).pow(-1) * (2 * c * dt ** (2) * laplacian(y1, h)) * grad_output
        if ctx.needs_input_grad[2]:
            c2_grad = (b * dt + 1) ** (-1) * c.pow(2) * grad_output
            grad_y1 = dt ** (2) * laplacian(c2_grad, h) + 2 * grad_output * (b * dt + 1).pow(-1)
        if ctx.needs_input_grad[3]:
            grad_y2 = (b * dt - 1) * (b * dt + 1).pow(-1) * grad_output
        return grad_b, grad_c, grad_y1, grad_y2, None, None


def saturable_damping(u, uth, b0):
    """cell.py:8-9."""
    return b0 / (1 + torch.abs(u / uth).pow(2))


def cell_forward(h1, h2, c_linear, rho, *, b_geom, dt, h, satdamp_b0, satdamp_uth, c_nl,
                 c_clamp=None):
    """cell.py:79-107.  ``c_clamp=(cmin, cmax)`` is an extension (not in the reference)."""
    if satdamp_b0 > 0:
        b = b_geom + rho * saturable_damping(h1, uth=satdamp_uth, b0=satdamp_b0)
    else:
        b = b_geom
    if c_nl != 0:
        c = c_linear + rho * c_nl * h1.pow(2)
    else:
        c = c_linear
    if c_clamp is not None:
        c = torch.clamp(c, c_clamp[0], c_clamp[1])
    y = TimeStepRef.apply(b, c, h1, h2, dt, h)
    return y, h1


def source_forward(Y, X, sx, sy, dt=1.0):
    """source.py:15-22: a fresh full-size zeros tensor per call."""
    X_expanded = torch.zeros(Y.size(), dtype=Y.dtype).detach()
    X_expanded[:, sx, sy] = X
    return Y + dt ** 2 * X_expanded


def probe_forward(u, px, py, intensity):
    """probe.py:14-15 (WaveProbe) and 26-27 (WaveIntensityProbe)."""
    v = u[:, px, py]
    return v.pow(2) if intensity else v


def rnn_forward(x, *, c, rho, b_geom, dt, h, sources, probes, output_fields=False,
                satdamp_b0=None, satdamp_uth=None, c_nl=None, c_clamp=None):
    """rnn.py:21-72.

    ``sources``: list of (sx, sy) index tensors (int64, 0-dim or 1-dim).
    ``probes``: list of (px, py, intensity: bool).
    Nonlinear parameters default to zero tensors like ``WaveCell`` buffers (cell.py:60-64).
    """
    dtype = x.dtype
    z = torch.tensor(0.0, dtype=dtype)
    satdamp_b0 = z if satdamp_b0 is None else satdamp_b0
    satdamp_uth = z if satdamp_uth is None else satdamp_uth
    c_nl = z if c_nl is None else c_nl
    batch_size = x.shape[0]
    hidden_state_shape = (batch_size,) + tuple(c.shape)
    h1 = torch.zeros(hidden_state_shape, dtype=dtype)
    h2 = torch.zeros(hidden_state_shape, dtype=dtype)
    y_all = []
    for i, xi in enumerate(x.chunk(x.size(1), dim=1)):
        h1, h2 = cell_forward(h1, h2, c, rho, b_geom=b_geom, dt=dt, h=h,
                              satdamp_b0=satdamp_b0, satdamp_uth=satdamp_uth, c_nl=c_nl,
                              c_clamp=c_clamp)
        for (sx, sy) in sources:
            h1 = source_forward(h1, xi.squeeze(-1), sx, sy)
        if len(probes) > 0 and not output_fields:
            probe_values = [probe_forward(h1, px, py, it) for (px, py, it) in probes]
            y_all.append(torch.stack(probe_values, dim=-1))
        else:
            y_all.append(h1)
    return torch.stack(y_all, dim=1)


# ---------------------------------------------------------------- geometry


def pml_b(domain_shape, abs_N, abs_sig, abs_p, dtype=torch.float32):
    """geom.py:63-85 (abs_p arrives as the uint8 buffer value, geom.py:26)."""
    Nx, Ny = domain_shape
    b_vals = abs_sig * torch.linspace(0.0, 1.0, abs_N + 1, dtype=dtype) ** abs_p
    b_x = torch.zeros(Nx, Ny, dtype=dtype)
    b_y = torch.zeros(Nx, Ny, dtype=dtype)
    if abs_N > 0:
        b_x[0:abs_N + 1, :] = torch.flip(b_vals, [0]).repeat(Ny, 1).transpose(0, 1)
        b_x[(Nx - abs_N - 1):Nx, :] = b_vals.repeat(Ny, 1).transpose(0, 1)
        b_y[:, 0:abs_N + 1] = torch.flip(b_vals, [0]).repeat(Nx, 1)
        b_y[:, (Ny - abs_N - 1):Ny] = b_vals.repeat(Nx, 1)
    return torch.sqrt(b_x ** 2 + b_y ** 2)


def disk_mask(radius):
    """Pixels selected by ``skimage.draw.circle(r, r, r+1)`` (geom.py:149).

    scikit-image's rule: (dr/R)^2 + (dc/R)^2 < 1 with R = radius+1, strict,
    over the (2r+1)^2 box centred at (r, r).  (scikit-image is not
    installed here; this is the published rule.)
    """
    R = radius + 1
    k = np.zeros((2 * radius + 1, 2 * radius + 1), dtype=bool)
    for i in range(2 * radius + 1):
        for j in range(2 * radius + 1):
            if ((i - radius) / R) ** 2 + ((j - radius) / R) ** 2 < 1.0:
                k[i, j] = True
    return k


def blur_kernel(radius, dtype=torch.float32):
    """geom.py:149-154."""
    m = disk_mask(radius)
    k = torch.zeros((2 * radius + 1, 2 * radius + 1), dtype=dtype)
    k[torch.from_numpy(m)] = 1
    k = k / k.sum()
    return k.unsqueeze(0).unsqueeze(0)


def apply_blur(rho, kernel, blur_N):
    """geom.py:207-215."""
    for _ in range(blur_N):
        rho = conv2d(rho.unsqueeze(0).unsqueeze(0), kernel, padding=kernel.shape[-1] // 2).squeeze()
    return rho


def apply_projection(rho, eta, beta):
    """geom.py:217-222."""
    return (np.tanh(beta * eta) + torch.tanh(beta * (rho - eta))) / (
        np.tanh(beta * eta) + np.tanh(beta * (1 - eta)))


def freeform_c(rho, c0, c1, kernel, blur_N, eta, beta):
    """geom.py:224-233."""
    r = apply_projection(apply_blur(rho, kernel, blur_N), eta, beta)
    return c0 + (c1 - c0) * r


def holey_rho(hx, hy, hr, domain_shape, eta, beta):
    """geom.py:109-128: the projected density of the Holey geometry,
    P(sum_h exp(-|p - p_h| / r_h)), holes accumulated in order; dtype and
    device follow the hole tensors."""
    dt, dev = hx.dtype, hx.device
    xv = torch.arange(0, domain_shape[0], dtype=dt, device=dev)
    yv = torch.arange(0, domain_shape[1], dtype=dt, device=dev)
    x, y = torch.meshgrid(xv, yv, indexing="ij")
    rho = torch.zeros(domain_shape, dtype=dt, device=dev)
    for ri, xi, yi in zip(hr, hx, hy):
        r = torch.sqrt((x - xi).pow(2) + (y - yi).pow(2))
        rho = rho + torch.exp(-r / ri)
    return apply_projection(rho, eta, beta)


def holey_c(hx, hy, hr, domain_shape, eta, beta, c0, c1):
    """geom.py:133-135."""
    return c0 + (c1 - c0) * holey_rho(hx, hy, hr, domain_shape, eta, beta)


def normalize_power(X):
    """utils.py:35-36."""
    return X / torch.sum(X, dim=1, keepdim=True)
