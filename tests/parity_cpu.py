"""CPU legs of tests/parity_report.py: numpy and the C oracle only.

Test infrastructure.  This module imports neither torch nor the HIP library,
so the worker processes that run these functions never load the GPU stack.

A problem `spec` is a plain dict: b (Nx,Ny) fp64, x (B,T) fp64, dt, h, src
and taps (lists of [x, y, k]), labels (B,) int or None, nl (None or dict of
sat_b0, sat_uth, c_nl, clamp) and, for free-form geometries, geom (c0, c1,
eta, beta, blur_N, kernel (1,1,k,k) fp64).
"""
from __future__ import annotations

import numpy as np

import oracle.c_oracle as co


def c_map(sp, rho):
    """rho -> c (geom.py:207-233): blur_N passes of the disk blur (conv2d
    with zero padding, operators as in torch's cross-correlation), then the
    tanh projection, in fp64."""
    gm = sp["geom"]
    k = np.asarray(gm["kernel"], np.float64).reshape(gm["kernel"].shape[-2:])
    r = k.shape[0] // 2
    Nx, Ny = rho.shape
    out = np.asarray(rho, np.float64)
    for _ in range(gm["blur_N"]):
        pad = np.zeros((Nx + 2 * r, Ny + 2 * r))
        pad[r:r + Nx, r:r + Ny] = out
        acc = np.zeros((Nx, Ny))
        for a in range(k.shape[0]):
            for b in range(k.shape[1]):
                if k[a, b] != 0:
                    acc += k[a, b] * pad[a:a + Nx, b:b + Ny]
        out = acc
    eta, beta = gm["eta"], gm["beta"]
    proj = (np.tanh(beta * eta) + np.tanh(beta * (out - eta))) / \
        (np.tanh(beta * eta) + np.tanh(beta * (1 - eta)))
    return gm["c0"] + (gm["c1"] - gm["c0"]) * proj


def ce_loss(sp, y):
    """Cross-entropy of the normalised integrated intensities (utils.py:35-36
    then torch's cross_entropy, mean over the batch) in fp64."""
    s = np.asarray(y, np.float64).sum(axis=1)
    p = s / s.sum(axis=1, keepdims=True)
    m = p.max(axis=1, keepdims=True)
    lse = (m + np.log(np.exp(p - m).sum(axis=1, keepdims=True)))[:, 0]
    lab = sp["labels"]
    return float(np.mean(lse - p[np.arange(len(lab)), lab]))


def _nl(sp, rho):
    return None if sp["nl"] is None else dict(rho=rho, **sp["nl"])


def task_fwd64(sp, rho64, samples, c_given=None):
    """fp64 probe series of `samples` (c from the fp64 map unless given)."""
    c = c_map(sp, rho64) if c_given is None else c_given
    y, _ = co.forward(sp["b"], c, sp["x"][samples], sp["dt"], sp["h"], sp["src"], sp["taps"],
                      nl=_nl(sp, rho64), dtype=np.float64)
    return y


def task_fwd32(sp, c32, rho32, samples, stride=0):
    """The fp32 oracle loop given the GPU's c: probe series, or the fields
    after every `stride`-th step."""
    y, f = co.forward(sp["b"].astype(np.float32), c32, sp["x"][samples].astype(np.float32),
                      sp["dt"], sp["h"], sp["src"], sp["taps"], nl=_nl(sp, rho32),
                      fields=stride > 0, field_stride=max(stride, 1))
    return f if stride > 0 else y


def task_fields64(sp, c64, stride):
    _, f = co.forward(sp["b"], c64, sp["x"], sp["dt"], sp["h"], sp["src"], sp["taps"],
                      fields=True, field_stride=stride, dtype=np.float64)
    return f


def task_adj64(sp, rho64, ybar, samples, c_given=None):
    """fp64 adjoint of `samples`: (dJ/dc summed, direct dJ/drho summed, dJ/dx)."""
    c = c_map(sp, rho64) if c_given is None else c_given
    g = co.adjoint(sp["b"], c, sp["x"][samples], sp["dt"], sp["h"], sp["src"], sp["taps"],
                   ybar=ybar[samples], nl=_nl(sp, rho64))
    return g["gc"], g["grho"], g["gx"]


def task_J64(sp, rho64):
    """fp64 loss of the whole batch at `rho64` (one central-difference leg)."""
    return ce_loss(sp, task_fwd64(sp, rho64, slice(None)))


def task_lens(sp, cg, ybar):
    """One batched instance: fp32 probe series given the GPU's c; fp64 series
    and adjoint given the same c (upcast)."""
    y32, _ = co.forward(sp["b"].astype(np.float32), cg, sp["x"].astype(np.float32), sp["dt"],
                        sp["h"], sp["src"], sp["taps"])
    c64 = cg.astype(np.float64)
    y64, _ = co.forward(sp["b"], c64, sp["x"], sp["dt"], sp["h"], sp["src"], sp["taps"],
                        dtype=np.float64)
    g = co.adjoint(sp["b"], c64, sp["x"], sp["dt"], sp["h"], sp["src"], sp["taps"], ybar=ybar)
    return y32, y64, g["gc"], g["gx"]


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a))


def task_lens_fwd32(sp, cg, yg):
    """One batched instance: is the GPU's probe series `yg` (T, P) bit for bit
    the fp32 oracle's, given the GPU's c?"""
    y32, _ = co.forward(sp["b"].astype(np.float32), cg, sp["x"].astype(np.float32), sp["dt"],
                        sp["h"], sp["src"], sp["taps"])
    return bool(np.array_equal(yg, y32[0]))


def task_lens_check(sp, cg, yg, gcg, gxg):
    """One batched instance with J = -sum of its focal intensity: the forward
    bit for bit against the fp32 oracle given the GPU's c, and the GPU's
    dJ/dc, dJ/dx relative to the fp64 adjoint taken at the fp32 trajectory
    (`_at`: the exact adjoint of the computed forward) and at the fp64
    forward's (`64`)."""
    T = sp["x"].shape[1]
    ybar = -np.ones((1, T, 1))
    y32, f32 = co.forward(sp["b"].astype(np.float32), cg, sp["x"].astype(np.float32), sp["dt"],
                          sp["h"], sp["src"], sp["taps"], fields=True)
    c64 = cg.astype(np.float64)
    g_at = co.adjoint(sp["b"], c64, sp["x"], sp["dt"], sp["h"], sp["src"], sp["taps"],
                      ybar=ybar, at_fields=f32)
    del f32
    g64 = co.adjoint(sp["b"], c64, sp["x"], sp["dt"], sp["h"], sp["src"], sp["taps"], ybar=ybar)
    return dict(bitwise=bool(np.array_equal(yg, y32[0])),
                gc_at=_rel(gcg, g_at["gc"]), gx_at=_rel(gxg, g_at["gx"][0]),
                gc64=_rel(gcg, g64["gc"]), gx64=_rel(gxg, g64["gx"][0]))
