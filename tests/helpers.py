"""Shared problem builders for the tests (seeded, small)."""
from __future__ import annotations

import numpy as np
import torch

from oracle import ref_torch


def small_problem(Nx=40, Ny=36, B=2, T=120, abs_N=5, seed=0, amp=1.0):
    """A C0-like problem at small size: PML ramp, random c in the interior,
    one point source with a windowed sine, three probes (two intensity, one
    amplitude)."""
    rng = np.random.default_rng(seed)
    b = ref_torch.pml_b((Nx, Ny), abs_N, 1.0, 2).numpy()
    c = np.ones((Nx, Ny), dtype=np.float32)
    c[abs_N + 1:Nx - abs_N - 1, abs_N + 1:Ny - abs_N - 1] = rng.uniform(
        0.5, 1.0, (Nx - 2 * abs_N - 2, Ny - 2 * abs_N - 2)).astype(np.float32)
    n = np.arange(T)
    x = np.stack([amp * np.sin(2 * np.pi * 0.05 * n + 0.3 * k) *
                  np.exp(-((n - T / 3) / (T / 5)) ** 2) for k in range(B)]).astype(np.float32)
    src = [[abs_N + 4, Ny // 2, 0]]
    taps = [[Nx - abs_N - 5, Ny // 2 - 6, 1], [Nx - abs_N - 5, Ny // 2, 1],
            [Nx - abs_N - 6, Ny // 2 + 5, 0]]
    return dict(b=b, c=c, x=x, src=src, taps=taps, dt=0.707, h=1.0, Nx=Nx, Ny=Ny, B=B, T=T)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
