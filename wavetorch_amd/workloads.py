"""Seeded synthetic workloads for benchmarks, tests and the smoke check.

Each builder returns a ``WaveRNN`` (random-initialised geometry), its input
batch and labels, built the way the reference's studies build theirs
(study/vowel_train.py:101-119, study/vowel_helpers.py:4-35,
study/propagate.py:16-51, study/optimize_lens.py:14-70), with synthetic
signals instead of recordings (no datasets are available offline):

* ``train_150_b64``      150x150 grid with a 20-cell PML, free-form rho on the
                         110x110 interior, batch 64, 3000 steps of vowel-like
                         10 kHz waveforms, 3 intensity probes, cross-entropy
* ``propagate_100_b4``   100x100 grid, batch 4, 1000 steps, windowed 500 Hz sine
* ``nonlinear_150_b32``  the 150x150 problem at batch 32 with saturable damping
                         and a Kerr-like speed
* ``scatter_1024``       1024x1024 grid, 20 circular scatterers, one
                         Gaussian-pulse source, forward only
* ``lens_batch_256``     256 independent lens-style problems, grid side drawn
                         from {64, 96, 128, 192, 256} (16-cell PML), length
                         drawn from [1000, 4000], each with its own rho, a
                         line source and one focal intensity probe
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .cell import WaveCell
from .geom import WaveGeometryFreeForm
from .probe import WaveIntensityProbe
from .rnn import WaveRNN
from .source import WaveLineSource, WaveSource


def vowel_like(B: int, T: int, sr: float = 10000.0, seed: int = 2, snr_db: float = 30.0):
    """Synthetic vowels: f0 ~ U[100,250] Hz, 15 harmonics weighted by three
    Gaussian formants F1 ~ U[300,900], F2 ~ U[900,2500], F3 ~ U[2300,3200] Hz
    (bandwidths 90/110/150 Hz), Hann window, peak-normalised, white noise at
    `snr_db`."""
    rng = np.random.default_rng(seed)
    t = np.arange(T) / sr
    out = np.zeros((B, T))
    bw = (90.0, 110.0, 150.0)
    for b in range(B):
        f0 = rng.uniform(100, 250)
        F = (rng.uniform(300, 900), rng.uniform(900, 2500), rng.uniform(2300, 3200))
        s = np.zeros(T)
        for k in range(1, 16):
            fk = k * f0
            a = sum(math.exp(-0.5 * ((fk - Fi) / wi) ** 2) for Fi, wi in zip(F, bw)) + 1e-3
            s += a * np.sin(2 * np.pi * fk * t + rng.uniform(0, 2 * np.pi))
        s *= np.hanning(T)
        s /= np.abs(s).max()
        p = np.mean(s ** 2)
        s += rng.standard_normal(T) * math.sqrt(p / 10 ** (snr_db / 10))
        out[b] = s
    return torch.from_numpy(out.astype(np.float32))


def _freeform(N, pml, seed, sig=1.0, p=2.0, dt=0.707):
    g = torch.Generator().manual_seed(seed)
    rho = torch.rand(N, N, generator=g)
    geom = WaveGeometryFreeForm((N, N), 1.0, c0=1.0, c1=0.5, abs_N=pml, abs_sig=sig, abs_p=p,
                                eta=0.5, beta=100.0, rho=rho, blur_radius=1, blur_N=1)
    return geom


def train_150_b64(B=64, T=3000, N=150, pml=20, seed=2000):
    geom = _freeform(N, pml, seed)
    cell = WaveCell(0.707, geom)
    src = WaveSource(40, 75)
    probes = [WaveIntensityProbe(110, 55), WaveIntensityProbe(110, 75),
              WaveIntensityProbe(110, 95)]
    model = WaveRNN(cell, src, probes)
    x = vowel_like(B, T, seed=seed + 2)
    labels = torch.from_numpy(np.random.default_rng(seed + 1).integers(0, 3, B))
    return model, x, labels


def propagate_100_b4(B=4, T=1000, N=100, pml=20, seed=0):
    geom = _freeform(N, pml, seed)
    cell = WaveCell(0.707, geom)
    src = WaveSource(pml + 20, N // 2)
    pd = 20
    y0 = int((N - 2 * pd) / 2)
    probes = [WaveIntensityProbe(N - pml - 20, y0 + i * pd) for i in range(3)]
    model = WaveRNN(cell, src, probes)
    n = np.arange(T)
    sig = T / 6.0
    x = np.sin(2 * np.pi * 0.05 * n) * np.exp(-((n - T / 2) / sig) ** 2)
    x = torch.from_numpy(np.tile(x, (B, 1)).astype(np.float32))
    labels = torch.from_numpy(np.random.default_rng(seed + 1).integers(0, 3, B))
    return model, x, labels


def nonlinear_150_b32(B=32, T=3000, N=150, pml=20, seed=3000, dt=0.64):
    """Saturable damping b_nl = 0.5, I_sat = 0.1 (uth = sqrt(0.1)), Kerr-like
    c_nl = 0.05 with c clamped to [0.4, 1.1]; dt = 0.64 keeps c = 1.1 inside
    the 2-D CFL bound c*dt/h <= 1/sqrt(2)."""
    geom = _freeform(N, pml, seed)
    cell = WaveCell(dt, geom, satdamp_b0=0.5, satdamp_uth=math.sqrt(0.1), c_nl=0.05,
                    c_clamp=(0.4, 1.1))
    src = WaveSource(40, 75)
    probes = [WaveIntensityProbe(110, 55), WaveIntensityProbe(110, 75),
              WaveIntensityProbe(110, 95)]
    model = WaveRNN(cell, src, probes)
    x = vowel_like(B, T, seed=seed + 2)
    labels = torch.from_numpy(np.random.default_rng(seed + 1).integers(0, 3, B))
    return model, x, labels


def scatter_1024(T=4000, N=1024, pml=32, nscat=20, seed=1000):
    rng = np.random.default_rng(seed + 3)
    rho = np.zeros((N, N), dtype=np.float32)
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    for _ in range(nscat):
        r = rng.integers(8, 31)
        cx, cy = rng.integers(pml + r, N - pml - r, size=2)
        rho[(ii - cx) ** 2 + (jj - cy) ** 2 < r * r] = 1.0
    geom = WaveGeometryFreeForm((N, N), 1.0, c0=1.0, c1=0.5, abs_N=pml, abs_sig=1.0, abs_p=2.0,
                                eta=0.5, beta=100.0, rho=torch.from_numpy(rho), blur_radius=1,
                                blur_N=0)
    cell = WaveCell(0.707, geom)
    src = WaveSource(N // 2, N // 2)
    model = WaveRNN(cell, src, [])
    n = np.arange(T)
    x = np.exp(-((n - 60) / 20.0) ** 2).astype(np.float32)[None]
    return model, torch.from_numpy(x), None


def lens_instance(N, T, seed, pml=16):
    """One lens-style inverse-design problem (after study/optimize_lens.py:
    14-70): free-form rho on an N x N grid with a `pml`-cell PML, a horizontal
    line source across the middle half near the top edge carrying
    sin(2 pi 0.1 n), and one focal intensity probe on the centre line three
    quarters down.  Returns (model, x (T,))."""
    g = torch.Generator().manual_seed(seed)
    rho = torch.rand(N, N, generator=g)
    geom = WaveGeometryFreeForm((N, N), 1.0, c0=1.0, c1=0.5, abs_N=pml, abs_sig=1.0, abs_p=2.0,
                                eta=0.5, beta=100.0, rho=rho, blur_radius=1, blur_N=1)
    cell = WaveCell(0.707, geom)
    r = pml + 4
    src = WaveLineSource(r, N // 4, r, 3 * N // 4 - 1)
    probe = WaveIntensityProbe(N - pml - N // 4, N // 2)
    model = WaveRNN(cell, src, [probe])
    x = np.sin(2 * np.pi * 0.1 * np.arange(T)).astype(np.float32)
    return model, torch.from_numpy(x)


def lens_batch_256(n=256, seed=4000, sides=(64, 96, 128, 192, 256), tmin=1000, tmax=4000):
    """`n` lens problems with seeded sides and lengths; returns a list of
    (model, x) -- the loss of each is J = -sum_n y[n] (maximise the focal
    intensity), so dJ/dy = -1 over its length."""
    rng = np.random.default_rng(seed + 3)
    N = rng.choice(np.asarray(sides), n)
    T = rng.integers(tmin, tmax + 1, n)
    return [lens_instance(int(N[i]), int(T[i]), seed + 10 + i) for i in range(n)]
