"""Probes (reference: wavetorch/probe.py).

Inside ``WaveRNN`` the probe pixels are read by the fused time loop right
after source injection (rnn.py:59-64); the modules' own ``forward`` is kept
for direct use.
"""
from __future__ import annotations

import numpy as np
import torch

from .utils import to_tensor


class WaveProbe(torch.nn.Module):
    """Field amplitude at (x, y) (probe.py:6-19)."""

    intensity = False

    def __init__(self, x, y):
        super().__init__()
        self.register_buffer('x', to_tensor(x, dtype=torch.int64))
        self.register_buffer('y', to_tensor(y, dtype=torch.int64))

    def forward(self, x):
        return x[:, self.x, self.y]

    def taps(self):
        """(x, y) pixels this probe reads, in output order (flattened)."""
        xs = np.atleast_1d(self.x.detach().cpu().numpy()).astype(np.int64)
        ys = np.atleast_1d(self.y.detach().cpu().numpy()).astype(np.int64)
        xs, ys = np.broadcast_arrays(xs, ys)
        return list(zip(xs.ravel().tolist(), ys.ravel().tolist()))

    def out_shape(self):
        """Trailing shape of this probe's output per (batch, step)."""
        return tuple(np.broadcast_shapes(tuple(self.x.shape), tuple(self.y.shape)))

    def plot(self, ax, color='k'):
        marker, = ax.plot(self.x.numpy(), self.y.numpy(), 'o', markeredgecolor=color,
                          markerfacecolor='none', markeredgewidth=1.0, markersize=4)
        return marker


class WaveIntensityProbe(WaveProbe):
    """Field intensity u**2 at (x, y) (probe.py:22-27)."""

    intensity = True

    def __init__(self, x, y):
        super().__init__(x, y)

    def forward(self, x):
        return super().forward(x).pow(2)
