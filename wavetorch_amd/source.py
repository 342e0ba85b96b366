"""Sources (reference: wavetorch/source.py).

The modules keep the reference's buffers and ``forward`` so they can be used
on their own; inside ``WaveRNN`` their pixels are handed to the fused time
loop, which injects the samples in-kernel (source.py:19-22 semantics: the
post-step field gets ``+ x_n`` at every source pixel, scale dt**2 = 1.0).
"""
from __future__ import annotations

import numpy as np
import torch

from .utils import to_tensor


def bresenham_line(r0: int, c0: int, r1: int, c1: int):
    """Pixel rows/cols of the segment (r0,c0)-(r1,c1), one pixel per step along
    the major axis, endpoints included, in the order of the published
    algorithm of ``skimage.draw.line`` (used at source.py:31).  Axis-aligned
    and 45-degree segments have a unique pixel set.

    This restates the Bresenham loop of scikit-image's ``skimage/draw/_draw.pyx``
    (``_line``) step for step, so that the pixel order matches the one the
    reference gets.  scikit-image is distributed under the BSD 3-Clause
    license: Copyright (C) 2019, the scikit-image team.  All rights reserved.
    Redistribution and use in source and binary forms, with or without
    modification, are permitted provided that the copyright notice, this list
    of conditions and the following disclaimer are retained; neither the name
    of skimage nor the names of its contributors may be used to endorse or
    promote products derived from this software without specific prior
    written permission.  THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS
    AND CONTRIBUTORS "AS IS" AND ANY EXPRESS OR IMPLIED WARRANTIES ARE
    DISCLAIMED."""
    r, c = int(r0), int(c0)
    r1, c1 = int(r1), int(c1)
    steep = False
    dr = abs(r1 - r)
    dc = abs(c1 - c)
    sc = 1 if (c1 - c) > 0 else -1
    sr = 1 if (r1 - r) > 0 else -1
    if dr > dc:
        steep = True
        c, r = r, c
        dc, dr = dr, dc
        sc, sr = sr, sc
    d = (2 * dr) - dc
    rr = np.zeros(dc + 1, dtype=np.intp)
    cc = np.zeros(dc + 1, dtype=np.intp)
    for i in range(dc):
        if steep:
            rr[i] = c
            cc[i] = r
        else:
            rr[i] = r
            cc[i] = c
        while d >= 0:
            r = r + sr
            d = d - (2 * dc)
        c = c + sc
        d = d + (2 * dr)
    rr[dc] = r1
    cc[dc] = c1
    return rr, cc


class WaveSource(torch.nn.Module):
    """Additive point (or pixel-list) source (source.py:7-26)."""

    def __init__(self, x, y):
        super().__init__()
        # long buffers for advanced indexing (source.py:11-13)
        self.register_buffer('x', to_tensor(x, dtype=torch.int64))
        self.register_buffer('y', to_tensor(y, dtype=torch.int64))

    def forward(self, Y, X, dt=1.0):
        """Y + dt**2 * X at the source pixels (source.py:15-22)."""
        X_expanded = torch.zeros(Y.size(), dtype=Y.dtype, device=Y.device).detach()
        X_expanded[:, self.x, self.y] = X
        return Y + dt ** 2 * X_expanded

    def pixels(self):
        """Deduplicated (x, y) pixel list of this source (index_put assigns once)."""
        xs = np.atleast_1d(self.x.detach().cpu().numpy()).astype(np.int64)
        ys = np.atleast_1d(self.y.detach().cpu().numpy()).astype(np.int64)
        xs, ys = np.broadcast_arrays(xs, ys)
        seen = []
        for a, b in zip(xs.tolist(), ys.tolist()):
            if (a, b) not in seen:
                seen.append((a, b))
        return seen

    def plot(self, ax, color='r'):
        marker, = ax.plot(self.x.numpy(), self.y.numpy(), 'o', markeredgecolor=color,
                          markerfacecolor='none', markeredgewidth=1.0, markersize=4)
        return marker


class WaveLineSource(WaveSource):
    """Line source from (r0, c0) to (r1, c1) (source.py:29-41).

    Every pixel of the line receives the sample x[b, n].  (With batch > 1 the
    reference's index_put broadcasts (B,) against (B, L); this implementation
    gives each pixel of sample b the value x[b, n], which equals the
    reference whenever the reference is well defined, i.e. B == 1.)"""

    def __init__(self, r0, c0, r1, c1):
        x, y = bresenham_line(r0, c0, r1, c1)
        self.r0 = r0
        self.c0 = c0
        self.r1 = r1
        self.c1 = c1
        super().__init__(x, y)

    def plot(self, ax, color='r'):
        line, = ax.plot([self.r0, self.r1], [self.c0, self.c1], '-', color=color, lw=2)
        return line
