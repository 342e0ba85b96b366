"""WaveRNN (reference: wavetorch/rnn.py).

Same constructor and ``forward(x, output_fields=False)`` as the reference
(rnn.py:4-72).  Instead of a Python loop that calls the cell, every source
and every probe once per step, the whole time loop runs in gfx950 kernels
through one autograd Function (``_loop.WaveLoop``): leapfrog + source
injection + probe readout in temporally blocked LDS kernels forward, and a
checkpointed hand-written adjoint backward.
"""
from __future__ import annotations

import numpy as np
import torch

from ._loop import LoopConfig, Plan, WaveLoop
from .probe import WaveProbe


class WaveRNN(torch.nn.Module):
    def __init__(self, cell, sources, probes=[]):
        super().__init__()
        self.cell = cell
        if type(sources) is list:
            self.sources = torch.nn.ModuleList(sources)
        else:
            self.sources = torch.nn.ModuleList([sources])
        if type(probes) is list:
            self.probes = torch.nn.ModuleList(probes)
        else:
            self.probes = torch.nn.ModuleList([probes])
        # kernel tuning (wt_tuning fields); None = library heuristics.  Set by
        # wavetorch_amd.autotune.tune(model, ...) or by hand.
        self.tuning = None
        # checkpoint interval K of the adjoint (steps); None = the tuning's
        # ck_every, else 64.  A value set here overrides the tuning's.
        self._checkpoint_every = None
        # extension: keep only every field_stride-th field when output_fields
        self.field_stride = 1
        # extension: on-device non-finite checks of the loop's output and of
        # the gradients it returns (see nonfinite)
        self.check_finite = False
        self._nonfinite = None
        self._plans = {}
        self._cfg_cache = None

    @property
    def checkpoint_every(self) -> int:
        """Checkpoint interval K of the adjoint: the value set on the model,
        else the tuning's ``ck_every``, else 64."""
        if self._checkpoint_every is not None:
            return self._checkpoint_every
        return int((self.tuning or {}).get("ck_every") or 64)

    @checkpoint_every.setter
    def checkpoint_every(self, k):
        self._checkpoint_every = None if k is None else int(k)

    @property
    def nonfinite(self):
        """With ``check_finite`` set: a device int32 tensor of two flags,
        [output, gradient], set to 1 by any forward whose output (probe
        series or fields) or any backward whose returned gradients hold a
        NaN or an infinity since the last ``zero_()`` of this tensor.  The
        checks run on the device and never synchronise with the host; read
        the flags when convenient.  None until a checked forward has run."""
        return self._nonfinite

    def _nonfinite_flags(self, device):
        if not self.check_finite:
            return None
        f = self._nonfinite
        if f is None or f.device != device:
            f = self._nonfinite = torch.zeros(2, dtype=torch.int32, device=device)
        return f

    # ------------------------------------------------------------ config
    def _buffers_key(self):
        ts = [self.cell.dt, self.cell.geom.h, self.cell.satdamp_b0, self.cell.satdamp_uth,
              self.cell.c_nl]
        for s in self.sources:
            ts += [s.x, s.y]
        for p in self.probes:
            ts += [p.x, p.y]
        return tuple((t.data_ptr(), t._version) for t in ts) + (self.cell.c_clamp,)

    def loop_config(self) -> LoopConfig:
        """Host snapshot of dt, h, source pixels, probe taps and nonlinear
        parameters (cached until one of those buffers changes)."""
        key = self._buffers_key()
        if self._cfg_cache is not None and self._cfg_cache[0] == key:
            return self._cfg_cache[1]
        cell = self.cell
        Nx, Ny = cell.geom.domain_shape
        src = []
        for k, s in enumerate(self.sources):
            for (x, y) in s.pixels():
                src.append((int(x), int(y), k))
        taps = []
        if len(self.probes) > 0:
            shapes = [p.out_shape() for p in self.probes]
            if any(sh != shapes[0] for sh in shapes):
                raise RuntimeError("stack expects each probe output to be equal size "
                                   "(rnn.py:64)")
            per = [p.taps() for p in self.probes]
            L = len(per[0])
            # torch.stack(dim=-1) -> (..., L, P): column index = l * P + p
            for l in range(L):
                for p, prb in enumerate(self.probes):
                    x, y = per[p][l]
                    taps.append((int(x), int(y), 1 if prb.intensity else 0))
        sat_b0 = float(cell.satdamp_b0.item())
        cfg = LoopConfig(
            Nx=int(Nx), Ny=int(Ny), dt=float(cell.dt.item()), h=float(cell.geom.h.item()),
            src_px=tuple(src), taps=tuple(taps),
            sat_b0=sat_b0 if sat_b0 > 0 else 0.0,
            sat_uth=float(cell.satdamp_uth.item()),
            c_nl=float(cell.c_nl.item()),
            c_clamp=cell.c_clamp)
        self._cfg_cache = (key, cfg)
        return cfg

    def plan(self, B: int, T: int, device) -> Plan:
        cfg = self.loop_config()
        tun = dict(self.tuning) if self.tuning else {}
        tun["ck_every"] = self.checkpoint_every
        tkey = tuple(sorted(tun.items()))
        key = (cfg, B, T, str(device), tkey)
        p = self._plans.get(key)
        if p is None:
            if len(self._plans) > 8:
                self._plans.clear()
            p = Plan(cfg, B, T, self.cell.geom.cmax, device, tun)
            self._plans[key] = p
        return p

    # ------------------------------------------------------------ forward
    def forward(self, x, output_fields=False):
        """Propagate for the length of the inputs (rnn.py:21-72).

        x: (B, T) input sequences.  Returns (B, T, P) probe series (or
        (B, T, L, P) for multi-pixel probes), or the fields (B, T, Nx, Ny) when
        there are no probes or ``output_fields`` is set.
        """
        if not x.is_cuda:
            raise RuntimeError("wavetorch_amd.WaveRNN runs on the HIP device only; move the "
                               "model and inputs with .to('cuda')")
        if x.dtype not in (torch.float32, torch.float64):
            raise NotImplementedError(f"wavetorch_amd computes in float32 or float64 "
                                      f"(got {x.dtype})")
        if x.dtype != self.cell.geom.b.dtype:
            raise TypeError(f"wavetorch_amd: input is {x.dtype}, the model's buffers are "
                            f"{self.cell.geom.b.dtype} (set_dtype before building the model)")
        B, T = x.shape[0], x.shape[1]
        fields_mode = len(self.probes) == 0 or output_fields
        stride = int(self.field_stride) if fields_mode else 1
        if fields_mode and T % stride != 0:
            raise ValueError(f"field_stride {stride} must divide the number of steps {T}")
        geom = self.cell.geom
        c = geom.c
        b = geom.b
        rho = geom.rho if self.loop_config().nonlinear else None
        plan = self.plan(B, T, x.device)
        out = WaveLoop.apply(x, c, b, rho, plan, fields_mode, stride,
                             self._nonfinite_flags(x.device))
        if fields_mode:
            return out
        shp = self.probes[0].out_shape()  # from tensor shapes: no device read
        if len(shp) == 0:
            return out
        return out.reshape(B, T, *shp, len(self.probes))
