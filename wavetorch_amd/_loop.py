"""The fused time loop behind ``WaveRNN.forward`` (reference: rnn.py:21-72).

``WaveLoop`` is one ``torch.autograd.Function`` for the whole T-step loop:
its forward runs ``wt_loop_fwd`` (leapfrog + sources + probes, checkpoints
every K steps) and its backward runs ``wt_loop_adj`` (segment recompute from
checkpoints + hand-written adjoint), replacing the per-step autograd tape the
reference builds (cell.py:20-44 replayed T times).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib


@dataclass(frozen=True)
class LoopConfig:
    """Host-side description of one WaveRNN configuration (``wt_config``)."""
    Nx: int
    Ny: int
    dt: float
    h: float
    src_px: tuple          # ((x, y, source_index), ...)
    taps: tuple            # ((x, y, kind), ...), output column = position
    sat_b0: float = 0.0
    sat_uth: float = 0.0
    c_nl: float = 0.0
    c_clamp: tuple | None = None

    @property
    def nonlinear(self) -> bool:
        return self.sat_b0 > 0 or self.c_nl != 0

    def to_c(self):
        src = np.ascontiguousarray(np.asarray(self.src_px, dtype=np.int32).reshape(-1, 3))
        taps = np.ascontiguousarray(np.asarray(self.taps, dtype=np.int32).reshape(-1, 3))
        cfg = _lib.WtConfig()
        cfg.Nx, cfg.Ny = self.Nx, self.Ny
        cfg.dt, cfg.h = self.dt, self.h
        cfg.nsrc_px = len(src)
        cfg.src_px = src.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        cfg.ntap = len(taps)
        cfg.taps = taps.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        cfg.sat_b0, cfg.sat_uth, cfg.c_nl = self.sat_b0, self.sat_uth, self.c_nl
        if self.c_clamp is not None:
            cfg.c_clamp, cfg.c_min, cfg.c_max = 1, self.c_clamp[0], self.c_clamp[1]
        return cfg, (src, taps)  # keep the arrays alive while cfg is used


def scalars_f64(cfg: LoopConfig):
    """``wt_scalars_f64`` of a configuration: dt, h and the nonlinear
    coefficients in double precision (``wt_config`` holds them as float)."""
    cmin, cmax = cfg.c_clamp if cfg.c_clamp is not None else (0.0, 0.0)
    return _lib.WtScalarsF64(dt=cfg.dt, h=cfg.h, sat_b0=cfg.sat_b0, sat_uth=cfg.sat_uth,
                             c_nl=cfg.c_nl, c_min=cmin, c_max=cmax)


def resolve(cfg: LoopConfig, B: int, T: int, cmax: float, tuning=None):
    """Host-only plan resolution (``wt_plan_resolve``): validates ``cfg``
    (ValueError on CFL or bad indices, as the reference's constructors raise)
    and returns (resolved tuning dict, saved bytes, temp bytes) without
    touching a device."""
    c, keep = cfg.to_c()
    tun = _lib.WtTuning(**{k: int(v) for k, v in tuning.items()}) if tuning else None
    out = _lib.WtTuning()
    s0, t0 = _lib.c_i64(), _lib.c_i64()
    _lib.check(_lib.lib().wt_plan_resolve(ctypes.byref(c), B, T, float(cmax),
                                          ctypes.byref(tun) if tun is not None else None,
                                          ctypes.byref(out), ctypes.byref(s0), ctypes.byref(t0)),
               "wt_plan_resolve")
    del keep
    return out.as_dict(), s0.value, t0.value


class Plan:
    """Owner of a library plan (validated config + device tables + tiling)."""

    def __init__(self, cfg: LoopConfig, B: int, T: int, cmax: float, device, tuning=None):
        self.cfg, self.B, self.T, self.device = cfg, B, T, device
        c, keep = cfg.to_c()
        tun = None
        if tuning:
            tun = _lib.WtTuning(**{k: int(v) for k, v in tuning.items()})
        handle = ctypes.c_void_p()
        L = _lib.lib()
        with torch.cuda.device(device):
            _lib.check(L.wt_plan_create(ctypes.byref(c), B, T, float(cmax),
                                        ctypes.byref(tun) if tun is not None else None,
                                        ctypes.byref(handle)), "wt_plan_create")
        self.handle = handle
        del keep
        t = _lib.WtTuning()
        _lib.check(L.wt_plan_tuning(handle, ctypes.byref(t)), "wt_plan_tuning")
        self.tuning = t.as_dict()
        s0, t0 = _lib.c_i64(), _lib.c_i64()
        _lib.check(L.wt_workspace_bytes(handle, 1, ctypes.byref(s0), ctypes.byref(t0)),
                   "wt_workspace_bytes")
        self.saved_bytes, self.temp_bytes = s0.value, t0.value
        _lib.check(L.wt_workspace_bytes(handle, 0, ctypes.byref(s0), ctypes.byref(t0)),
                   "wt_workspace_bytes")
        self.temp_bytes_nograd = t0.value

    def bytes_f64(self, with_grad: bool):
        """(saved, temp) bytes of the float64 loop (wt_workspace_bytes_f64)."""
        s0, t0 = _lib.c_i64(), _lib.c_i64()
        _lib.check(_lib.lib().wt_workspace_bytes_f64(self.handle, int(with_grad),
                                                     ctypes.byref(s0), ctypes.byref(t0)),
                   "wt_workspace_bytes_f64")
        return s0.value, t0.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().wt_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


def _buf(nbytes, device):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def _check_finite(t, flag, slot, stream):
    if t is not None and flag is not None:
        name = "wt_check_finite_f64" if t.dtype == torch.float64 else "wt_check_finite"
        _lib.check(getattr(_lib.lib(), name)(t.data_ptr(), t.numel(),
                                             flag.data_ptr() + 4 * slot, stream), name)


class WaveLoop(torch.autograd.Function):
    """y = WaveRNN time loop (x, c, b, rho); backward = checkpointed adjoint.
    float64 inputs take the double-precision loop (wt_loop_fwd_f64 /
    wt_loop_adj_f64: every field kept, no recompute).

    ``nonfinite``: None, or a device int32 tensor of two entries: entry 0 is
    set to 1 when the forward's output holds a NaN or infinity, entry 1 when
    a gradient the backward returns does (checked on the device with
    ``wt_check_finite``; nothing waits for the host)."""

    @staticmethod
    def forward(ctx, x, c, b, rho, plan: Plan, output_fields: bool, field_stride: int,
                nonfinite=None):
        dev = x.device
        x = x.contiguous()
        c = c.contiguous()
        b = b.contiguous()
        rho_c = rho.contiguous() if rho is not None else None
        B, T = x.shape
        Nx, Ny = plan.cfg.Nx, plan.cfg.Ny
        need = ctx.needs_input_grad
        with_grad = bool(need[0] or need[1] or need[2] or (rho is not None and need[3]))
        ntap = len(plan.cfg.taps)
        f64 = c.dtype == torch.float64
        dtype = torch.float64 if f64 else torch.float32
        for t_, name in ((x, "x"), (c, "c"), (b, "b"), (rho_c, "rho")):
            if t_ is not None and t_.dtype != dtype:
                raise TypeError(f"wavetorch_amd: {name} is {t_.dtype}, the loop runs in {dtype}")
        y = None
        fields = None
        if output_fields:
            fields = torch.empty((B, T // field_stride, Nx, Ny), dtype=dtype, device=dev)
        else:
            y = torch.empty((B, T, ntap), dtype=dtype, device=dev)
        s = _lib.stream_ptr(dev)
        if f64:
            sb, tb = plan.bytes_f64(with_grad)
            saved = _buf(sb, dev) if with_grad else None
            # three rotating fields without gradients; the tape holds every
            # field otherwise
            temp = None if with_grad else _buf(plan.bytes_f64(False)[1], dev)
            sc = scalars_f64(plan.cfg)
            _lib.check(_lib.lib().wt_loop_fwd_f64(
                plan.handle, ctypes.byref(sc), b.data_ptr(), c.data_ptr(), _lib.ptr(rho_c),
                x.data_ptr(), _lib.ptr(y), _lib.ptr(fields), int(field_stride), int(with_grad),
                _lib.ptr(saved), _lib.ptr(temp), s), "wt_loop_fwd_f64")
        else:
            saved = _buf(plan.saved_bytes, dev) if with_grad else None
            temp = _buf(plan.temp_bytes_nograd, dev)  # the forward's scratch, with or without grad
            _lib.check(_lib.lib().wt_loop_fwd(
                plan.handle, b.data_ptr(), c.data_ptr(), _lib.ptr(rho_c), x.data_ptr(),
                _lib.ptr(y), _lib.ptr(fields), int(field_stride), int(with_grad),
                _lib.ptr(saved), temp.data_ptr(), s), "wt_loop_fwd")
        del temp
        _check_finite(fields if output_fields else y, nonfinite, 0, s)
        ctx.nonfinite = nonfinite
        if with_grad:
            ctx.save_for_backward(x, c, b, rho_c if rho_c is not None else x.new_empty(0))
            ctx.ws = saved
            ctx.plan = plan
            ctx.has_rho = rho_c is not None
            ctx.output_fields = output_fields
            ctx.field_stride = field_stride
        return fields if output_fields else y

    @staticmethod
    def backward(ctx, gout):
        x, c, b, rho = ctx.saved_tensors
        plan = ctx.plan
        if ctx.output_fields and ctx.field_stride != 1:
            raise NotImplementedError("wavetorch_amd: gradients through strided field snapshots")
        rho = rho if ctx.has_rho else None
        dev = x.device
        need = ctx.needs_input_grad
        Nx, Ny = plan.cfg.Nx, plan.cfg.Ny
        f64 = c.dtype == torch.float64
        dtype = c.dtype
        g = gout.to(dtype).contiguous()
        gc = torch.empty((Nx, Ny), dtype=dtype, device=dev) if need[1] else None
        gb = torch.empty((Nx, Ny), dtype=dtype, device=dev) if need[2] else None
        grho = torch.empty((Nx, Ny), dtype=dtype, device=dev) \
            if (rho is not None and need[3]) else None
        gx = torch.empty_like(x) if need[0] else None
        s = _lib.stream_ptr(dev)
        ybar = None if ctx.output_fields else g.data_ptr()
        fbar = g.data_ptr() if ctx.output_fields else None
        if f64:
            temp = _buf(plan.bytes_f64(True)[1], dev)
            sc = scalars_f64(plan.cfg)
            _lib.check(_lib.lib().wt_loop_adj_f64(
                plan.handle, ctypes.byref(sc), b.data_ptr(), c.data_ptr(), _lib.ptr(rho), ybar,
                fbar, ctx.ws.data_ptr(), temp.data_ptr(), _lib.ptr(gc), _lib.ptr(gb),
                _lib.ptr(grho), _lib.ptr(gx), s), "wt_loop_adj_f64")
        else:
            temp = _buf(plan.temp_bytes, dev)
            _lib.check(_lib.lib().wt_loop_adj(
                plan.handle, b.data_ptr(), c.data_ptr(), _lib.ptr(rho), x.data_ptr(), ybar, fbar,
                ctx.ws.data_ptr(), temp.data_ptr(), _lib.ptr(gc), _lib.ptr(gb), _lib.ptr(grho),
                _lib.ptr(gx), s), "wt_loop_adj")
        ctx.ws = None
        for gt in (gx, gc, gb, grho):
            _check_finite(gt, ctx.nonfinite, 1, s)
        return gx, gc, gb, grho, None, None, None, None
