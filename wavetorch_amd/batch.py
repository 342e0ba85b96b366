"""Batched instances: many independent WaveRNN problems of uneven grid size
and length, run together (``wt_batch_*`` in include/wavetorch_hip.h).

The reference runs one configuration at a time (``WaveRNN.forward``,
rnn.py:21-72, then autograd through cell.py:26-44); a set of independent
problems -- e.g. one inverse-design problem per lens size -- is a Python loop
over models.  ``InstanceBatch`` groups the instances by configuration (grid,
PML field, sources, probes, dt) and runs all of them as one sequence of
launches in which every launch covers every instance that still has steps to
run; each instance keeps its own wave speed c and its own number of steps.
The whole forward + adjoint for a loss with a fixed dJ/dy is one library call
(``fwd_adj``) and can be captured once and replayed as one HIP graph
(``capture`` / ``replay``).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._loop import LoopConfig


class Instance:
    """One problem: its loop configuration, PML field b (Nx,Ny), wave speed c
    (Nx,Ny) and number of steps."""

    def __init__(self, cfg: LoopConfig, b: torch.Tensor, length: int, cmax: float):
        self.cfg, self.b, self.length, self.cmax = cfg, b, int(length), float(cmax)

    @staticmethod
    def from_model(model, length: int):
        """The configuration of a WaveRNN (linear cell) for `length` steps."""
        return Instance(model.loop_config(), model.cell.geom.b, length, model.cell.geom.cmax)


class InstanceBatch:
    """Groups `instances` by configuration and owns the batch plan, the
    workspaces and padded per-group buffers on `device`.

    Per group g (instances longest first, or in the given order with
    `longest_first` False): c (B,Nx,Ny), x (B,T) with T the
    longest length of the batch (zero-padded), ybar (B,T,P), outputs y
    (B,T,P), grad_c (B,Nx,Ny), grad_x (B,T).  With `grad` False the batch
    runs forward only and holds no saved workspace (checkpoints and the
    forward's record, the bulk of the memory of a one-segment plan)."""

    def __init__(self, instances, device, tuning=None, checkpoint_every=64,
                 longest_first=True, grad=True):
        self.device = torch.device(device)
        self.instances = list(instances)
        groups = {}
        for i, inst in enumerate(self.instances):
            if inst.cfg.nonlinear:
                raise NotImplementedError("batched instances take the linear cell")
            key = (inst.cfg, inst.b.detach().cpu().numpy().tobytes())
            groups.setdefault(key, []).append(i)
        if len(groups) > 8:
            raise ValueError(f"{len(groups)} distinct configurations; at most 8 per batch")
        # longest first within a group: launches then cover only the prefix
        # of instances that still have steps to run (in any other order they
        # cover every instance of the group while its longest one runs)
        self.groups = [sorted(idx, key=lambda i: -self.instances[i].length) if longest_first
                       else idx for idx in groups.values()]
        self.where = {}
        for g, idx in enumerate(self.groups):
            for k, i in enumerate(idx):
                self.where[i] = (g, k)
        self.T = max(inst.length for inst in self.instances)
        G = len(self.groups)
        cfgs = (_lib.WtConfig * G)()
        keep = []
        for g, idx in enumerate(self.groups):
            c, k = self.instances[idx[0]].cfg.to_c()
            cfgs[g] = c
            keep.append(k)
        B = (ctypes.c_int32 * G)(*[len(idx) for idx in self.groups])
        lens = [self.instances[i].length for idx in self.groups for i in idx]
        lengths = (ctypes.c_int32 * len(lens))(*lens)
        cmax = (ctypes.c_float * G)(*[self.instances[idx[0]].cmax for idx in self.groups])
        tun = None
        if tuning is not None:
            tun = (_lib.WtTuning * G)()
            per = tuning if isinstance(tuning, (list, tuple)) else [tuning] * G
            for g in range(G):
                d = dict(per[g])
                d.setdefault("ck_every", checkpoint_every)
                tun[g] = _lib.WtTuning(**{k: int(v) for k, v in d.items()})
        elif checkpoint_every != 64:
            tun = (_lib.WtTuning * G)()
            for g in range(G):
                tun[g].ck_every = int(checkpoint_every)
        h = ctypes.c_void_p()
        L = _lib.lib()
        with torch.cuda.device(self.device):
            _lib.check(L.wt_batch_plan_create(G, cfgs, B, lengths, cmax, tun, ctypes.byref(h)),
                       "wt_batch_plan_create")
        self.handle = h
        del keep
        self.tuning = []
        for g in range(G):
            t = _lib.WtTuning()
            _lib.check(L.wt_batch_plan_tuning(h, g, ctypes.byref(t)), "wt_batch_plan_tuning")
            self.tuning.append(t.as_dict())
        s0, t0 = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(L.wt_batch_workspace_bytes(h, ctypes.byref(s0), ctypes.byref(t0)),
                   "wt_batch_workspace_bytes")
        dev = self.device
        # the saved workspace (checkpoints, the forward's record) only when
        # gradients are wanted: forward-only batches run without it
        self.grad = bool(grad)
        self.saved = (torch.empty(max(s0.value, 1), dtype=torch.uint8, device=dev)
                      if self.grad else None)
        self.temp = torch.empty(max(t0.value, 1), dtype=torch.uint8, device=dev)
        # per-group buffers
        self.buf = []
        for idx in self.groups:
            cfg = self.instances[idx[0]].cfg
            Bg, P = len(idx), len(cfg.taps)
            self.buf.append(dict(
                b=self.instances[idx[0]].b.detach().to(dev, torch.float32).contiguous(),
                c=torch.zeros((Bg, cfg.Nx, cfg.Ny), device=dev),
                x=torch.zeros((Bg, self.T), device=dev),
                y=torch.zeros((Bg, self.T, P), device=dev),
                ybar=torch.zeros((Bg, self.T, P), device=dev),
                grad_c=torch.zeros((Bg, cfg.Nx, cfg.Ny), device=dev),
                grad_x=torch.zeros((Bg, self.T), device=dev)))
        self.io = (_lib.WtBatchIO * G)()
        for g, bf in enumerate(self.buf):
            self.io[g] = _lib.WtBatchIO(*[bf[k].data_ptr() for k in
                                          ("b", "c", "x", "y", "ybar", "grad_c", "grad_x")])
        self._graph = None
        # the buffers were zeroed on the current stream; callers may run on
        # another one
        torch.cuda.synchronize(self.device)

    def __del__(self):
        try:
            if getattr(self, "_graph", None) is not None:
                _lib.lib().wt_graph_destroy(self._graph)
                self._graph = None
            if getattr(self, "handle", None) is not None and self.handle.value:
                _lib.lib().wt_batch_plan_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    # ------------------------------------------------------------ inputs
    def set_inputs(self, cs, xs, ybars=None):
        """Copies per-instance c (Nx,Ny), x (length,) and optionally dJ/dy
        (length, P) into the padded group buffers (entries past an instance's
        length are zero)."""
        with torch.no_grad():
            for i, inst in enumerate(self.instances):
                g, k = self.where[i]
                bf = self.buf[g]
                bf["c"][k].copy_(cs[i])
                bf["x"][k].zero_()
                bf["x"][k, :inst.length].copy_(xs[i][:inst.length])
                if ybars is not None:
                    bf["ybar"][k].zero_()
                    bf["ybar"][k, :inst.length].copy_(ybars[i][:inst.length])

    def outputs(self):
        """Per-instance (y (length, P), grad_c (Nx,Ny), grad_x (length,)) views."""
        out = []
        for i, inst in enumerate(self.instances):
            g, k = self.where[i]
            bf = self.buf[g]
            out.append((bf["y"][k, :inst.length], bf["grad_c"][k],
                        bf["grad_x"][k, :inst.length]))
        return out

    # ------------------------------------------------------------ runs
    def _stream(self):
        return _lib.stream_ptr(self.device)

    def _need_grad(self):
        if not self.grad:
            raise RuntimeError("InstanceBatch(grad=False) holds no saved workspace: forward "
                               "only (forward(with_grad=False))")

    def forward(self, with_grad=True):
        if with_grad:
            self._need_grad()
        L = _lib.lib()
        saved = self.saved.data_ptr() if self.saved is not None else None
        _lib.check(L.wt_batch_fwd(self.handle, self.io, int(with_grad), saved,
                                  self.temp.data_ptr(), self._stream()), "wt_batch_fwd")

    def backward(self):
        self._need_grad()
        L = _lib.lib()
        _lib.check(L.wt_batch_adj(self.handle, self.io, self.saved.data_ptr(),
                                  self.temp.data_ptr(), self._stream()), "wt_batch_adj")

    def fwd_adj(self):
        self._need_grad()
        L = _lib.lib()
        _lib.check(L.wt_batch_fwd_adj(self.handle, self.io, self.saved.data_ptr(),
                                      self.temp.data_ptr(), self._stream()), "wt_batch_fwd_adj")

    def capture(self):
        """Captures fwd_adj (on the current stream, which must not be the
        default stream) into a HIP graph; replay() launches it.  The graph
        reads and writes this object's buffers: update them with set_inputs
        between replays."""
        L = _lib.lib()
        s = self._stream()
        self.fwd_adj()  # first launches of each kernel variant set their LDS limits
        torch.cuda.current_stream(self.device).synchronize()
        _lib.check(L.wt_graph_begin(s), "wt_graph_begin")
        try:
            self.fwd_adj()
        finally:
            g = ctypes.c_void_p()
            rc = L.wt_graph_end(s, ctypes.byref(g))
        _lib.check(rc, "wt_graph_end")
        if self._graph is not None:
            L.wt_graph_destroy(self._graph)
        self._graph = g

    def replay(self):
        _lib.check(_lib.lib().wt_graph_launch(self._graph, self._stream()), "wt_graph_launch")

    def cell_steps(self):
        return sum(inst.cfg.Nx * inst.cfg.Ny * inst.length for inst in self.instances)
