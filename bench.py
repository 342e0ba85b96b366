"""Benchmark: one training step of the 150x150, batch-64, 3000-step WaveRNN.

A step is the hot path over one synthetic batch: the fused forward time loop
(leapfrog + source + 3 intensity probes), the cross-entropy loss head, the
checkpointed adjoint (recompute every K=64 steps) back to rho through the
blur/projection, and an Adam update of rho.  Inputs are resident on the GPU
before the timed region.

Prints ONE JSON line:
  value          fwd+adjoint Gcell-steps/s = B*Nx*Ny*T / step time / 1e9
  roofline       the dominant kernel (by summed time) at 12 B (forward) or
                 16 B (adjoint) of algorithmic HBM traffic per cell-step,
                 timed with HIP events on its launch stream over the timed
                 region, against MI355X's 8 TB/s
  cpu_baseline   the torch-CPU restatement of the reference (oracle/ref_torch)
                 timed forward+backward on a bounded sample of the same
                 workload on this host's cores
  plain          the same step with the one-launch-per-step kernels
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12
BYTES = {0: 12.0, 1: 12.0, 2: 16.0}  # algorithmic bytes per cell-step per kernel class
CLASS_NAME = {0: "k_fwd_tiled (forward)", 1: "k_fwd_tiled (recompute, record lap)",
              2: "k_adj_tiled (adjoint)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build(dev, tuning=None, workload="train_150_b64", T=None):
    from wavetorch_amd import workloads
    kw = {} if T is None else {"T": T}
    model, x, labels = getattr(workloads, workload)(**kw)
    model = model.to(dev)
    if tuning is not None:
        model.tuning = tuning
    return model, x.to(dev), labels.to(dev)


def train_step(model, x, labels, opt):
    from wavetorch_amd.utils import normalize_power
    opt.zero_grad(set_to_none=True)
    y = model(x)
    loss = torch.nn.functional.cross_entropy(normalize_power(y.sum(dim=1)), labels)
    loss.backward()
    opt.step()
    return loss


def time_steps(model, x, labels, steps, warmup, profile=False):
    from wavetorch_amd import _lib
    opt = torch.optim.Adam(model.parameters(), lr=1e-3)
    for _ in range(warmup):
        train_step(model, x, labels, opt)
    torch.cuda.synchronize()
    L = _lib.lib()
    if profile:
        L.wt_profile(1)
        _lib.check(L.wt_profile_read(None, None, None), "wt_profile_read")
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = train_step(model, x, labels, opt)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    prof = None
    if profile:
        import ctypes
        ms = (ctypes.c_double * 3)()
        n = (ctypes.c_int64 * 3)()
        w = (ctypes.c_int64 * 3)()
        _lib.check(L.wt_profile_read(ms, n, w), "wt_profile_read")
        L.wt_profile(0)
        prof = {c: (ms[c], n[c], w[c]) for c in range(3)}
    return dt, float(loss.item()), prof


def cpu_baseline(budget_s=15.0):
    """The reference's op sequence (torch-CPU restatement, fp32) on the same
    workload with T cut to fit ~budget_s of forward+backward."""
    from oracle import ref_torch as rt
    from wavetorch_amd import workloads
    model, x, labels = workloads.train_150_b64()
    g = model.cell.geom
    ncores = os.cpu_count() or 1
    torch.set_num_threads(ncores)
    rho = g.rho.detach().clone().requires_grad_(True)
    k = g.blur_kernel
    dt = torch.tensor(0.707)
    h = torch.tensor(1.0)
    src = [(torch.tensor(40), torch.tensor(75))]
    probes = [(torch.tensor(110), torch.tensor(p), True) for p in (55, 75, 95)]

    def run(T):
        c = rt.freeform_c(rho, 1.0, 0.5, k, 1, 0.5, 100.0)
        y = rt.rnn_forward(x[:, :T].contiguous(), c=c, rho=rho, b_geom=g.b, dt=dt, h=h,
                           sources=src, probes=probes)
        loss = torch.nn.functional.cross_entropy(rt.normalize_power(y.sum(dim=1)), labels)
        loss.backward()

    t0 = time.perf_counter()
    run(8)
    est = (time.perf_counter() - t0) / 8
    T = int(max(16, min(3000, budget_s / max(est, 1e-6))))
    t0 = time.perf_counter()
    run(T)
    el = time.perf_counter() - t0
    B, N = x.shape[0], g.domain_shape[0]
    cells = B * N * N * T
    return {"value": cells / el / 1e9, "unit": "Gcell-steps/s", "cores": torch.get_num_threads(),
            "kind": "port",
            "sample": f"batch {B}, {N}x{N}, first {T} of 3000 steps, fwd+autograd bwd through "
                      f"the restated TimeStep.backward, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-kernel comparison")
    ap.add_argument("--tuning", type=str, default=None, help="JSON wt_tuning override")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    dev = torch.device("cuda:0")
    torch.cuda.init()
    tuning = json.loads(args.tuning) if args.tuning else None
    if tuning is None:
        from wavetorch_amd import autotune
        tuning = autotune.cached("train_150_b64")
    model, x, labels = build(dev, tuning)
    B, T = x.shape
    N = model.cell.geom.domain_shape[0]
    cells = B * N * N * T
    log(f"[bench] tuning {tuning}")
    dt, loss, prof = time_steps(model, x, labels, args.steps, args.warmup, profile=True)
    plan = next(iter(model._plans.values()))
    log(f"[bench] fused: {dt * 1e3:.2f} ms/step  loss {loss:.4f}  plan {plan.tuning}")
    value = cells / dt / 1e9
    # dominant kernel class by summed time
    cls = max(prof, key=lambda c: prof[c][0])
    ms, nl, work = prof[cls]
    avg_s = ms / max(nl, 1) / 1e3
    per_launch = work / max(nl, 1)
    achieved = BYTES[cls] * per_launch / avg_s
    for c, (m_, n_, w_) in prof.items():
        log(f"[bench] class {c} {CLASS_NAME[c]}: {n_} launches, {m_:.2f} ms total, "
            f"{(m_ / max(n_, 1)) * 1e3:.1f} us avg")
    out = {
        "metric": "fwd+adjoint Gcell-steps/s (150x150, batch 64, 3000 steps, training step)",
        "value": round(value, 3),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded vowel-like 10 kHz waveforms, random rho)",
        "config": {"workload": "train_150_b64: 150x150 grid incl. 20-cell PML, free-form rho "
                               "on the 110x110 interior, batch 64, 3000 steps, 3 intensity "
                               "probes, CE loss, adjoint with checkpoints every 64 steps, Adam",
                   "B": B, "Nx": N, "Ny": N, "T": T, "tuning": plan.tuning},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": None,
                     "kernel": CLASS_NAME[cls], "bytes_per_cell_step": BYTES[cls],
                     "cell_steps_per_launch": per_launch, "avg_launch_us": avg_s * 1e6,
                     "launches": nl},
        "whole_step_alg_GBps": 28.0 * cells / dt / 1e9,
    }
    if not args.no_plain:
        pm, px, pl = build(dev, {"plain": 1})
        pdt, _, _ = time_steps(pm, px, pl, 1, 1)
        log(f"[bench] plain: {pdt * 1e3:.2f} ms/step")
        out["plain_ms_per_step"] = round(pdt * 1e3, 3)
        out["speedup_vs_plain"] = round(pdt / dt, 3)
        del pm, px, pl
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
