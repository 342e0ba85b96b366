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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_train_150_b64.json")


def class_name(cls, tun):
    """Kernel (as rocprofv3 names it) of a kernel class under plan tuning `tun`."""
    if tun["plain"]:
        return ["wt::k_fwd_plain", "wt::k_fwd_plain (record)", "wt::k_adj_plain"][cls]
    if tun["kind"] == 1:
        return ["wt::k_fwd_tiled", "wt::k_fwd_tiled (record)", "wt::k_adj_tiled"][cls]
    if cls == 2:
        return f"wt::k_adj_cs<{tun['a_nt']}, {tun['a_cpt']}>"
    return f"wt::k_fwd_cs<{tun['nt']}, {tun['cpt']}, 0, {'true' if cls == 1 else 'false'}>"


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
    """Times `steps` training steps after `warmup`.  With `profile`, a second
    timed region of the same length runs with the library's HIP-event timing
    of every loop-kernel launch (events recorded on the launch stream); its
    per-class sums give the roofline, while the first region (no events)
    gives the step time."""
    from wavetorch_amd import _lib
    opt = torch.optim.Adam(model.parameters(), lr=1e-3)
    for _ in range(warmup):
        train_step(model, x, labels, opt)
    torch.cuda.synchronize()
    L = _lib.lib()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = train_step(model, x, labels, opt)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    prof = None
    if profile:
        import ctypes
        L.wt_profile(1)
        _lib.check(L.wt_profile_read(None, None, None), "wt_profile_read")
        t0 = time.perf_counter()
        for _ in range(steps):
            train_step(model, x, labels, opt)
        torch.cuda.synchronize()
        dt_p = (time.perf_counter() - t0) / steps
        ms = (ctypes.c_double * 3)()
        n = (ctypes.c_int64 * 3)()
        w = (ctypes.c_int64 * 3)()
        _lib.check(L.wt_profile_read(ms, n, w), "wt_profile_read")
        L.wt_profile(0)
        prof = {c: (ms[c], n[c], w[c]) for c in range(3)}
        prof["step_s"] = dt_p
    return dt, float(loss.item()), prof


def traffic(kname, tun):
    """Measured HBM bytes per launch of `kname` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary written by --summarize-pmc, when it was taken on the
    same kernel and tuning; else None."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    k = d.get("kernels", {}).get(kname)
    if k is None or d.get("tuning") != tun:
        return None
    return k["hbm_bytes_per_launch"]


def summarize_pmc(fetch_dir, write_dir, calib_bytes):
    """Per-kernel average FETCH_SIZE / WRITE_SIZE (KiB per dispatch) from two
    rocprofv3 --pmc runs, each scaled by what the counter reads for the
    bench's own device copy of `calib_bytes` bytes; writes TRAFFIC_FILE."""
    import collections
    import csv
    import glob

    def read(d, counter):
        agg = collections.defaultdict(list)
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] == counter:
                    agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        return agg

    fa, wa = read(fetch_dir, "FETCH_SIZE"), read(write_dir, "WRITE_SIZE")
    fe = {k: sum(v) / len(v) for k, v in fa.items()}
    wr = {k: sum(v) / len(v) for k, v in wa.items()}
    copy = [k for k in fa if "copy" in k.lower()]
    if not copy:
        raise SystemExit("no copy kernel in the FETCH_SIZE run (run bench.py with --calib-copy)")
    # the calibration copies are the largest dispatches of the copy kernel
    ck = max(copy, key=lambda k: max(fa[k]))
    cf, cw = max(fa[ck]), max(wa[ck])
    f_scale = calib_bytes / (cf * 1024.0)
    w_scale = calib_bytes / (cw * 1024.0)
    meta = json.load(open(os.path.join(fetch_dir, "bench_meta.json")))
    out = {"tuning": meta["tuning"], "calibration": {
        "kernel": ck, "bytes": calib_bytes, "fetch_kib": cf, "write_kib": cw,
        "fetch_scale": f_scale, "write_scale": w_scale}, "kernels": {}}
    lines = [f"copy calibration: {ck}: {calib_bytes} B -> FETCH {cf:.1f} KiB "
             f"(scale {f_scale:.4f}), WRITE {cw:.1f} KiB (scale {w_scale:.4f})"]
    for k in sorted(fe, key=lambda k: -fe[k]):
        if k == ck or k not in wr:
            continue
        name = k.split("(")[0].strip()
        name = name[5:] if name.startswith("void ") else name
        fb, wb = fe[k] * 1024.0 * f_scale, wr[k] * 1024.0 * w_scale
        out["kernels"][name] = {"dispatches": len(fa[k]), "fetch_kib": fe[k], "write_kib": wr[k],
                                "fetch_bytes": fb, "write_bytes": wb,
                                "hbm_bytes_per_launch": fb + wb}
        lines.append(f"{name}: {len(fa[k])} dispatches, FETCH {fe[k]:.1f} KiB, WRITE "
                     f"{wr[k]:.1f} KiB -> fetch {fb / 1e6:.3f} MB  write {wb / 1e6:.3f} MB "
                     f"per launch")
    os.makedirs(os.path.dirname(TRAFFIC_FILE), exist_ok=True)
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(out, f, indent=1)
    print("\n".join(lines))


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
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the timed region, copy MIB MiB device-to-device 3 times (the "
                         "known-size copy that scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--meta-out", type=str, default=None,
                    help="write the resolved tuning to this JSON file")
    ap.add_argument("--summarize-pmc", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "BYTES"),
                    help="summarise two rocprofv3 --pmc runs into the traffic file and exit")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], int(args.summarize_pmc[2]))
        return
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
    prof_step_s = prof.pop("step_s")
    cls = max(prof, key=lambda c: prof[c][0])
    ms, nl, work = prof[cls]
    avg_s = ms / max(nl, 1) / 1e3
    per_launch = work / max(nl, 1)
    achieved = BYTES[cls] * per_launch / avg_s
    for c, (m_, n_, w_) in prof.items():
        log(f"[bench] class {c} {class_name(c, plan.tuning)}: {n_} launches, {m_:.2f} ms total, "
            f"{(m_ / max(n_, 1)) * 1e3:.1f} us avg")
    kname = class_name(cls, plan.tuning)
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
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1),
                     "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": round(achieved / PEAK_HBM, 4),
                     "traffic": traffic(kname, plan.tuning),
                     "kernel": kname, "bytes_per_cell_step": BYTES[cls],
                     "cell_steps_per_launch": per_launch, "avg_launch_us": round(avg_s * 1e6, 2),
                     "launches": nl, "events_ms_per_step": round(prof_step_s * 1e3, 3),
                     "kernel_ms_per_step": {class_name(c, plan.tuning): round(prof[c][0] / args.steps, 3)
                                            for c in prof}},
        # the whole training step at 28 B per cell-step (12 forward + 16 adjoint)
        "step_roofline": {"achieved": round(28.0 * cells / dt / 1e9, 1), "unit": "GB/s",
                          "frac": round(28.0 * cells / dt / PEAK_HBM, 4)},
    }
    if args.meta_out:
        with open(args.meta_out, "w") as f:
            json.dump({"tuning": plan.tuning}, f)
    if args.calib_copy > 0:
        n = args.calib_copy * 2 ** 20 // 4
        a = torch.ones(n, device=dev)
        bb = torch.empty_like(a)
        for _ in range(3):
            bb.copy_(a)
        torch.cuda.synchronize()
        del a, bb
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
