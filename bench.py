"""Benchmark of the WaveRNN hot path on one MI355X.

Default workload: one training step of the 150x150, batch-64, 3000-step
WaveRNN -- the fused forward time loop (leapfrog + source + 3 intensity
probes), the cross-entropy loss head, the checkpointed adjoint (recompute
every K=64 steps) back to rho through the blur/projection, and an Adam update
of rho.  `--workload` selects the other configurations (100x100 batch 4,
nonlinear 150x150 batch 32, forward-only 1024x1024 with snapshots).  Inputs
are resident on the GPU before the timed region.

Prints ONE JSON line:
  value          Gcell-steps/s = B*Nx*Ny*T / step time / 1e9
  roofline       the dominant kernel (by summed time), its average launch
                 time from HIP events on its launch stream (a second timed
                 region of the same length), against the unit that binds it:
                 HBM (the bytes rocprofv3's FETCH_SIZE + WRITE_SIZE count per
                 launch, calibrated on a device copy, against 8 TB/s) or the
                 vector ALUs (SQ_INSTS_VALU per launch x 64 lanes against one
                 fp32 lane-operation per lane per clock: 256 CUs x 64 lanes
                 x 2.4 GHz); `frac` is the larger of the two.  The counters
                 come from profiles/traffic_<workload>.json (--calib-copy /
                 --summarize-pmc) when it was taken on the same plan.
                 `effective_bw` keeps the algorithmic figure (12 B forward,
                 16 B adjoint per cell-step x cell-steps per launch / time),
                 which temporal blocking lets exceed the HBM peak: it is not
                 a bound.
  step_roofline  the whole step at its algorithmic bytes per cell-step (the
                 same caveat)
  cpu_baseline   the torch-CPU restatement of the reference (oracle/ref_torch)
                 on a bounded sample of the same workload on this host's cores
  plain          the same step with the one-launch-per-step kernels

`--dtype float64` times a training workload in double precision instead (the
per-step float64 kernels; its roofline is algorithmic bytes over the launch
time of the longer sweep, from HIP events).
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
# non-packed fp32 VALU issue: 256 CUs x 4 SIMDs x 16 lanes per clock at the
# 2.4 GHz peak engine clock (the 157.3 TFLOP/s figure counts packed FMAs)
PEAK_VALU_LANE_OPS = 256 * 64 * 2.4e9
BYTES = {0: 12.0, 1: 12.0, 2: 16.0}  # algorithmic bytes per cell-step per kernel class

def traffic_file(workload):
    return os.path.join(ROOT, "profiles", f"traffic_{workload}.json")


def class_name(cls, tun, nonlinear=False):
    """Kernel (as rocprofv3 names it) of a kernel class under plan tuning `tun`."""
    if tun["plain"]:
        return ["wt::k_loop_fwd_plain", "wt::k_loop_fwd_plain (record)",
                "wt::k_loop_adj_nl_plain" if nonlinear else "wt::k_loop_adj_plain"][cls]
    if tun["kind"] == 1:
        return ["wt::k_fwd_tiled", "wt::k_fwd_tiled (record)",
                "wt::k_adj_nl_plain" if nonlinear else "wt::k_adj_tiled"][cls]
    if cls == 2:
        if nonlinear:
            return f"wt::k_adj_cs_nl<{tun['a_nt']}, {tun['a_cpt']}, false>"
        return f"wt::k_adj_cs<{tun['a_nt']}, {tun['a_cpt']}>"
    return f"wt::k_fwd_cs<{tun['nt']}, {tun['cpt']}, {8 if nonlinear else 0}, " \
        f"{'true' if cls == 1 else 'false'}>"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Workloads bench.py can time.  The default is the training step the headline
# metric is quoted on; the others are the remaining configurations of the
# hot path, timed with the same protocol.
#   kind     "train": forward + CE loss + checkpointed adjoint + Adam step;
#            "forward": forward only (no autograd tape), fields snapshotted
#   bytes    algorithmic HBM bytes per cell-step of one step
WORKLOADS = {
    "train_150_b64": dict(
        kind="train", bytes=28.0,
        metric="fwd+adjoint Gcell-steps/s (150x150, batch 64, 3000 steps, training step)",
        desc="150x150 grid incl. 20-cell PML, free-form rho on the 110x110 interior, batch 64, "
             "3000 steps, 3 intensity probes, CE loss, adjoint with checkpoints every 64 steps, "
             "Adam"),
    "nonlinear_150_b32": dict(
        kind="train", bytes=28.0,
        metric="fwd+adjoint Gcell-steps/s (150x150, batch 32, 3000 steps, nonlinear cell)",
        desc="as train_150_b64 at batch 32 with saturable damping b_nl=0.5, I_sat=0.1 and Kerr "
             "speed c_nl=0.05 clamped to [0.4, 1.1], dt=0.64"),
    "propagate_100_b4": dict(
        kind="train", bytes=28.0,
        metric="fwd+adjoint Gcell-steps/s (100x100, batch 4, 1000 steps)",
        desc="100x100 grid incl. 20-cell PML, free-form rho on the 60x60 interior, batch 4, "
             "1000 steps, windowed 500 Hz sine, 3 intensity probes, CE loss"),
    "lens_batch_256": dict(
        kind="batch", bytes=28.0,
        metric="fwd+adjoint Gcell-steps/s (256 lens instances of uneven size and length, "
               "one graph)",
        desc="256 independent lens problems, grid side from {64,96,128,192,256} incl. 16-cell "
             "PML, 1000-4000 steps, each with its own rho, a line source and a focal intensity "
             "probe, J = -sum_t u(focal)^2; forward + checkpointed adjoint of every instance as "
             "one captured HIP graph"),
    "holey_150_h64": dict(
        kind="geom", N=150, holes=64,
        metric="Holey geometry c forward+backward time (150x150, 64 holes)",
        desc="WaveGeometryHoley c = c0 + (c1-c0) P(sum_h exp(-d_h/r_h)) on a 150x150 grid with "
             "64 holes (beta 100), forward and backward to the hole centres and radii"),
    "scatter_1024": dict(
        kind="forward", bytes=12.0, field_stride=500,
        metric="forward Gcell-steps/s (1024x1024, 4000 steps, snapshots every 500)",
        desc="1024x1024 grid incl. 32-cell PML, c=1 with 20 circular scatterers (c=0.5), "
             "batch 1, 4000 steps, Gaussian pulse at the centre, field snapshot every 500 steps"),
}


def build(dev, tuning=None, workload="train_150_b64", T=None):
    from wavetorch_amd import workloads
    kw = {} if T is None else {"T": T}
    model, x, labels = getattr(workloads, workload)(**kw)
    model = model.to(dev)
    if tuning is not None:
        model.tuning = tuning
    fs = WORKLOADS[workload].get("field_stride")
    if fs:
        model.field_stride = fs
    return model, x.to(dev), (labels.to(dev) if labels is not None else None)


def train_step(model, x, labels, opt, fused_loss=False):
    """One training step: forward, the reference's criterion
    cross_entropy(normalize_power(y.sum(dim=1)), labels) (train.py:61-62) --
    torch ops, or the fused loss-head kernel with `fused_loss` -- the
    checkpointed adjoint and an Adam step."""
    from wavetorch_amd.loss import power_cross_entropy
    from wavetorch_amd.utils import normalize_power
    if labels is None:  # forward-only workload: snapshot fields, no tape
        with torch.no_grad():
            return model(x, output_fields=True)[0, -1, 0, 0]
    opt.zero_grad(set_to_none=True)
    y = model(x)
    if fused_loss:
        loss = power_cross_entropy(y, labels)
    else:
        loss = torch.nn.functional.cross_entropy(normalize_power(y.sum(dim=1)), labels)
    loss.backward()
    opt.step()
    return loss


def time_graph_steps(model, x, labels, steps, warmup):
    """Times `steps` replays of the training step captured as one HIP graph
    (wavetorch_amd.GraphedTrainStep: fused loss head, capturable Adam)."""
    from wavetorch_amd import GraphedTrainStep
    opt = torch.optim.Adam(model.parameters(), lr=1e-3, capturable=True)
    st = GraphedTrainStep(model, opt, x, labels, warmup=max(1, warmup), constrain=False)
    for _ in range(warmup):
        st(x, labels)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = st(x, labels)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, float(loss.item())


def time_steps(model, x, labels, steps, warmup, profile=False, fused_loss=False):
    """Times `steps` training steps after `warmup`.  With `profile`, a second
    timed region of the same length runs with the library's HIP-event timing
    of every loop-kernel launch (events recorded on the launch stream); its
    per-class sums give the roofline, while the first region (no events)
    gives the step time."""
    from wavetorch_amd import _lib
    opt = torch.optim.Adam(model.parameters(), lr=1e-3)
    for _ in range(warmup):
        train_step(model, x, labels, opt, fused_loss)
    torch.cuda.synchronize()
    L = _lib.lib()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = train_step(model, x, labels, opt, fused_loss)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    prof = None
    if profile:
        import ctypes
        L.wt_profile(1)
        _lib.check(L.wt_profile_read(None, None, None), "wt_profile_read")
        t0 = time.perf_counter()
        for _ in range(steps):
            train_step(model, x, labels, opt, fused_loss)
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


def kernel_counters(workload, kname, tun):
    """Per-launch counters of `kname` from the rocprofv3 summary written by
    --summarize-pmc (hbm_bytes_per_launch, and SQ_* counters when those runs
    were taken), when it was taken on the same kernel and tuning; else {}."""
    try:
        with open(traffic_file(workload)) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return {}
    k = d.get("kernels", {}).get(kname)
    if k is None or d.get("tuning") != tun:
        return {}
    return k


def traffic(workload, kname, tun):
    """Measured HBM bytes per launch of `kname` (FETCH_SIZE + WRITE_SIZE,
    calibrated), or None."""
    return kernel_counters(workload, kname, tun).get("hbm_bytes_per_launch")


def roofline(workload, kname, tun, cls, per_launch, avg_s, launches, extra=None):
    """The roofline object of the dominant kernel: the fraction of HBM and of
    VALU issue its measured counters reach at its measured launch time,
    `frac` the larger (the binding unit), and the algorithmic-equivalent
    bandwidth under its own key."""
    k = kernel_counters(workload, kname, tun)
    out = {"kernel": kname, "avg_launch_us": round(avg_s * 1e6, 2), "launches": launches,
           "cell_steps_per_launch": per_launch}
    units = {}
    if k.get("hbm_bytes_per_launch"):
        bw = k["hbm_bytes_per_launch"] / avg_s
        units["hbm"] = {"achieved": round(bw / 1e9, 1), "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                        "frac": round(bw / PEAK_HBM, 4),
                        "bytes_per_launch": k["hbm_bytes_per_launch"]}
    if k.get("SQ_INSTS_VALU"):
        ops = k["SQ_INSTS_VALU"] * 64 / avg_s
        units["valu"] = {"achieved": round(ops / 1e12, 2),
                         "peak": round(PEAK_VALU_LANE_OPS / 1e12, 2),
                         "unit": "Tlane-op/s", "frac": round(ops / PEAK_VALU_LANE_OPS, 4),
                         "insts_per_launch": k["SQ_INSTS_VALU"]}
    for key_ in ("SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_WAIT_INST_ANY", "SQ_BUSY_CYCLES",
                 "SQ_WAVE_CYCLES", "SQ_ACTIVE_INST_VALU", "SQ_INSTS_SALU"):
        if key_ in k:
            out.setdefault("counters_per_launch", {})[key_] = k[key_]
    if units:
        bound = max(units, key=lambda u: units[u]["frac"])
        b = units[bound]
        out.update({"bound": bound, "achieved": b["achieved"], "peak": b["peak"],
                    "unit": b["unit"], "frac": b["frac"]})
        out.update(units)
    else:  # no counter summary for this plan: no measured fraction
        out.update({"bound": "hbm", "achieved": None, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                    "frac": None})
    out["traffic"] = k.get("hbm_bytes_per_launch")
    eff = BYTES[cls] * per_launch / avg_s
    out["effective_bw"] = {"basis": f"{BYTES[cls]:g} algorithmic B per cell-step x cell-steps "
                                    "per launch / launch time (not a bound)",
                           "achieved": round(eff / 1e9, 1), "unit": "GB/s",
                           "frac_of_hbm_peak": round(eff / PEAK_HBM, 4)}
    if extra:
        out.update(extra)
    return out


def summarize_pmc(fetch_dir, write_dir, calib_bytes, extra_dirs=()):
    """Per-kernel average FETCH_SIZE / WRITE_SIZE (KiB per dispatch) from two
    rocprofv3 --pmc runs, each scaled by what the counter reads for the
    bench's own device copy of `calib_bytes` bytes, plus the per-dispatch
    average of every counter found in `extra_dirs` (further --pmc runs, e.g.
    SQ_INSTS_VALU, SQ_INSTS_LDS, SQ_LDS_BANK_CONFLICT, SQ_WAIT_INST_ANY);
    writes profiles/traffic_<workload>.json."""
    import collections
    import csv
    import glob

    def short(k):
        name = k.split("(")[0].strip()
        return name[5:] if name.startswith("void ") else name

    meta = json.load(open(os.path.join(fetch_dir, "bench_meta.json")))
    # kernels whose summary covers only their last n dispatches (the bench's
    # own timed eager pass, when the run also holds other launches of them)
    tail = meta.get("tail_dispatches", {})

    def keep(k, pairs):
        vals = [v for _, v in sorted(pairs)]
        n = tail.get(short(k))
        return vals if n is None else vals[max(len(vals) - n, 0):]

    def read(d, counter):
        agg = collections.defaultdict(list)
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] == counter:
                    agg[r["Kernel_Name"]].append((int(r.get("Dispatch_Id") or 0),
                                                  float(r["Counter_Value"])))
        return {k: kv for k, kv in ((k, keep(k, v)) for k, v in agg.items()) if kv}

    def read_all(d):
        agg = collections.defaultdict(lambda: collections.defaultdict(list))
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                agg[r["Kernel_Name"]][r["Counter_Name"]].append(
                    (int(r.get("Dispatch_Id") or 0), float(r["Counter_Value"])))
        out = {}
        for k, c in agg.items():
            kc = {cn: kv for cn, kv in ((cn, keep(k, v)) for cn, v in c.items()) if kv}
            if kc:
                out[k] = kc
        return out

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
    out = {"tuning": meta["tuning"], "calibration": {
        "kernel": ck, "bytes": calib_bytes, "fetch_kib": cf, "write_kib": cw,
        "fetch_scale": f_scale, "write_scale": w_scale}, "kernels": {}}
    lines = [f"copy calibration: {ck}: {calib_bytes} B -> FETCH {cf:.1f} KiB "
             f"(scale {f_scale:.4f}), WRITE {cw:.1f} KiB (scale {w_scale:.4f})"]
    for k in sorted(fe, key=lambda k: -fe[k]):
        if k == ck or k not in wr:
            continue
        name = short(k)
        fb, wb = fe[k] * 1024.0 * f_scale, wr[k] * 1024.0 * w_scale
        out["kernels"][name] = {"dispatches": len(fa[k]), "fetch_kib": fe[k], "write_kib": wr[k],
                                "fetch_bytes": fb, "write_bytes": wb,
                                "hbm_bytes_per_launch": fb + wb}
        lines.append(f"{name}: {len(fa[k])} dispatches, FETCH {fe[k]:.1f} KiB, WRITE "
                     f"{wr[k]:.1f} KiB -> fetch {fb / 1e6:.3f} MB  write {wb / 1e6:.3f} MB "
                     f"per launch")
    for d in extra_dirs:
        for k, cnt in read_all(d).items():
            name = short(k)
            if name not in out["kernels"]:
                continue
            for cn, vals in cnt.items():
                out["kernels"][name][cn] = sum(vals) / len(vals)
            lines.append(f"{name}: " + ", ".join(f"{cn} {sum(v) / len(v):.4g}"
                                                 for cn, v in sorted(cnt.items())))
    path = traffic_file(meta.get("workload", "train_150_b64"))
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("\n".join(lines))


def run_f64(args, dev):
    """The training step of a "train" workload in float64 (set_dtype('float64'),
    the precision the reference's example configurations train in): geometry
    and loss as torch ops in float64, the double-precision time loop
    (k64_fwd one launch per step, the adjoint as one k64_adj launch per
    step over the stored fields).  HIP events on the launch stream bracket
    the forward sweep and the backward of each timed step; the roofline is
    that of the sweep that takes longer, at its algorithmic bytes (24 B per
    cell-step forward: read u_n, u_{n-1}, write u_{n+1}; 32 B adjoint: read
    lambda_{m+1}, lambda_{m+2}, u_m, write lambda_m)."""
    from wavetorch_amd.utils import normalize_power, set_dtype
    wl = WORKLOADS[args.workload]
    set_dtype("float64")
    try:
        model, x, labels = build(dev, None, args.workload)
        x = x.double()
        B, T = x.shape
        Nx, Ny = model.cell.geom.domain_shape
        cells = B * Nx * Ny * T
        opt = torch.optim.Adam(model.parameters(), lr=1e-3)
        for _ in range(args.warmup):
            train_step(model, x, labels, opt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            loss = train_step(model, x, labels, opt)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        log(f"[bench] float64: {dt * 1e3:.2f} ms/step  loss {float(loss.detach()):.4f}")
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        fwd_ms = bwd_ms = 0.0
        for _ in range(args.steps):
            opt.zero_grad(set_to_none=True)
            ev[0].record()
            y = model(x)
            ev[1].record()
            torch.nn.functional.cross_entropy(normalize_power(y.sum(dim=1)), labels).backward()
            ev[2].record()
            opt.step()
            torch.cuda.synchronize()
            fwd_ms += ev[0].elapsed_time(ev[1]) / args.steps
            bwd_ms += ev[1].elapsed_time(ev[2]) / args.steps
        log(f"[bench] float64 sweeps: forward {fwd_ms:.2f} ms, backward {bwd_ms:.2f} ms")
        fwd = fwd_ms >= bwd_ms
        kname = "k64_fwd" if fwd else "k64_adj"
        avg_s = (fwd_ms if fwd else bwd_ms) / 1e3 / T
        per_launch = B * Nx * Ny
        byts = (24.0 if fwd else 32.0) * per_launch
        eff = {"basis": f"{byts / per_launch:g} algorithmic B per cell-step x cells / launch "
                        "time", "achieved": round(byts / avg_s / 1e9, 1), "unit": "GB/s",
               "frac_of_hbm_peak": round(byts / avg_s / PEAK_HBM, 4)}
        # counter bytes (profiles/traffic_<workload>_f64.json, --summarize-pmc)
        knames = ["wt::k64_fwd"] if fwd else ["wt::k64_adj"]
        ks = [kernel_counters(f"{args.workload}_f64", k, "float64") for k in knames]
        traffic_b = (sum(k["hbm_bytes_per_launch"] for k in ks)
                     if all(k.get("hbm_bytes_per_launch") for k in ks) else None)
        if traffic_b:
            roof = {"achieved": round(traffic_b / avg_s / 1e9, 1),
                    "frac": round(traffic_b / avg_s / PEAK_HBM, 4),
                    "basis": "HBM bytes counted by FETCH_SIZE + WRITE_SIZE (calibrated) per "
                             "launch / launch time from HIP events over the sweep"}
        else:
            roof = {"achieved": eff["achieved"], "frac": eff["frac_of_hbm_peak"],
                    "basis": eff["basis"]}
        out = {
            "metric": wl["metric"].replace("Gcell-steps/s (", "Gcell-steps/s (float64, "),
            "value": round(cells / dt / 1e9, 3),
            "unit": "Gcell-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(dt * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic (seeded inputs and rho, see wavetorch_amd/workloads.py)",
            "config": {"workload": f"{args.workload} in float64: {wl['desc']}", "B": B,
                       "Nx": Nx, "Ny": Ny, "T": T},
            "roofline": {"kernel": kname, "bound": "hbm", "achieved": roof["achieved"],
                         "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": roof["frac"],
                         "basis": roof["basis"], "avg_launch_us": round(avg_s * 1e6, 2),
                         "traffic": traffic_b, "effective_bw": eff,
                         "forward_ms_per_step": round(fwd_ms, 3),
                         "backward_ms_per_step": round(bwd_ms, 3)},
            "loss_finite": bool(np.isfinite(float(loss.detach()))),
        }
        if args.meta_out:
            with open(args.meta_out, "w") as f:
                json.dump({"tuning": "float64", "workload": f"{args.workload}_f64"}, f)
        if args.calib_copy > 0:
            a = torch.ones(args.calib_copy * 2 ** 20 // 4, device=dev, dtype=torch.float32)
            bb = torch.empty_like(a)
            for _ in range(3):
                bb.copy_(a)
            torch.cuda.synchronize()
            del a, bb
        del model, x, labels, opt, y
        if not args.no_cpu:
            log("[bench] cpu baseline (float64) ...")
            out["cpu_baseline"] = cpu_baseline(args.workload)
    finally:
        set_dtype("float32")
    print(json.dumps(out), flush=True)


def cpu_baseline(workload="train_150_b64", budget_s=15.0):
    """The reference's op sequence (oracle/ref_torch: the torch-CPU restatement
    of rnn.py/cell.py/source.py/probe.py/geom.py with its autograd tape) on
    the same workload, with T cut to fit ~budget_s on this host's cores."""
    from oracle import ref_torch as rt
    from wavetorch_amd import workloads
    spec = WORKLOADS[workload]
    model, x, labels = getattr(workloads, workload)()
    x = x.to(torch.get_default_dtype())  # float64 under set_dtype('float64')
    g, cell = model.cell.geom, model.cell
    torch.set_num_threads(os.cpu_count() or 1)
    train = spec["kind"] == "train"
    rho = g.rho.detach().clone().requires_grad_(train)
    src = [(s_.x.clone(), s_.y.clone()) for s_ in model.sources]
    probes = [(p_.x.clone(), p_.y.clone(), bool(p_.intensity)) for p_ in model.probes]
    nlkw = {}
    if float(cell.satdamp_b0) > 0 or float(cell.c_nl) != 0:
        nlkw = dict(satdamp_b0=cell.satdamp_b0, satdamp_uth=cell.satdamp_uth, c_nl=cell.c_nl,
                    c_clamp=cell.c_clamp)

    def run(T):
        with torch.set_grad_enabled(train):
            c = rt.freeform_c(rho, g.c0.item(), g.c1.item(), g.blur_kernel, int(g.blur_N.item()),
                              g.eta.item(), g.beta.item())
            y = rt.rnn_forward(x[:, :T].contiguous(), c=c, rho=rho, b_geom=g.b, dt=cell.dt,
                               h=g.h, sources=src, probes=probes, output_fields=not train,
                               **nlkw)
            if train:
                loss = torch.nn.functional.cross_entropy(rt.normalize_power(y.sum(dim=1)),
                                                         labels)
                loss.backward()

    B, Tfull = x.shape
    Nx, Ny = g.domain_shape
    t0 = time.perf_counter()
    run(4)
    est = (time.perf_counter() - t0) / 4
    T = int(max(8, min(Tfull, budget_s / max(est, 1e-6))))
    if not train:  # the reference stacks every field: keep the sample within ~2 GB
        T = min(T, max(8, int(2e9 // (B * Nx * Ny * 4))))
    t0 = time.perf_counter()
    run(T)
    el = time.perf_counter() - t0
    cells = B * Nx * Ny * T
    what = "fwd+autograd bwd through the restated TimeStep.backward" if train else \
        "forward keeping every field (rnn.py:66-70)"
    prec = "float64, " if x.dtype == torch.float64 else ""
    return {"value": cells / el / 1e9, "unit": "Gcell-steps/s", "cores": torch.get_num_threads(),
            "kind": "port",
            "sample": f"{prec}batch {B}, {Nx}x{Ny}, first {T} of {Tfull} steps, {what}, "
                      f"{el:.1f} s"}


def _holey_torch_ops(hx, hy, hr, shape, eta, beta, c0, c1):
    """geom.py:109-135 as the reference computes it -- a few torch ops per hole
    -- run on the device: the one-op-per-launch form the kernels replace."""
    xv = torch.arange(0, shape[0], dtype=hx.dtype, device=hx.device)
    yv = torch.arange(0, shape[1], dtype=hx.dtype, device=hx.device)
    x, y = torch.meshgrid(xv, yv, indexing="ij")
    rho = torch.zeros(shape, dtype=hx.dtype, device=hx.device)
    for ri, xi, yi in zip(hr, hx, hy):
        r = torch.sqrt((x - xi).pow(2) + (y - yi).pow(2))
        rho = rho + torch.exp(-r / ri)
    p = (np.tanh(beta * eta) + torch.tanh(beta * (rho - eta))) / (
        np.tanh(beta * eta) + np.tanh(beta * (1 - eta)))
    return c0 + (c1 - c0) * p


def run_geom(args, dev):
    """Times WaveGeometryHoley.c forward + backward (the wt_geom_holey_c
    kernels) against the reference's torch ops on the same device."""
    from wavetorch_amd import WaveGeometryHoley
    wl = WORKLOADS[args.workload]
    N, H = wl["N"], wl["holes"]
    rng = np.random.default_rng(17)
    hx = (rng.uniform(25, N - 25, H) + 0.3).astype(np.float32)
    hy = (rng.uniform(25, N - 25, H) + 0.2).astype(np.float32)
    hr = rng.uniform(2.0, 8.0, H).astype(np.float32)
    g = WaveGeometryHoley((N, N), 1.0, c0=1.0, c1=0.5, abs_N=20, eta=0.5, beta=100.0, x=hx,
                          y=hy, r=hr).to(dev)
    w = torch.from_numpy(rng.standard_normal((N, N)).astype(np.float32)).to(dev)
    pt = [torch.tensor(v, device=dev, requires_grad=True) for v in (hx, hy, hr)]

    def kern():
        g.zero_grad(set_to_none=True)
        (g.c * w).sum().backward()

    def ops():
        for q in pt:
            q.grad = None
        (_holey_torch_ops(*pt, (N, N), 0.5, 100.0, 1.0, 0.5) * w).sum().backward()

    def timed(f, reps):
        for _ in range(max(args.warmup, 1)):
            f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    inner = 20
    dt = timed(kern, args.steps * inner)
    out = {"metric": wl["metric"], "value": round(dt * 1e6, 1), "unit": "us",
           "n_gpus": 1, "steps": args.steps * inner, "warmup": args.warmup,
           "ms_per_step": round(dt * 1e3, 4), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded hole centres and radii)",
           "config": {"workload": f"{args.workload}: {wl['desc']}", "Nx": N, "Ny": N,
                      "holes": H}}
    if not args.no_plain:
        pdt = timed(ops, max(1, args.steps))
        log(f"[bench] torch ops: {pdt * 1e6:.1f} us")
        out["plain_ms_per_step"] = round(pdt * 1e3, 4)
        out["speedup_vs_plain"] = round(pdt / dt, 3)
    print(json.dumps(out), flush=True)


def run_batch(args, dev):
    """Batched instances: the forward + adjoint of every instance, captured
    once as a HIP graph and replayed; inputs (c of every instance from its
    rho, x, dJ/dy) resident in HBM."""
    import ctypes
    from wavetorch_amd import _lib, autotune, workloads
    from wavetorch_amd._loop import Plan, WaveLoop
    from wavetorch_amd.batch import Instance, InstanceBatch
    wl = WORKLOADS[args.workload]
    items = [(m.to(dev), x.to(dev)) for m, x in workloads.lens_batch_256()]
    with torch.no_grad():
        cs = [m.cell.geom.c.detach().clone() for m, _ in items]
    xs = [x for _, x in items]
    ybars = [-torch.ones((x.shape[0], 1), device=dev) for x in xs]
    insts = [Instance.from_model(m, x.shape[0]) for m, x in items]
    tuning = json.loads(args.tuning) if args.tuning else autotune.cached_batch(insts)
    bt = InstanceBatch(insts, dev, tuning=tuning)
    bt.set_inputs(cs, xs, ybars)
    cells = bt.cell_steps()
    log(f"[bench] {args.workload}: {len(insts)} instances in {len(bt.groups)} groups, "
        f"{cells:.3g} cell-steps, tuning {bt.tuning[0]}")
    s = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(s):
        bt.capture()
        for _ in range(args.warmup):
            bt.replay()
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bt.replay()
        s.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        bt.fwd_adj()
        s.synchronize()
        dt_eager = time.perf_counter() - t0
    # per-kernel-class HIP-event timing of one eager run (events cannot sit
    # inside the graph) of the same batch as ONE part, so that each launch
    # runs alone on the chip: with several parts the launches overlap and an
    # event pair on a part's stream also times the wait for CUs held by the
    # other parts' launches.  The timed plan's workspaces go first (a plan
    # that keeps the whole record can hold a large share of HBM).
    bt_tuning = bt.tuning
    bt_groups = [{"side": bt.instances[idx[0]].cfg.Nx, "instances": len(idx),
                  "strips": bt.tuning[g]["strips"], "tcols": bt.tuning[g]["tcols"]}
                 for g, idx in enumerate(bt.groups)]
    bt_ws = (bt.saved.numel() + bt.temp.numel()) / 2 ** 30
    del bt
    torch.cuda.empty_cache()
    bt1 = InstanceBatch(insts, dev, tuning=dict(tuning or {}, streams=1))
    bt1.set_inputs(cs, xs, ybars)
    with torch.cuda.stream(s):
        bt1.fwd_adj()  # the first launches of each variant set their LDS limits
        s.synchronize()
        L = _lib.lib()
        L.wt_profile(1)
        _lib.check(L.wt_profile_read(None, None, None), "wt_profile_read")
        t0 = time.perf_counter()
        bt1.fwd_adj()
        s.synchronize()
        dt_eager1 = time.perf_counter() - t0
        ms = (ctypes.c_double * 3)()
        n = (ctypes.c_int64 * 3)()
        w = (ctypes.c_int64 * 3)()
        _lib.check(L.wt_profile_read(ms, n, w), "wt_profile_read")
        L.wt_profile(0)
    prof = {c: (ms[c], n[c], w[c]) for c in range(3)}
    cls = max(prof, key=lambda c: prof[c][0])
    msc, nl, work = prof[cls]
    avg_s = msc / max(nl, 1) / 1e3
    per_launch = work / max(nl, 1)
    t0g = bt1.tuning[0]
    names = [f"wt::k_fwd_cs_batch<{t0g['nt']}, {t0g['cpt']}, 0, false>",
             f"wt::k_fwd_cs_batch<{t0g['nt']}, {t0g['cpt']}, 0, true>",
             f"wt::k_adj_cs_batch<{t0g['a_nt']}, {t0g['a_cpt']}>"]
    for c in range(3):
        log(f"[bench] class {c} {names[c]}: {n[c]} launches, {ms[c]:.2f} ms, "
            f"{ms[c] / max(n[c], 1) * 1e3:.1f} us avg")
    log(f"[bench] graph replay {dt * 1e3:.2f} ms, eager {dt_eager * 1e3:.2f} ms "
        f"({bt_tuning[0].get('streams', 1)} parts, workspaces {bt_ws:.1f} GiB); one part: "
        f"eager {dt_eager1 * 1e3:.2f} ms with HIP events")
    out = {
        "metric": wl["metric"], "value": round(cells / dt / 1e9, 3), "unit": "Gcell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded sizes, lengths and rho, see wavetorch_amd/workloads.py)",
        "config": {"workload": f"{args.workload}: {wl['desc']}", "instances": len(insts),
                   "groups": bt_groups, "cell_steps": cells, "tuning": bt_tuning[0],
                   "workspace_gib": round(bt_ws, 2)},
        # the dominant kernel's launches in the one-part eager run (the
        # counter summary covers the same dispatches, the last of the run)
        "roofline": roofline(args.workload, names[cls], bt1.tuning[0], cls, per_launch, avg_s,
                             nl, {"measured_on": "the batch as one part, eager, HIP events",
                                  "eager_one_part_ms": round(dt_eager1 * 1e3, 3),
                                  "eager_ms": round(dt_eager * 1e3, 3),
                                  "parts": bt_tuning[0].get("streams", 1)}),
        "step_roofline": {"bytes_per_cell_step": wl["bytes"],
                          "basis": "algorithmic bytes (not a bound)",
                          "achieved": round(wl["bytes"] * cells / dt / 1e9, 1), "unit": "GB/s",
                          "frac_of_hbm_peak": round(wl["bytes"] * cells / dt / PEAK_HBM, 4)},
    }
    if args.meta_out:
        with open(args.meta_out, "w") as f:
            json.dump({"tuning": bt1.tuning[0], "workload": args.workload,
                       "tail_dispatches": {names[c]: int(n[c]) for c in range(3)}}, f)
    if args.calib_copy > 0:
        nn = args.calib_copy * 2 ** 20 // 4
        a = torch.ones(nn, device=dev)
        bb = torch.empty_like(a)
        for _ in range(3):
            bb.copy_(a)
        torch.cuda.synchronize()
        del a, bb
    del bt1
    if not args.no_plain:
        # the plain per-instance path: one plan per instance, one launch per
        # step, instance after instance (plans made before the clock starts)
        plans = [Plan(m.loop_config(), 1, x.shape[0], m.cell.geom.cmax, dev, {"plain": 1})
                 for m, x in items]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for (m, x), c, yb, p in zip(items, cs, ybars, plans):
            ct = c.clone().requires_grad_(True)
            y = WaveLoop.apply(x[None], ct, m.cell.geom.b, None, p, False, 1)
            y.backward(yb[None])
        torch.cuda.synchronize()
        pdt = time.perf_counter() - t0
        log(f"[bench] plain per-instance loop: {pdt * 1e3:.1f} ms")
        out["plain_ms_per_step"] = round(pdt * 1e3, 3)
        out["speedup_vs_plain"] = round(pdt / dt, 3)
        del plans
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_batch(items, cells)
    print(json.dumps(out), flush=True)


def cpu_baseline_batch(items, total_cells, budget_s=15.0):
    """The reference's op sequence (oracle/ref_torch) on a seeded subset of
    the instances (each truncated in time), extrapolated to the whole batch
    by cell-steps."""
    from oracle import ref_torch as rt
    torch.set_num_threads(os.cpu_count() or 1)
    rng = np.random.default_rng(0)
    order = rng.permutation(len(items))
    done_cells, el = 0, 0.0
    picked = []
    for i in order:
        m, x = items[i]
        g, cell = m.cell.geom, m.cell
        N = g.domain_shape[0]
        T = min(x.shape[0], 200)
        rho = g.rho.detach().cpu().clone().requires_grad_(True)
        src = [(s_.x.cpu().clone(), s_.y.cpu().clone()) for s_ in m.sources]
        probes = [(p_.x.cpu().clone(), p_.y.cpu().clone(), True) for p_ in m.probes]
        t0 = time.perf_counter()
        c = rt.freeform_c(rho, g.c0.item(), g.c1.item(), g.blur_kernel.cpu(),
                          int(g.blur_N.item()), g.eta.item(), g.beta.item())
        y = rt.rnn_forward(x[:T].cpu()[None], c=c, rho=rho, b_geom=g.b.cpu(), dt=cell.dt.cpu(),
                           h=g.h.cpu(), sources=src, probes=probes)
        (-y.sum()).backward()
        el += time.perf_counter() - t0
        done_cells += N * N * T
        picked.append(f"{N}x{N}x{T}")
        if el > budget_s:
            break
    rate = done_cells / el
    return {"value": rate / 1e9, "unit": "Gcell-steps/s", "cores": torch.get_num_threads(),
            "kind": "port",
            "sample": f"{len(picked)} of {len(items)} instances (seeded pick), first 200 steps "
                      f"each, fwd+autograd bwd, {el:.1f} s; the whole batch "
                      f"({total_cells:.3g} cell-steps) would take {total_cells / rate:.0f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-kernel comparison")
    ap.add_argument("--workload", default="train_150_b64", choices=sorted(WORKLOADS),
                    help="configuration to time (default: the headline training step)")
    ap.add_argument("--tuning", type=str, default=None, help="JSON wt_tuning override")
    ap.add_argument("--dtype", choices=("float32", "float64"), default="float32",
                    help="float64: the training step in double precision (train workloads)")
    ap.add_argument("--loss", choices=("torch", "fused"), default="torch",
                    help="loss head of the training step: the reference's torch ops (default) "
                         "or the fused wt_loss_power_ce kernel")
    ap.add_argument("--graph", action="store_true",
                    help="also time the step captured as one HIP graph (GraphedTrainStep)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the timed region, copy MIB MiB device-to-device 3 times (the "
                         "known-size copy that scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--meta-out", type=str, default=None,
                    help="write the resolved tuning to this JSON file")
    ap.add_argument("--summarize-pmc", nargs="+", metavar="DIR",
                    help="FETCH_DIR WRITE_DIR BYTES [COUNTER_DIR ...]: summarise the "
                         "FETCH_SIZE and WRITE_SIZE rocprofv3 --pmc runs (calibrated on the "
                         "BYTES-byte copy) and any further --pmc runs into the traffic file "
                         "and exit")
    args = ap.parse_args()
    if args.summarize_pmc:
        if len(args.summarize_pmc) < 3:
            ap.error("--summarize-pmc needs FETCH_DIR WRITE_DIR BYTES")
        summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], int(args.summarize_pmc[2]),
                      args.summarize_pmc[3:])
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    dev = torch.device("cuda:0")
    torch.cuda.init()
    wl = WORKLOADS[args.workload]
    if args.dtype == "float64":
        if wl["kind"] != "train":
            ap.error("--dtype float64 times the training workloads")
        run_f64(args, dev)
        return
    if wl["kind"] == "batch":
        run_batch(args, dev)
        return
    if wl["kind"] == "geom":
        run_geom(args, dev)
        return
    tuning = json.loads(args.tuning) if args.tuning else None
    if tuning is None:
        from wavetorch_amd import autotune
        tuning = autotune.cached(args.workload)
    model, x, labels = build(dev, tuning, args.workload)
    B, T = x.shape
    Nx, Ny = model.cell.geom.domain_shape
    cells = B * Nx * Ny * T
    log(f"[bench] {args.workload}: tuning {tuning}")
    fused = args.loss == "fused" and labels is not None
    dt, loss, prof = time_steps(model, x, labels, args.steps, args.warmup, profile=True,
                                fused_loss=fused)
    plan = next(iter(model._plans.values()))
    log(f"[bench] fused: {dt * 1e3:.2f} ms/step  loss {loss:.4f}  plan {plan.tuning}")
    value = cells / dt / 1e9
    # dominant kernel class by summed time
    prof_step_s = prof.pop("step_s")
    cls = max(prof, key=lambda c: prof[c][0])
    ms, nl, work = prof[cls]
    avg_s = ms / max(nl, 1) / 1e3
    per_launch = work / max(nl, 1)
    nonlin = plan.cfg.nonlinear
    for c, (m_, n_, w_) in prof.items():
        log(f"[bench] class {c} {class_name(c, plan.tuning, nonlin)}: {n_} launches, "
            f"{m_:.2f} ms total, {(m_ / max(n_, 1)) * 1e3:.1f} us avg")
    kname = class_name(cls, plan.tuning, nonlin)
    out = {
        "metric": wl["metric"],
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
        "data": "synthetic (seeded inputs and rho, see wavetorch_amd/workloads.py)",
        "config": {"workload": f"{args.workload}: {wl['desc']}", "B": B, "Nx": Nx, "Ny": Ny,
                   "T": T, "tuning": plan.tuning},
        "roofline": roofline(args.workload, kname, plan.tuning, cls, per_launch, avg_s, nl,
                             {"events_ms_per_step": round(prof_step_s * 1e3, 3),
                              "kernel_ms_per_step": {class_name(c, plan.tuning, nonlin):
                                                     round(prof[c][0] / args.steps, 3)
                                                     for c in prof if prof[c][1]}}),
        # the whole step at its algorithmic bytes per cell-step (28 B = 12 forward
        # + 16 adjoint for a training step, 12 B forward only)
        "step_roofline": {"bytes_per_cell_step": wl["bytes"],
                          "basis": "algorithmic bytes (not a bound)",
                          "achieved": round(wl["bytes"] * cells / dt / 1e9, 1), "unit": "GB/s",
                          "frac_of_hbm_peak": round(wl["bytes"] * cells / dt / PEAK_HBM, 4)},
    }
    if labels is not None:
        out["loss_head"] = args.loss
        # the timed step's loss, then one more step with the on-device
        # non-finite checks of the loop's output and gradients
        out["loss_finite"] = bool(np.isfinite(loss))
        model.check_finite = True
        train_step(model, x, labels, torch.optim.Adam(model.parameters(), lr=1e-3), fused)
        flags = model.nonfinite.tolist()
        model.check_finite = False
        out["output_finite"], out["grad_finite"] = flags[0] == 0, flags[1] == 0
        if not (out["loss_finite"] and flags == [0, 0]):
            log(f"[bench] non-finite values: loss {loss}, flags (output, gradient) {flags}")
    if args.meta_out:
        with open(args.meta_out, "w") as f:
            json.dump({"tuning": plan.tuning, "workload": args.workload}, f)
    if args.graph and labels is not None:
        gm, gx, gl = build(dev, tuning, args.workload)
        gdt, _ = time_graph_steps(gm, gx, gl, args.steps, args.warmup)
        log(f"[bench] graph: {gdt * 1e3:.2f} ms/step")
        out["graph_ms_per_step"] = round(gdt * 1e3, 3)
        del gm, gx, gl
    if args.calib_copy > 0:
        n = args.calib_copy * 2 ** 20 // 4
        a = torch.ones(n, device=dev)
        bb = torch.empty_like(a)
        for _ in range(3):
            bb.copy_(a)
        torch.cuda.synchronize()
        del a, bb
    del model
    if not args.no_plain:
        pm, px, pl = build(dev, {"plain": 1}, args.workload)
        pdt, _, _ = time_steps(pm, px, pl, max(5, args.steps), 2)
        log(f"[bench] plain: {pdt * 1e3:.2f} ms/step")
        out["plain_ms_per_step"] = round(pdt * 1e3, 3)
        out["speedup_vs_plain"] = round(pdt / dt, 3)
        del pm, px, pl
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
