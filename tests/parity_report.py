"""Per-workload parity report of the HIP path.

Test infrastructure: the oracle is the checker here, never the thing
measured.  For each workload of wavetorch_amd/workloads.py this runs the
library on cuda:0 with the plan bench.py uses and prints one JSON line:

* forward -- the probe series (the snapshot fields for the forward-only
  scatter_1024) against the C oracle's fp32 loop given the GPU's c (bit for
  bit), and the relative L2 error against the fp64 oracle (c from the fp64
  rho -> c map) and against the reference's fp32 op sequence (whose c comes
  from torch's CPU tanh, so it agrees with the GPU's to rounding only);
* gradient -- dJ/drho of the workload's loss (cross-entropy of the
  normalised integrated probe intensities, as bench.py) relative to (a) the
  fp64 oracle adjoint chained through the fp64 rho -> c map, (b) the
  reference's fp32 op sequence differentiated by torch autograd
  (oracle/ref_torch.py: the reference's own gradient), (c) central
  differences of the fp64 loss on 32 random rho pixels of the design region
  (b == 0);
* nonlinear_150_b32: the same over its first 300 steps, then the gradient
  error of the GPU and of the reference's fp32 op sequence against the fp64
  oracle at 450 and 600 steps (the exact gradient of this batch grows
  ill-conditioned past ~400 steps: one sample's adjoint grows by orders of
  magnitude, and no fp32 computation, the reference's included, stays
  within 1e-4 of fp64), plus the workspace bytes, peak device memory and
  step time of the full 3000-step training step for checkpoint intervals K
  in {16, 64, 256};
* lens_batch_256: per instance, the probe series bitwise against the fp32
  oracle and dJ/dc, dJ/dx against the fp64 oracle given the GPU's c (max and
  median over the 256 instances); for the instances with the largest
  errors, the same for the reference's fp32 op sequence.

The CPU legs (tests/parity_cpu.py: numpy and the C oracle) run in worker
processes forked before this process imports torch, so only this process
ever loads the GPU stack.

    python -m tests.parity_report [workload ...] [--workers 16] [--out FILE]
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from tests import parity_cpu as pc  # noqa: E402  (numpy and the C oracle only)

FD_PIXELS = 32
FD_DELTA = 1e-5
NL_STEPS = 300
NL_CONDITIONING = (450, 600)
WORKLOADS = ("propagate_100_b4", "train_150_b64", "nonlinear_150_b32", "scatter_1024",
             "lens_batch_256")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    """||a - b|| / ||b|| in fp64."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a))


# ---------------------------------------------------------------- problem spec

def spec_of(model, x, labels):
    """Plain-data description of a WaveRNN problem for the CPU legs."""
    cfg = model.loop_config()
    g = model.cell.geom
    sp = dict(b=g.b.detach().cpu().numpy().astype(np.float64),
              x=x.detach().cpu().numpy().astype(np.float64),
              dt=float(cfg.dt), h=float(cfg.h),
              src=[list(p) for p in cfg.src_px], taps=[list(t) for t in cfg.taps],
              labels=None if labels is None else labels.cpu().numpy().astype(np.int64),
              nl=None)
    if hasattr(g, "blur_kernel"):
        sp["geom"] = dict(c0=float(g.c0), c1=float(g.c1), eta=float(g.eta), beta=float(g.beta),
                          blur_N=int(g.blur_N),
                          kernel=g.blur_kernel.detach().cpu().numpy().astype(np.float64))
    if cfg.nonlinear:
        sp["nl"] = dict(sat_b0=float(cfg.sat_b0), sat_uth=float(cfg.sat_uth),
                        c_nl=float(cfg.c_nl), clamp=cfg.c_clamp)
    return sp


def c_map_vjp(sp, rho64, gc):
    """(dc/drho)^T gc through the reference's rho -> c ops in fp64
    (oracle/ref_torch.freeform_c, torch autograd)."""
    import torch
    from oracle import ref_torch as rt
    gm = sp["geom"]
    rho = torch.from_numpy(rho64.copy()).requires_grad_(True)
    c = rt.freeform_c(rho, gm["c0"], gm["c1"], torch.from_numpy(gm["kernel"]), gm["blur_N"],
                      gm["eta"], gm["beta"])
    c.backward(torch.from_numpy(gc))
    return rho.grad.numpy()


def loss64(sp, y):
    """bench.py's loss in fp64 with torch autograd: (J, dJ/dy)."""
    import torch
    yt = torch.from_numpy(np.asarray(y, np.float64)).requires_grad_(True)
    s = yt.sum(dim=1)
    J = torch.nn.functional.cross_entropy(s / s.sum(dim=1, keepdim=True),
                                          torch.from_numpy(sp["labels"]))
    J.backward()
    return float(J.detach()), yt.grad.numpy()


def make_pool(n):
    """`n` forked workers for the CPU legs; they run tests/parity_cpu.py
    functions only (no torch)."""
    return mp.get_context("fork").Pool(n)


def _chunks(B, n):
    n = max(1, min(n, B))
    e = np.linspace(0, B, n + 1).astype(int)
    return [slice(int(e[i]), int(e[i + 1])) for i in range(n) if e[i + 1] > e[i]]


# ---------------------------------------------------------------- CPU legs

def grad64_chain(sp, rho64, pool, nchunk):
    """fp64 forward, loss and adjoint of the whole batch, chained through
    the fp64 rho -> c map.  Returns (y64, J64, dJ/drho)."""
    B = sp["x"].shape[0]
    parts = [pool.apply_async(pc.task_fwd64, (sp, rho64, s)) for s in _chunks(B, nchunk)]
    y64 = np.concatenate([p.get() for p in parts], axis=0)
    J64, ybar = loss64(sp, y64)
    parts = [pool.apply_async(pc.task_adj64, (sp, rho64, ybar, s)) for s in _chunks(B, nchunk)]
    res = [p.get() for p in parts]
    gc = sum(r[0] for r in res)
    grho_direct = sum(r[1] for r in res)
    return y64, J64, c_map_vjp(sp, rho64, gc) + grho_direct


def fd_pixels(sp, seed=0, n=FD_PIXELS):
    """`n` distinct random pixels of the design region (b == 0)."""
    ii, jj = np.nonzero(sp["b"] == 0)
    pick = np.random.default_rng(seed).choice(len(ii), size=n, replace=False)
    return [(int(ii[k]), int(jj[k])) for k in pick]


def fd_submit(sp, rho64, pixels, pool, delta=FD_DELTA):
    jobs = []
    for (i, j) in pixels:
        for s in (1.0, -1.0):
            r = rho64.copy()
            r[i, j] += s * delta
            jobs.append(pool.apply_async(pc.task_J64, (sp, r)))
    return jobs


def fd_collect(jobs, delta=FD_DELTA):
    J = [j.get() for j in jobs]
    return np.array([(J[2 * k] - J[2 * k + 1]) / (2 * delta) for k in range(len(J) // 2)])


def ref32_grad(model, x, labels, T=None):
    """The reference's fp32 op sequence with torch autograd (oracle/ref_torch):
    probe series and dJ/drho."""
    import torch
    from oracle import ref_torch as rt
    g, cell = model.cell.geom, model.cell
    rho = g.rho.detach().cpu().clone().requires_grad_(True)
    c = rt.freeform_c(rho, g.c0.item(), g.c1.item(), g.blur_kernel.cpu(), int(g.blur_N.item()),
                      g.eta.item(), g.beta.item())
    src = [(s.x.cpu(), s.y.cpu()) for s in model.sources]
    probes = [(p.x.cpu(), p.y.cpu(), bool(p.intensity)) for p in model.probes]
    xs = x.cpu() if T is None else x.cpu()[:, :T]
    y = rt.rnn_forward(xs, c=c, rho=rho, b_geom=g.b.cpu(), dt=cell.dt.cpu(), h=g.h.cpu(),
                       sources=src, probes=probes, satdamp_b0=cell.satdamp_b0.cpu(),
                       satdamp_uth=cell.satdamp_uth.cpu(), c_nl=cell.c_nl.cpu(),
                       c_clamp=cell.c_clamp)
    s = y.sum(dim=1)
    J = torch.nn.functional.cross_entropy(rt.normalize_power(s), labels.cpu())
    J.backward()
    return y.detach().numpy(), rho.grad.numpy()


# ---------------------------------------------------------------- GPU legs

def _cached(workload):
    from wavetorch_amd import autotune
    return autotune.cached(workload)


def gpu_training_grad(model, x, labels, dev, tuning):
    """One forward + loss + backward through the module API on the GPU:
    (y, dJ/drho, c, J)."""
    import torch
    from wavetorch_amd.utils import normalize_power
    model = model.to(dev)
    if tuning is not None:
        model.tuning = tuning
    model.zero_grad(set_to_none=True)
    y = model(x.to(dev))
    J = torch.nn.functional.cross_entropy(normalize_power(y.sum(dim=1)), labels.to(dev))
    J.backward()
    with torch.no_grad():
        c = model.cell.geom.c.detach().cpu().numpy()
    torch.cuda.synchronize(dev)
    return (y.detach().cpu().numpy(), model.cell.geom.rho.grad.detach().cpu().numpy(), c,
            float(J))


# ---------------------------------------------------------------- workloads

def report_training(name, pool, dev, nchunk, T=None):
    from wavetorch_amd import workloads
    kw = {} if T is None else {"T": T}
    model, x, labels = getattr(workloads, name)(**kw)
    sp = spec_of(model, x, labels)
    B = x.shape[0]
    rho64 = model.cell.geom.rho.detach().cpu().numpy().astype(np.float64)
    pixels = fd_pixels(sp)
    t0 = time.time()
    log(f"[{name}] central differences on {len(pixels)} pixels, fp64 chain: submitted")
    fd_jobs = fd_submit(sp, rho64, pixels, pool)
    log(f"[{name}] GPU forward + backward")
    yg, grho_g, cg, Jg = gpu_training_grad(model, x, labels, dev, _cached(name) if T is None
                                           else None)
    log(f"[{name}] fp32 oracle given the GPU's c")
    rho32 = rho64.astype(np.float32)
    parts = [pool.apply_async(pc.task_fwd32, (sp, cg, rho32, s)) for s in _chunks(B, nchunk)]
    y32 = np.concatenate([p.get() for p in parts], axis=0)
    log(f"[{name}] fp64 oracle forward + adjoint")
    y64, J64, grho64 = grad64_chain(sp, rho64, pool, nchunk)
    log(f"[{name}] waiting for central differences ({time.time() - t0:.0f} s so far)")
    fd = fd_collect(fd_jobs)
    log(f"[{name}] reference fp32 autograd")
    import torch
    torch.set_num_threads(nchunk)
    yr, grho_r = ref32_grad(model.cpu(), x, labels)
    gpx = np.array([grho_g[i, j] for (i, j) in pixels])
    opx = np.array([grho64[i, j] for (i, j) in pixels])
    big = np.abs(fd) > 1e-3 * np.abs(fd).max()
    return {
        "workload": name, "B": int(B), "T": int(x.shape[1]),
        "Nx": int(sp["b"].shape[0]), "Ny": int(sp["b"].shape[1]),
        "forward": {"bitwise_vs_oracle_fp32_given_gpu_c": bool(np.array_equal(yg, y32)),
                    "rel_l2_vs_reference_fp32": rel(yg, yr),
                    "rel_l2_vs_oracle_fp64": rel(yg, y64),
                    "rel_l2_reference_fp32_vs_oracle_fp64": rel(yr, y64)},
        "loss": {"gpu": Jg, "oracle_fp64": J64},
        "grad_rho": {"rel_vs_oracle_fp64": rel(grho_g, grho64),
                     "rel_vs_reference_fp32": rel(grho_g, grho_r),
                     "rel_reference_fp32_vs_oracle_fp64": rel(grho_r, grho64)},
        "central_differences": {
            "pixels": len(pixels), "delta": FD_DELTA,
            "rel_gpu_vs_fd": rel(gpx, fd), "rel_oracle_fp64_vs_fd": rel(opx, fd),
            "max_pixel_rel_gpu_vs_fd": float(np.max(np.abs(gpx - fd)[big] / np.abs(fd)[big])),
            "pixels_above_1e-3_of_max": int(big.sum())},
        "seconds": round(time.time() - t0, 1),
    }


def conditioning(name, pool, dev, nchunk, lengths):
    """dJ/drho of the GPU and of the reference's fp32 op sequence against the
    fp64 oracle chain at each length."""
    from wavetorch_amd import workloads
    out = []
    for T in lengths:
        model, x, labels = getattr(workloads, name)(T=T)
        sp = spec_of(model, x, labels)
        rho64 = model.cell.geom.rho.detach().cpu().numpy().astype(np.float64)
        _, grho_g, _, _ = gpu_training_grad(model, x, labels, dev, None)
        _, _, grho64 = grad64_chain(sp, rho64, pool, nchunk)
        _, grho_r = ref32_grad(model.cpu(), x, labels)
        out.append({"T": T, "grad_rho_rel_gpu_vs_oracle_fp64": rel(grho_g, grho64),
                    "grad_rho_rel_reference_fp32_vs_oracle_fp64": rel(grho_r, grho64)})
        log(f"[{name}] conditioning T={T}: {out[-1]}")
    return out


def memory_vs_K(name, dev, Ks=(16, 64, 256), steps=2):
    """Workspace bytes (the plan's), peak device memory and step time of the
    full training step for each checkpoint interval."""
    import torch
    import bench
    from wavetorch_amd._loop import resolve
    out = []
    base = _cached(name) or {}
    for K in Ks:
        tun = {**base, "ck_every": K}
        model, x, labels = bench.build(dev, tun, name)
        cfg = model.loop_config()
        _, saved, temp = resolve(cfg, x.shape[0], x.shape[1], model.cell.geom.cmax, tun)
        opt = torch.optim.Adam(model.parameters(), lr=1e-3)
        bench.train_step(model, x, labels, opt)
        torch.cuda.synchronize(dev)
        torch.cuda.reset_peak_memory_stats(dev)
        t = time.perf_counter()
        for _ in range(steps):
            bench.train_step(model, x, labels, opt)
        torch.cuda.synchronize(dev)
        ms = (time.perf_counter() - t) / steps * 1e3
        out.append({"K": K, "saved_bytes": int(saved), "temp_bytes": int(temp),
                    "peak_allocated_bytes": int(torch.cuda.max_memory_allocated(dev)),
                    "ms_per_step": round(ms, 2)})
        log(f"[{name}] K={K}: saved {saved / 2**20:.0f} MiB temp {temp / 2**20:.0f} MiB "
            f"{ms:.1f} ms")
        del model, x, labels, opt
        torch.cuda.empty_cache()
    return out


def report_scatter(pool, dev):
    import torch
    import bench
    from wavetorch_amd import workloads
    name = "scatter_1024"
    stride = bench.WORKLOADS[name]["field_stride"]
    model, x, _ = workloads.scatter_1024()
    sp = spec_of(model, x, None)
    rho64 = model.cell.geom.rho.detach().cpu().numpy().astype(np.float64)
    c64 = pc.c_map(sp, rho64)
    t0 = time.time()
    log(f"[{name}] fp64 oracle fields: submitted")
    j64 = pool.apply_async(pc.task_fields64, (sp, c64, stride))
    model, x, _ = bench.build(dev, _cached(name), name)
    with torch.no_grad():
        f = model(x, output_fields=True).detach().cpu().numpy()
        cg = model.cell.geom.c.detach().cpu().numpy()
    log(f"[{name}] fp32 oracle fields given the GPU's c")
    f32 = pool.apply_async(pc.task_fwd32, (sp, cg, None, slice(None), stride)).get()
    f64 = j64.get()
    return {"workload": name, "B": 1, "T": int(x.shape[1]), "Nx": 1024, "Ny": 1024,
            "snapshots": int(f.shape[1]),
            "forward": {"bitwise_vs_oracle_fp32_given_gpu_c": bool(np.array_equal(f, f32)),
                        "rel_l2_vs_oracle_fp64": rel(f, f64),
                        "rel_l2_last_snapshot_vs_oracle_fp64": rel(f[:, -1], f64[:, -1])},
            "seconds": round(time.time() - t0, 1)}


def report_lens(pool, dev):
    import torch
    from wavetorch_amd import workloads
    from wavetorch_amd.batch import Instance, InstanceBatch
    name = "lens_batch_256"
    items = workloads.lens_batch_256()
    t0 = time.time()
    models = [m.to(dev) for m, _ in items]
    xs = [x.to(dev) for _, x in items]
    with torch.no_grad():
        cs = [m.cell.geom.c.detach().clone() for m in models]
    ybars = [-torch.ones((x.shape[0], 1), device=dev) for x in xs]
    insts = [Instance.from_model(m, x.shape[0]) for m, x in zip(models, xs)]
    bt = InstanceBatch(insts, dev, tuning=_cached_batch(insts))
    bt.set_inputs(cs, xs, ybars)
    log(f"[{name}] GPU forward + adjoint of {len(insts)} instances")
    bt.fwd_adj()
    torch.cuda.synchronize(dev)
    outs = [(y.cpu().numpy(), gc.cpu().numpy(), gx.cpu().numpy()) for y, gc, gx in bt.outputs()]
    jobs = []
    for (m, x), c in zip(items, cs):
        sp = spec_of(m.cpu(), x[None], None)
        jobs.append(pool.apply_async(pc.task_lens, (sp, c.cpu().numpy(),
                                                  -np.ones((1, x.shape[0], 1)))))
    log(f"[{name}] oracle per instance: submitted")
    bit, fwd, egc, egx, gold = [], [], [], [], []
    for k, ((y, gc, gx), j) in enumerate(zip(outs, jobs)):
        y32, y64, gc64, gx64 = j.get()
        gold.append((gc64, gx64))
        bit.append(bool(np.array_equal(y, y32[0])))
        fwd.append(rel(y, y64[0]))
        egc.append(rel(gc, gc64))
        egx.append(rel(gx, gx64[0]))
        if (k + 1) % 64 == 0:
            log(f"[{name}] {k + 1} instances checked")
    st = lambda v: {"max": float(np.max(v)), "median": float(np.median(v))}  # noqa: E731
    worst = sorted(set(np.argsort(egc)[-3:].tolist() + np.argsort(egx)[-3:].tolist()))
    log(f"[{name}] reference fp32 autograd on instances {worst}")
    import torch
    torch.set_num_threads(16)
    ref = []
    for k in worst:
        m, x = items[k]
        gc64, gx64 = gold[k]
        gcr, gxr = ref32_instance(m.cpu(), x, cs[k].cpu())
        ref.append({"instance": int(k), "N": int(m.cell.geom.b.shape[0]), "T": int(x.shape[0]),
                    "grad_c_rel_gpu_vs_oracle_fp64": egc[k],
                    "grad_c_rel_reference_fp32_vs_oracle_fp64": rel(gcr, gc64),
                    "grad_x_rel_gpu_vs_oracle_fp64": egx[k],
                    "grad_x_rel_reference_fp32_vs_oracle_fp64": rel(gxr, gx64)})
    return {"workload": name, "instances": len(outs),
            "forward": {"bitwise_vs_oracle_fp32_given_gpu_c": int(sum(bit)),
                        "rel_l2_vs_oracle_fp64": st(fwd)},
            "grad_c_rel_vs_oracle_fp64": st(egc), "grad_x_rel_vs_oracle_fp64": st(egx),
            "largest_errors_vs_reference_fp32": ref,
            "seconds": round(time.time() - t0, 1)}


def ref32_instance(model, x, c):
    """The reference's fp32 op sequence for one batched instance given its
    c, with J = -(sum of the focal intensity) as in the batch's dJ/dy = -1:
    (dJ/dc, dJ/dx)."""
    import torch
    from oracle import ref_torch as rt
    g = model.cell.geom
    ct = c.clone().requires_grad_(True)
    xt = x.clone()[None].requires_grad_(True)
    src = [(s.x.cpu(), s.y.cpu()) for s in model.sources]
    probes = [(p.x.cpu(), p.y.cpu(), bool(p.intensity)) for p in model.probes]
    y = rt.rnn_forward(xt, c=ct, rho=g.rho.detach(), b_geom=g.b, dt=model.cell.dt, h=g.h,
                       sources=src, probes=probes)
    (-y.sum()).backward()
    return ct.grad.numpy(), xt.grad.numpy()


def _cached_batch(insts):
    from wavetorch_amd import autotune
    return autotune.cached_batch(insts)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("workloads", nargs="*", default=list(WORKLOADS))
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    for w in args.workloads:
        if w not in WORKLOADS:
            raise SystemExit(f"unknown workload {w}; choose from {', '.join(WORKLOADS)}")
    # the workers are forked before this process imports torch
    assert "torch" not in sys.modules
    pool = make_pool(args.workers)
    import torch
    dev = torch.device("cuda:0")
    try:
        for w in args.workloads:
            if w == "scatter_1024":
                r = report_scatter(pool, dev)
            elif w == "lens_batch_256":
                r = report_lens(pool, dev)
            elif w == "nonlinear_150_b32":
                r = report_training(w, pool, dev, args.workers, T=NL_STEPS)
                r["note"] = f"gradients over the first {NL_STEPS} of 3000 steps"
                r["conditioning"] = conditioning(w, pool, dev, args.workers, NL_CONDITIONING)
                r["memory_vs_K"] = memory_vs_K(w, dev)
            else:
                r = report_training(w, pool, dev, args.workers)
            line = json.dumps(r)
            print(line, flush=True)
            if args.out:
                with open(args.out, "a") as f:
                    f.write(line + "\n")
    finally:
        pool.terminate()
        pool.join()


if __name__ == "__main__":
    main()
