"""Autotuner for the temporally blocked kernels.

Searches row strips per sample x time-block depth x (threads, cells per
thread) for one problem shape by timing the forward and the adjoint on the
GPU, and stores the winner in an on-disk cache keyed by
(Nx, Ny, B, T, K, gfx arch).  ``WaveRNN.tuning`` takes the winning
``wt_tuning`` fields; without a cached entry the library's heuristic applies.

The cache file is ``wavetorch_amd/tuning_cache.json`` (JSON, written
atomically; a corrupt file is ignored and rewritten).
"""
from __future__ import annotations

import itertools
import json
import os
import threading
import time

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(_HERE, "tuning_cache.json")
_lock = threading.Lock()

VARIANTS = [(1024, 8), (1024, 12), (1024, 16), (512, 16), (512, 24), (512, 32), (256, 32),
            (256, 48)]


def _arch():
    try:
        return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    except Exception:
        return "gfx950"


def _load():
    try:
        with open(CACHE) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(d):
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def key(Nx, Ny, B, T, K, arch=None):
    return f"{arch or _arch()}:{Nx}x{Ny}:B{B}:T{T}:K{K}"


def lookup(Nx, Ny, B, T, K=64, arch="gfx950"):
    with _lock:
        e = _load().get(key(Nx, Ny, B, T, K, arch))
    return dict(e["tuning"]) if e else None


def cached(workload: str):
    """Tuning for a named workload of wavetorch_amd.workloads, or None."""
    shapes = {"train_150_b64": (150, 150, 64, 3000, 64),
              "propagate_100_b4": (100, 100, 4, 1000, 64),
              "nonlinear_150_b32": (150, 150, 32, 3000, 64)}
    if workload not in shapes:
        return None
    return lookup(*shapes[workload])


def candidates(Nx, B, max_strips=8):
    out = []
    for strips, tb, (nt, cpt) in itertools.product(range(1, max_strips + 1), (4, 8, 12, 16, 24, 32),
                                                   VARIANTS):
        if strips == 1 and tb != 32:
            continue  # whole-domain regions have no halo: depth only sets the launch count
        out.append({"strips": strips, "tblock": tb, "nt": nt, "cpt": cpt})
    return out


def tune(model, x, labels=None, steps=1, verbose=False, max_strips=8, budget_s=120.0):
    """Times fwd+adjoint for every feasible candidate on ``model``/``x`` (on
    the GPU) and caches the fastest; returns it."""
    from ._loop import Plan  # noqa: F401  (plan creation validates feasibility)
    from .utils import normalize_power
    B, T = x.shape
    Nx, Ny = model.cell.geom.domain_shape
    K = int(model.checkpoint_every)
    best, best_t = None, float("inf")
    t_start = time.perf_counter()
    for cand in candidates(Nx, B, max_strips):
        if time.perf_counter() - t_start > budget_s:
            break
        model.tuning = dict(cand)
        try:
            model.plan(B, T, x.device)
        except (ValueError, RuntimeError):
            continue
        try:
            def step():
                model.zero_grad(set_to_none=True)
                y = model(x)
                if labels is None:
                    loss = -y.sum()
                else:
                    loss = torch.nn.functional.cross_entropy(normalize_power(y.sum(dim=1)),
                                                             labels)
                loss.backward()
            step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                step()
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / steps
        except RuntimeError as e:  # an infeasible variant reported by the library
            if verbose:
                print("skip", cand, e)
            continue
        if verbose:
            print(f"{cand} {t * 1e3:.2f} ms", flush=True)
        if t < best_t:
            best, best_t = dict(cand), t
    model.tuning = best
    if best is not None:
        with _lock:
            d = _load()
            d[key(Nx, Ny, B, T, K)] = {"tuning": best, "ms": best_t * 1e3}
            _store(d)
    return best


# ------------------------------------------------------------ batched instances

BATCH_CANDIDATES = [
    {"tblock": 16, "nt": 1024, "cpt": 10, "a_tblock": 12, "a_nt": 1024, "a_cpt": 8},
    {"tblock": 12, "nt": 1024, "cpt": 8, "a_tblock": 12, "a_nt": 1024, "a_cpt": 8},
    {"tblock": 24, "nt": 1024, "cpt": 12, "a_tblock": 12, "a_nt": 1024, "a_cpt": 8},
    {"tblock": 16, "nt": 1024, "cpt": 10, "a_tblock": 8, "a_nt": 1024, "a_cpt": 6},
    {"tblock": 8, "nt": 1024, "cpt": 8, "a_tblock": 8, "a_nt": 1024, "a_cpt": 6},
]


def batch_key(insts, arch=None):
    """Cache key of a batch: the sorted (Nx, Ny, length) of its instances."""
    import hashlib
    desc = sorted((i.cfg.Nx, i.cfg.Ny, i.length) for i in insts)
    return f"{arch or _arch()}:batch:{len(insts)}:" + \
        hashlib.sha1(repr(desc).encode()).hexdigest()[:16]


def cached_batch(insts, arch=None):
    with _lock:
        e = _load().get(batch_key(insts, arch))
    return dict(e["tuning"]) if e else None


def tune_batch(insts, cs, xs, ybars, device, candidates=None, reps=2, verbose=False):
    """Times the forward + adjoint of the whole batch (wt_batch_fwd_adj) for
    each candidate block variant / time-block depth (the per-group tiling
    follows from each: the largest square tiles that fit) and caches the
    fastest under batch_key(insts)."""
    from .batch import InstanceBatch
    best, best_t = None, float("inf")
    for cand in candidates or BATCH_CANDIDATES:
        try:
            bt = InstanceBatch(insts, device, tuning=cand)
        except (ValueError, RuntimeError) as e:
            if verbose:
                print("skip", cand, e, flush=True)
            continue
        bt.set_inputs(cs, xs, ybars)
        bt.fwd_adj()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            bt.fwd_adj()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        if verbose:
            print(f"{cand} {t * 1e3:.2f} ms", flush=True)
        if t < best_t:
            best, best_t = dict(cand), t
        del bt
    if best is not None:
        with _lock:
            d = _load()
            d[batch_key(insts)] = {"tuning": best, "ms": best_t * 1e3}
            _store(d)
    return best
