"""Autotuner for the temporally blocked kernels.

``tune`` searches tilings per sample x time-block depth x rows per thread (for
the forward and the adjoint separately) for one problem shape by timing whole
training steps on the GPU, starting from the library's own plan; ``tune_batch``
does the same for batched instances.  Winners go to an on-disk cache keyed by
(Nx, Ny, B, T, K, gfx arch) or by the batch's composition; ``WaveRNN.tuning``
takes the winning ``wt_tuning`` fields, and without a cached entry the
library's heuristic applies.

The cache file is ``wavetorch_amd/tuning_cache.json`` (JSON, written
atomically; a corrupt file is ignored and rewritten).
"""
from __future__ import annotations

import json
import math
import os
import threading
import time

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(_HERE, "tuning_cache.json")
_lock = threading.Lock()

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


def key(Nx, Ny, B, T, K, arch=None, forward_only=False):
    """Cache key of a problem shape; K is the checkpoint interval, or "open"
    for tunings whose search chose it (they carry ``ck_every``)."""
    return f"{arch or _arch()}:{Nx}x{Ny}:B{B}:T{T}:K{K}" + (":fwd" if forward_only else "")


def lookup(Nx, Ny, B, T, K=64, arch="gfx950", forward_only=False):
    with _lock:
        e = _load().get(key(Nx, Ny, B, T, K, arch, forward_only))
    return dict(e["tuning"]) if e else None


def cached(workload: str):
    """Tuning for a named workload of wavetorch_amd.workloads, or None."""
    shapes = {"train_150_b64": (150, 150, 64, 3000, 64),
              "propagate_100_b4": (100, 100, 4, 1000, "open"),
              "nonlinear_150_b32": (150, 150, 32, 3000, 64),
              "scatter_1024": (1024, 1024, 1, 4000, 64, "gfx950", True)}
    if workload not in shapes:
        return None
    return lookup(*shapes[workload])


# Options of each tuning field the search moves through (column-segment
# family; rows per thread `cpt` must be a compiled variant).  "tiling" and
# "a_tiling" are (strips, tcols) pairs: s x c tiles per sample for the
# forward and the adjoint; their options are built per problem (tilings).
AXES = {
    "tiling": (),
    "tblock": (4, 6, 8, 12, 16, 20, 24, 32),
    "cpt": (4, 6, 8, 10, 12),
    "a_tiling": (),
    "a_tblock": (4, 5, 6, 8, 10, 12, 13, 16, 20, 24),
    "a_cpt": (4, 6, 8, 10),
}


def tilings(Nx, Ny, s0, c0):
    """(strips, tcols) options around the plan's own s0 x c0: every pair with
    s, c in 1..6 or within 0.5x-1.5x of s0, c0 whose tiles are at most 2:1
    and whose tile count is within 4x of s0*c0 either way."""
    def near(v0):
        return set(range(1, 7)) | {max(1, round(v0 * r)) for r in
                                   (0.5, 0.625, 0.75, 0.875, 1.125, 1.25, 1.5)} | {v0}
    out = []
    for s in sorted(v for v in near(s0) if v <= Nx):
        for c in sorted(v for v in near(c0) if v <= Ny):
            if (s, c) != (s0, c0):
                if not s0 * c0 / 4 <= s * c <= s0 * c0 * 4:
                    continue
                if max((Nx / s) / (Ny / c), (Ny / c) / (Nx / s)) > 2.0 + 1e-9 and s * c > 1:
                    continue
            out.append((s, c))
    return tuple(out)


def _to_tuning(st):
    t = {"kind": 0, "strips": st["tiling"][0], "tcols": st["tiling"][1],
         "tblock": st["tblock"], "nt": 1024, "cpt": st["cpt"],
         "a_strips": st["a_tiling"][0], "a_tcols": st["a_tiling"][1],
         "a_tblock": st["a_tblock"], "a_nt": 1024, "a_cpt": st["a_cpt"]}
    if "ck_every" in st:
        t["ck_every"] = st["ck_every"]
    return t


def _from_tuning(t, with_k=False):
    st = {"tiling": (t["strips"], t["tcols"]), "tblock": t["tblock"], "cpt": t["cpt"],
          "a_tiling": (t["a_strips"], t["a_tcols"]), "a_tblock": t["a_tblock"],
          "a_cpt": t["a_cpt"]}
    if with_k:
        st["ck_every"] = t["ck_every"]
    return st


def k_options(K0, T):
    """Checkpoint intervals the search tries from the model's K0: doublings
    up to T, and T itself (one segment: the forward then writes the only
    record and the reverse sweep recomputes nothing)."""
    return tuple(sorted({k for k in (K0, 2 * K0, 4 * K0, 8 * K0, T) if K0 <= k <= T}))


def tune(model, x, labels=None, steps=2, verbose=False, budget_s=300.0, rounds=2,
         forward_only=False, reps=2, tune_k=False):
    """Coordinate search over the tuning fields of AXES for ``model`` on
    batch ``x``, timing a full training step (forward, loss, checkpointed
    adjoint) on the GPU, or with `forward_only` the forward without gradients
    (fields every ``model.field_stride`` steps when the model has no probes).
    Each candidate's time is the best of `reps` runs of `steps` steps (a
    single run can carry a host-side hiccup as large as the differences
    between good plans).  Starts from the library's own choice, tries every
    option of one field at a time (the rest fixed), keeps any improvement and
    repeats for `rounds` passes or until `budget_s`; caches the winner under
    key(Nx, Ny, B, T, K) and returns it.  With `tune_k` the checkpoint
    interval is searched too, from the model's K (k_options; the winner
    carries ``ck_every`` and is cached under K = "open") -- for
    configurations that leave K open: more memory for fewer recomputed
    segments.  A checkpoint interval set on the model is set aside during
    that search and restored after it (it then still overrides the
    tuning's)."""
    from .utils import normalize_power
    B, T = x.shape
    Nx, Ny = model.cell.geom.domain_shape
    K = int(model.checkpoint_every)
    user_k = model._checkpoint_every
    if tune_k:
        model.checkpoint_every = None
    t_start = time.perf_counter()

    def step():
        if forward_only:
            with torch.no_grad():
                model(x)
            return
        model.zero_grad(set_to_none=True)
        y = model(x)
        if labels is None:
            loss = -y.sum()
        else:
            loss = torch.nn.functional.cross_entropy(normalize_power(y.sum(dim=1)), labels)
        loss.backward()

    def measure(tun):
        model.tuning = dict(tun)
        try:
            plan = model.plan(B, T, x.device)
            step()
            t = float("inf")
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    step()
                torch.cuda.synchronize()
                t = min(t, (time.perf_counter() - t0) / steps)
        except (ValueError, RuntimeError) as e:  # a tiling the variants cannot hold
            if verbose:
                print("skip", tun, str(e)[:120], flush=True)
            return None, float("inf")
        if verbose:
            print(f"{ {k: v for k, v in plan.tuning.items() if v} } {t * 1e3:.2f} ms", flush=True)
        return plan.tuning, t

    model.tuning = None
    base, best_t = measure({"ck_every": K} if tune_k else {})
    if base is None:
        model.checkpoint_every = user_k
        return None
    state = _from_tuning(base, with_k=tune_k)
    tried = {tuple(sorted(state.items()))}
    def consider(cand):
        nonlocal state, best_t
        key_ = tuple(sorted(cand.items()))
        if key_ in tried or time.perf_counter() - t_start > budget_s:
            return False
        tried.add(key_)
        _, t = measure(_to_tuning(cand))
        if t < best_t * 0.995:
            state, best_t = cand, t
            return True
        return False

    # tilings around the library's own (large grids start far above the
    # small counts)
    axes = dict(AXES)
    for f in ("tiling", "a_tiling"):
        axes[f] = tilings(Nx, Ny, *state[f])
    if forward_only:  # the adjoint's fields do not take part
        axes = {f: o for f, o in axes.items() if not f.startswith("a_")}
    if tune_k and not forward_only:
        axes["ck_every"] = k_options(K, T)
    for _ in range(rounds):
        improved = False
        for field, options in axes.items():
            for v in options:
                improved |= consider(dict(state, **{field: v}))
        # block depth and rows per thread are coupled (deeper blocks need more
        # rows per thread to fit): also move them together
        pairs = (("tblock", "cpt"),) if forward_only else (("tblock", "cpt"),
                                                           ("a_tblock", "a_cpt"))
        for f1, f2 in pairs:
            base_state = dict(state)
            for v1 in axes[f1]:
                for v2 in axes[f2]:
                    improved |= consider(dict(base_state, **{f1: v1, f2: v2}))
        if not improved:
            break
    best = _to_tuning(state)
    model.tuning = best
    model.checkpoint_every = user_k
    with _lock:
        d = _load()
        d[key(Nx, Ny, B, T, "open" if tune_k else K, forward_only=forward_only)] = {
            "tuning": best, "ms": best_t * 1e3}
        _store(d)
    return best


def tune_native(model, x, with_grad=True, reps=2, budget_s=120.0, store=True):
    """The library's own search (``wt_autotune``, the C ABI's tuner: the same
    fields, options and acceptance rule as :func:`tune`, timed with HIP events
    on the loop kernels alone -- forward, and with ``with_grad`` the
    checkpointed adjoint -- rather than on a whole training step) for
    ``model`` on batch ``x`` at the model's checkpoint interval.  Returns the
    winning tuning dict (also set as ``model.tuning``) and its time in ms;
    with ``store`` the winner is cached under key(Nx, Ny, B, T, K) like
    :func:`tune`'s."""
    import ctypes
    from . import _lib
    B, T = x.shape
    cfg = model.loop_config()
    geom = model.cell.geom
    with torch.no_grad():
        c = geom.c.detach().contiguous()
        b = geom.b.detach().contiguous()
        rho = geom.rho.detach().contiguous() if cfg.nonlinear else None
    xx = x.detach().contiguous()
    K = int(model.checkpoint_every)
    start = _lib.WtTuning(ck_every=K)
    out = _lib.WtTuning()
    ms = ctypes.c_double(0.0)
    cc, keep = cfg.to_c()
    with torch.cuda.device(x.device):
        _lib.check(_lib.lib().wt_autotune(
            ctypes.byref(cc), B, T, float(geom.cmax), ctypes.byref(start), int(bool(with_grad)),
            b.data_ptr(), c.data_ptr(), rho.data_ptr() if rho is not None else None,
            xx.data_ptr(), int(reps), float(budget_s * 1e3), ctypes.byref(out),
            ctypes.byref(ms), _lib.stream_ptr(x.device)), "wt_autotune")
    del keep
    best = {k: v for k, v in out.as_dict().items() if v}
    best.pop("ck_every", None)  # the model's K stays the model's
    model.tuning = dict(best)
    if store:
        with _lock:
            d = _load()
            d[key(cfg.Nx, cfg.Ny, B, T, K, forward_only=not with_grad)] = {
                "tuning": best, "ms": ms.value, "search": "wt_autotune"}
            _store(d)
    return best, ms.value


# ------------------------------------------------------------ batched instances

def _aligned_k(tb, atb, kmin=64):
    """The smallest checkpoint interval >= kmin that both block depths
    divide: no segment then ends in a shorter launch."""
    step = tb * atb // math.gcd(tb, atb)
    return -(-kmin // step) * step


# Batched instances: block variant and depth of forward and adjoint, and the
# checkpoint interval (64, or the smallest multiple of both depths from 64 up;
# the lens batch's configuration leaves K open).
BATCH_CANDIDATES = [
    {"tblock": tb, "nt": 1024, "cpt": r, "a_tblock": atb, "a_nt": 1024, "a_cpt": ar,
     "ck_every": k}
    for tb, r in ((16, 10), (24, 10), (24, 12), (32, 12))
    for atb, ar in ((13, 8), (16, 8), (16, 10), (20, 10), (24, 10))
    for k in sorted({64, _aligned_k(tb, atb)})
    if k <= 256
]


def batch_record_bytes(insts, K):
    """Device bytes of the saved record of one segment of K steps (lap u of
    every cell and step, rows of Ny + 1) over the batch's instances, at K
    step slots per instance (an upper bound: a one-segment record is packed
    by each instance's own length)."""
    return sum(4 * K * i.cfg.Nx * (i.cfg.Ny + 1) for i in insts)


def batch_candidates(insts, max_record_bytes):
    """BATCH_CANDIDATES plus, for each block pair, one segment over the whole
    batch (K = the smallest multiple of both depths >= the longest length):
    the forward then records every step and the reverse sweep recomputes
    nothing, for a record of batch_record_bytes(insts, K) -- offered only
    within `max_record_bytes`."""
    T = max(i.length for i in insts)
    out = list(BATCH_CANDIDATES)
    pairs = sorted({(c["tblock"], c["cpt"], c["a_tblock"], c["a_cpt"]) for c in BATCH_CANDIDATES})
    for tb, r, atb, ar in pairs:
        k = _aligned_k(tb, atb, kmin=T)
        if batch_record_bytes(insts, k) <= max_record_bytes:
            out.append({"tblock": tb, "nt": 1024, "cpt": r, "a_tblock": atb, "a_nt": 1024,
                        "a_cpt": ar, "ck_every": k})
    return out


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


def tune_batch(insts, cs, xs, ybars, device, candidates=None, reps=2, verbose=False,
               mem_fraction=0.5):
    """Times the forward + adjoint of the whole batch (wt_batch_fwd_adj) for
    each candidate block variant / time-block depth / checkpoint interval (the
    per-group tiling follows from each: the largest square tiles that fit) and
    caches the fastest under batch_key(insts).  Without `candidates`:
    batch_candidates with a record of at most `mem_fraction` of the device's
    memory."""
    from .batch import InstanceBatch
    if candidates is None:
        total = torch.cuda.get_device_properties(torch.device(device)).total_memory
        candidates = batch_candidates(insts, mem_fraction * total)
    best, best_t = None, float("inf")
    for cand in candidates:
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
        torch.cuda.empty_cache()
    if best is not None:
        with _lock:
            d = _load()
            d[batch_key(insts)] = {"tuning": best, "ms": best_t * 1e3}
            _store(d)
    return best
