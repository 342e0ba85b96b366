"""ctypes binding of libwavetorch_hip.so (the C ABI in include/wavetorch_hip.h).

The library is built in-tree by ``wavetorch_amd.csrc.build`` into
``wavetorch_amd/_lib/libwavetorch_hip.so``.  There is no fallback: if the
library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WAVETORCH_AMD_LIB",
                          os.path.join(_HERE, "_lib", "libwavetorch_hip.so"))

WT_OK = 0
WT_EINVAL = -1
WT_ECFL = -2
WT_EHIP = -3
WT_EUNSUPPORTED = -4
WT_ENOMEM = -5
ABI_VERSION = 2

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_f32 = ctypes.c_float
c_f64 = ctypes.c_double
c_p = ctypes.c_void_p


class WtConfig(ctypes.Structure):
    """Mirror of ``wt_config``."""
    _fields_ = [
        ("Nx", c_i32), ("Ny", c_i32), ("dt", c_f32), ("h", c_f32),
        ("nsrc_px", c_i32), ("src_px", ctypes.POINTER(c_i32)),
        ("ntap", c_i32), ("taps", ctypes.POINTER(c_i32)),
        ("sat_b0", c_f32), ("sat_uth", c_f32), ("c_nl", c_f32),
        ("c_clamp", c_i32), ("c_min", c_f32), ("c_max", c_f32),
    ]


class WtTuning(ctypes.Structure):
    """Mirror of ``wt_tuning``."""
    _fields_ = [("plain", c_i32), ("strips", c_i32), ("tblock", c_i32), ("nt", c_i32),
                ("cpt", c_i32), ("ck_every", c_i32), ("kind", c_i32),
                ("tcols", c_i32), ("a_strips", c_i32), ("a_tcols", c_i32), ("a_tblock", c_i32),
                ("a_nt", c_i32), ("a_cpt", c_i32), ("streams", c_i32)]

    def as_dict(self):
        d = {k: int(getattr(self, k)) for k, _ in self._fields_}
        if d["streams"] == 0:  # single plans: no part streams
            del d["streams"]
        return d


class WtScalarsF64(ctypes.Structure):
    """Mirror of ``wt_scalars_f64`` (the float64 loop's scalars)."""
    _fields_ = [(n, ctypes.c_double) for n in ("dt", "h", "sat_b0", "sat_uth", "c_nl", "c_min",
                                               "c_max")]


class WtBatchIO(ctypes.Structure):
    """Mirror of ``wt_batch_group_io``."""
    _fields_ = [("b", c_p), ("c", c_p), ("x", c_p), ("y", c_p), ("ybar", c_p),
                ("grad_c", c_p), ("grad_x", c_p)]


_lock = threading.Lock()
_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "wt_abi_version": (ctypes.c_int, []),
    "wt_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "wt_check_finite": (ctypes.c_int, [c_p, c_i64, c_p, c_p]),
    "wt_check_finite_f64": (ctypes.c_int, [c_p, c_i64, c_p, c_p]),
    "wt_coef_linear": (ctypes.c_int, [c_p, c_p, c_i64, c_f32, c_p, c_p, c_p, c_p]),
    "wt_laplacian": (ctypes.c_int, [c_p, c_p, c_i32, c_i32, c_i32, c_f32, c_p]),
    "wt_laplacian_f64": (ctypes.c_int, [c_p, c_p, c_i32, c_i32, c_i32, c_f64, c_p]),
    "wt_step_fwd_f64": (ctypes.c_int, [c_p, c_p, c_i32, c_p, c_p, c_i32, c_i32, c_i32, c_f64,
                                       c_f64, c_p, c_p]),
    "wt_step_bwd_f64": (ctypes.c_int, [c_p, c_p, c_i32, c_p, c_p, c_p, c_i32, c_i32, c_i32,
                                       c_f64, c_f64, c_p, c_p, c_p, c_p, c_p]),
    "wt_step_fwd": (ctypes.c_int, [c_p, c_p, c_i32, c_p, c_p, c_i32, c_i32, c_i32, c_f32, c_f32,
                                   c_p, c_p]),
    "wt_step_bwd": (ctypes.c_int, [c_p, c_p, c_i32, c_p, c_p, c_p, c_i32, c_i32, c_i32, c_f32,
                                   c_f32, c_p, c_p, c_p, c_p, c_p]),
    "wt_plan_create": (ctypes.c_int, [ctypes.POINTER(WtConfig), c_i32, c_i32, c_f32,
                                      ctypes.POINTER(WtTuning), ctypes.POINTER(c_p)]),
    "wt_plan_resolve": (ctypes.c_int, [ctypes.POINTER(WtConfig), c_i32, c_i32, c_f32,
                                       ctypes.POINTER(WtTuning), ctypes.POINTER(WtTuning),
                                       ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]),
    "wt_plan_destroy": (ctypes.c_int, [c_p]),
    "wt_plan_tuning": (ctypes.c_int, [c_p, ctypes.POINTER(WtTuning)]),
    "wt_workspace_bytes": (ctypes.c_int, [c_p, c_i32, ctypes.POINTER(c_i64),
                                          ctypes.POINTER(c_i64)]),
    "wt_loop_fwd": (ctypes.c_int, [c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_i32, c_i32, c_p, c_p,
                                   c_p]),
    "wt_loop_adj": (ctypes.c_int, [c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                   c_p, c_p]),
    "wt_geom_freeform_c": (ctypes.c_int, [c_p, c_p, c_i32, c_i32, c_i32, c_i32, c_f32, c_f32,
                                          c_f32, c_f32, c_p, c_p, c_p, c_p, c_p]),
    "wt_geom_freeform_c_bwd": (ctypes.c_int, [c_p, c_p, c_p, c_i32, c_i32, c_i32, c_i32, c_f32,
                                              c_f32, c_f32, c_f32, c_p, c_p, c_p]),
    "wt_geom_holey_c": (ctypes.c_int, [c_p, c_p, c_p, c_i32, c_i32, c_i32, c_f32, c_f32, c_f32,
                                       c_f32, c_p, c_p, c_p]),
    "wt_geom_holey_c_bwd": (ctypes.c_int, [c_p, c_p, c_p, c_p, c_p, c_i32, c_i32, c_i32, c_f32,
                                           c_f32, c_f32, c_f32, c_p, c_p, c_p, c_p, c_p]),
    "wt_geom_holey_workspace_bytes": (ctypes.c_int, [c_i32, c_i32, c_i32,
                                                     ctypes.POINTER(c_i64)]),
    "wt_geom_freeform_c_f64": (ctypes.c_int, [c_p, c_p, c_i32, c_i32, c_i32, c_i32, c_f64,
                                              c_f64, c_f64, c_f64, c_p, c_p, c_p, c_p, c_p]),
    "wt_geom_freeform_c_bwd_f64": (ctypes.c_int, [c_p, c_p, c_p, c_i32, c_i32, c_i32, c_i32,
                                                  c_f64, c_f64, c_f64, c_f64, c_p, c_p, c_p]),
    "wt_geom_holey_c_f64": (ctypes.c_int, [c_p, c_p, c_p, c_i32, c_i32, c_i32, c_f64, c_f64,
                                           c_f64, c_f64, c_p, c_p, c_p]),
    "wt_geom_holey_c_bwd_f64": (ctypes.c_int, [c_p, c_p, c_p, c_p, c_p, c_i32, c_i32, c_i32,
                                               c_f64, c_f64, c_f64, c_f64, c_p, c_p, c_p, c_p,
                                               c_p]),
    "wt_geom_holey_workspace_bytes_f64": (ctypes.c_int, [c_i32, c_i32, c_i32,
                                                         ctypes.POINTER(c_i64)]),
    "wt_loss_power_ce": (ctypes.c_int, [c_p, c_p, c_i32, c_i32, c_i32, c_p, c_p, c_p, c_p]),
    "wt_loss_power_ce_workspace_bytes": (ctypes.c_int, [c_i32, c_i32, c_i32,
                                                        ctypes.POINTER(c_i64)]),
    "wt_loss_power_ce_static": (ctypes.c_int, [c_p, c_p, c_i32, c_i32, c_i32, c_p, c_p, c_p]),
    "wt_workspace_bytes_f64": (ctypes.c_int, [c_p, c_i32, ctypes.POINTER(c_i64),
                                              ctypes.POINTER(c_i64)]),
    "wt_loop_fwd_f64": (ctypes.c_int, [c_p, ctypes.POINTER(WtScalarsF64), c_p, c_p, c_p, c_p, c_p,
                                       c_p, c_i32, c_i32, c_p, c_p, c_p]),
    "wt_loop_adj_f64": (ctypes.c_int, [c_p, ctypes.POINTER(WtScalarsF64), c_p, c_p, c_p, c_p, c_p,
                                       c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "wt_autotune": (ctypes.c_int, [ctypes.POINTER(WtConfig), c_i32, c_i32, c_f32,
                                   ctypes.POINTER(WtTuning), c_i32, c_p, c_p, c_p, c_p, c_i32,
                                   ctypes.c_double, ctypes.POINTER(WtTuning),
                                   ctypes.POINTER(ctypes.c_double), c_p]),
    "wt_batch_plan_create": (ctypes.c_int, [c_i32, ctypes.POINTER(WtConfig),
                                            ctypes.POINTER(c_i32), ctypes.POINTER(c_i32),
                                            ctypes.POINTER(c_f32), ctypes.POINTER(WtTuning),
                                            ctypes.POINTER(c_p)]),
    "wt_batch_plan_tuning": (ctypes.c_int, [c_p, c_i32, ctypes.POINTER(WtTuning)]),
    "wt_batch_workspace_bytes": (ctypes.c_int, [c_p, ctypes.POINTER(c_i64),
                                                ctypes.POINTER(c_i64)]),
    "wt_batch_fwd": (ctypes.c_int, [c_p, ctypes.POINTER(WtBatchIO), c_i32, c_p, c_p, c_p]),
    "wt_batch_adj": (ctypes.c_int, [c_p, ctypes.POINTER(WtBatchIO), c_p, c_p, c_p]),
    "wt_batch_fwd_adj": (ctypes.c_int, [c_p, ctypes.POINTER(WtBatchIO), c_p, c_p, c_p]),
    "wt_batch_plan_destroy": (ctypes.c_int, [c_p]),
    "wt_profile": (ctypes.c_int, [c_i32]),
    "wt_profile_read": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(c_i64),
                                       ctypes.POINTER(c_i64)]),
    "wt_graph_begin": (ctypes.c_int, [c_p]),
    "wt_graph_end": (ctypes.c_int, [c_p, ctypes.POINTER(c_p)]),
    "wt_graph_launch": (ctypes.c_int, [c_p, c_p]),
    "wt_graph_destroy": (ctypes.c_int, [c_p]),
}


class WaveLibError(RuntimeError):
    pass


def lib():
    """Loads the library once; raises if it is absent (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise WaveLibError(
                    f"{LIB_PATH} is missing: build it with `python -m wavetorch_amd.csrc.build` "
                    "(wavetorch_amd has no CPU fallback)")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            v = L.wt_abi_version()
            if v != ABI_VERSION:
                raise WaveLibError(f"libwavetorch_hip ABI {v}, expected {ABI_VERSION}")
            _lib = L
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().wt_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    """Maps a wt_status to the reference's exception types (ValueError for
    CFL / bad arguments, RuntimeError otherwise)."""
    if rc == WT_OK:
        return
    msg = last_error()
    if rc in (WT_EINVAL, WT_ECFL):
        raise ValueError(msg or f"{what} failed ({rc})")
    raise RuntimeError(f"{what} failed ({rc}): {msg}")


def ptr(t) -> int | None:
    """Device pointer of a contiguous tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_ptr(device) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream
