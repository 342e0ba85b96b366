"""ctypes binding of ``wave_oracle.c`` (test infrastructure only).

The shared object is built by :func:`build` (called from
``__graft_entry__.build()``) with ``-ffp-contract=off`` so that every fp32
operation is rounded individually, as torch's CPU kernels round them.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "wave_oracle.c")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libwave_oracle.so")

_lib = None


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["gcc", "-O2", "-std=c99", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
           "-o", LIB + ".tmp", SRC, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        # a library older than its source would bind the wrong signatures
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        i = ctypes.c_int
        f = ctypes.c_float
        d = ctypes.c_double
        L.wo_laplacian_f32.argtypes = [P, P, i, i, i, f]
        L.wo_laplacian_f64.argtypes = [P, P, i, i, i, d]
        L.wo_forward_f32.argtypes = [i, i, i, i, P, P, f, f, P, i, P, i, P, P, P, i, i, P,
                                     f, f, f, f, f]
        L.wo_forward_f64.argtypes = [i, i, i, i, P, P, d, d, P, i, P, i, P, P, P, i, i, P,
                                     d, d, d, d, d]
        L.wo_adjoint_f64.argtypes = [i, i, i, i, P, P, d, d, P, i, P, i, P, i, P,
                                     d, d, d, d, d, P, P, P, P, P, P, P]
        for fn in (L.wo_forward_f32, L.wo_forward_f64, L.wo_adjoint_f64):
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def laplacian(u, h=1.0):
    """operators.py:5-11 on a (B,Nx,Ny) array (fp32 or fp64)."""
    u = np.ascontiguousarray(u)
    B, Nx, Ny = u.shape
    out = np.empty_like(u)
    if u.dtype == np.float32:
        lib().wo_laplacian_f32(_p(u), _p(out), B, Nx, Ny, float(h))
    else:
        u = u.astype(np.float64)
        out = np.empty_like(u)
        lib().wo_laplacian_f64(_p(u), _p(out), B, Nx, Ny, float(h))
    return out


def _nl_args(nl):
    """nl: None or dict(rho, sat_b0, sat_uth, c_nl, clamp=(cmin,cmax) or None)."""
    if nl is None:
        return 0, None, 0.0, 0.0, 0.0, 0.0, 0.0
    flags = 0
    if nl.get("sat_b0", 0.0) > 0:
        flags |= 1
    if nl.get("c_nl", 0.0) != 0:
        flags |= 2
    cl = nl.get("clamp")
    if cl is not None:
        flags |= 4
    cmin, cmax = cl if cl is not None else (0.0, 0.0)
    return flags, nl["rho"], nl.get("sat_b0", 0.0), nl.get("sat_uth", 1.0), nl.get("c_nl", 0.0), cmin, cmax


def forward(b, c, x, dt, h, src_px, taps, fields=False, nl=None, dtype=np.float32,
            field_stride=1):
    """rnn.py:21-72 forward.

    b, c: (Nx,Ny); x: (B,T); src_px: (n,3) int (x, y, source index);
    taps: (m,3) int (x, y, kind).  Returns (y (B,T,m), fields or None), the
    fields being u after the steps n with (n+1) % field_stride == 0, shaped
    (B, T // field_stride, Nx, Ny).
    """
    b = _c(b, dtype)
    c = _c(c, dtype)
    x = _c(x, dtype)
    B, T = x.shape
    Nx, Ny = b.shape
    src_px = np.ascontiguousarray(np.asarray(src_px, dtype=np.int32).reshape(-1, 3))
    taps = np.ascontiguousarray(np.asarray(taps, dtype=np.int32).reshape(-1, 3))
    y = np.zeros((B, T, len(taps)), dtype=dtype)
    fo = np.zeros((B, T // field_stride, Nx, Ny), dtype=dtype) if fields else None
    flags, rho, sb0, suth, cnl, cmin, cmax = _nl_args(nl)
    rho = _c(rho, dtype) if rho is not None else None
    fn = lib().wo_forward_f32 if dtype == np.float32 else lib().wo_forward_f64
    rc = fn(B, Nx, Ny, T, _p(b), _p(c), dt, h, _p(x), len(src_px), _p(src_px), len(taps),
            _p(taps), _p(y), _p(fo), int(field_stride), flags, _p(rho), sb0, suth, cnl, cmin,
            cmax)
    if rc != 0:
        raise MemoryError("wo_forward failed")
    return y, fo


def adjoint(b, c, x, dt, h, src_px, taps, ybar=None, fbar=None, nl=None, at_fields=None):
    """fp64 discrete adjoint.  Returns dict(gc, gb, grho, gx).

    at_fields: None (the adjoint of the fp64 forward), or the fp32 fields
    u_1..u_T, shaped (B, T, Nx, Ny), of a forward computed elsewhere: the
    fp64 reverse sweep is then taken at that trajectory (the exact adjoint of
    the computed forward)."""
    b = _c(b, np.float64)
    c = _c(c, np.float64)
    x = _c(x, np.float64)
    B, T = x.shape
    Nx, Ny = b.shape
    src_px = np.ascontiguousarray(np.asarray(src_px, dtype=np.int32).reshape(-1, 3))
    taps = np.ascontiguousarray(np.asarray(taps, dtype=np.int32).reshape(-1, 3))
    flags, rho, sb0, suth, cnl, cmin, cmax = _nl_args(nl)
    rho = _c(rho, np.float64) if rho is not None else None
    ybar = _c(ybar, np.float64)
    fbar = _c(fbar, np.float64)
    gc = np.zeros((Nx, Ny))
    gb = np.zeros((Nx, Ny))
    grho = np.zeros((Nx, Ny))
    gx = np.zeros((B, T))
    u32 = None
    if at_fields is not None:
        u32 = np.ascontiguousarray(at_fields, dtype=np.float32)
        if u32.shape != (B, T, Nx, Ny):
            raise ValueError(f"at_fields shape {u32.shape}, expected {(B, T, Nx, Ny)}")
    rc = lib().wo_adjoint_f64(B, Nx, Ny, T, _p(b), _p(c), dt, h, _p(x), len(src_px), _p(src_px),
                              len(taps), _p(taps), flags, _p(rho), sb0, suth, cnl, cmin, cmax,
                              _p(ybar), _p(fbar), _p(u32), _p(gc), _p(gb), _p(grho), _p(gx))
    if rc != 0:
        raise MemoryError("wo_adjoint failed")
    return dict(gc=gc, gb=gb, grho=grho, gx=gx)
