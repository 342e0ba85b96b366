"""Builds libwavetorch_hip.so for gfx950 in-tree with hipcc.

Usage: ``python -m wavetorch_amd.csrc.build`` (also called by
``__graft_entry__.build()``).  Every translation unit is compiled with

* ``--offload-arch=gfx950``               MI355X only
* ``-ffp-contract=off``                   no implicit FMA: the forward kernels
                                          reproduce the reference's per-op
                                          fp32 rounding; FMAs in the adjoint
                                          are explicit ``fmaf`` calls
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``
                                          no packed-FP32 VALU instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "_lib")
LIBNAME = "libwavetorch_hip.so"
SOURCES = ["wt_runtime.hip", "wt_plain.hip", "wt_tiled.hip", "wt_colseg.hip", "wt_geom.hip",
           "wt_loss.hip", "wt_autotune.hip", "wt_f64.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


# Host-side AddressSanitizer build (plan, tiling and table construction,
# argument checks): the sanitizer applies to the host pass only, device code
# is compiled as in the product build.  The library's __asan_* references are
# resolved at load time from the clang runtime, which the loading process
# preloads (tests/test_host_asan.py).
ASAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer",
              "-Xarch_host", "-g"]
ASAN_OUT = os.path.join(OUT, "asan")


def asan_runtime() -> str:
    """Path of clang's shared AddressSanitizer runtime next to hipcc."""
    import glob
    cands = sorted(glob.glob(os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "lib",
                                          "clang", "*", "lib", "linux",
                                          "libclang_rt.asan-x86_64.so")))
    if not cands:
        cands = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not cands:
        raise FileNotFoundError("libclang_rt.asan-x86_64.so not found next to hipcc")
    return os.path.realpath(cands[-1])


def _compile(src: str, obj: str, verbose: bool, extra=()) -> None:
    cmd = [HIPCC, *FLAGS, *extra, "-c", os.path.join(HERE, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(line for line in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msg}")
    if verbose and msg.strip():
        print(msg)


def build(verbose: bool = False, jobs: int = 4, out_dir: str | None = None,
          extra_flags=()) -> str:
    """Compiles every kernel source and links the shared library; returns its path.
    ``out_dir``/``extra_flags`` build a variant elsewhere (diagnostic builds)."""
    out = out_dir or OUT
    os.makedirs(out, exist_ok=True)
    objdir = os.path.join(out, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, s.replace(".hip", ".o")) for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, o, verbose, tuple(extra_flags))
                for s, o in zip(SOURCES, objs)]
        for f in futs:
            f.result()
    lib = os.path.join(out, LIBNAME)
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, lib)
    return lib


def build_asan(verbose: bool = False, jobs: int = 4) -> str:
    """The host-AddressSanitizer variant of the library, in ``_lib/asan``."""
    return build(verbose=verbose, jobs=jobs, out_dir=ASAN_OUT, extra_flags=ASAN_FLAGS)


if __name__ == "__main__":
    if "--asan" in sys.argv:
        print(build_asan(verbose="-v" in sys.argv))
    else:
        print(build(verbose="-v" in sys.argv))
