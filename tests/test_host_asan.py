"""Host-side AddressSanitizer run of the library's CPU-reachable code.

The library is built a second time with AddressSanitizer on the host pass
(``python -m wavetorch_amd.csrc.build --asan``: plan resolution, tiling and
table construction, argument checks; device code unchanged) and
tests/test_host.py runs against it in a child process that preloads clang's
ASan runtime.  Any invalid host memory access there aborts the child with an
AddressSanitizer report, which fails this test.  No kernel is launched.
"""
import os
import subprocess
import sys

import pytest

from wavetorch_amd.csrc import build as hipbuild

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN_LIB = os.path.join(hipbuild.ASAN_OUT, hipbuild.LIBNAME)


def _stale() -> bool:
    if not os.path.exists(ASAN_LIB):
        return True
    t = os.path.getmtime(ASAN_LIB)
    srcs = [os.path.join(hipbuild.HERE, s) for s in hipbuild.SOURCES]
    srcs += [os.path.join(hipbuild.HERE, "wt_internal.h"),
             os.path.join(ROOT, "include", "wavetorch_hip.h")]
    return any(os.path.getmtime(s) > t for s in srcs)


@pytest.fixture(scope="module")
def asan_lib():
    try:
        rt = hipbuild.asan_runtime()
    except FileNotFoundError as e:
        pytest.skip(str(e))
    if _stale():
        hipbuild.build_asan()
    return ASAN_LIB, rt


def test_asan_library_is_instrumented(asan_lib):
    lib, _ = asan_lib
    out = subprocess.run(["nm", "-D", "--undefined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    assert "__asan_report_load4" in out and "__asan_report_store4" in out


def test_host_tests_clean_under_asan(asan_lib):
    lib, rt = asan_lib
    env = dict(os.environ, LD_PRELOAD=rt, WAVETORCH_AMD_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:abort_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_host.py"),
                        "-x", "-q", "-p", "no:cacheprovider"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert " passed" in out
