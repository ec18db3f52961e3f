"""The host side of the C ABI under AddressSanitizer and UndefinedBehaviorSanitizer, no GPU needed.

myplacement_amd/libmp_gp_hip_hostsan.so is the product library's host code built with both
sanitizers (myplacement_amd.build.hostsan_flags; no device code).  The host tests of
tests/test_host.py (layout, fillers, static density, preconditioner and random placement bit-equal to
the oracle, the error paths, the service and batch argument checks, the no-allocation entry point's
refusals) run again in a child process that preloads the sanitizer runtime and loads that library;
a memory error (out-of-bounds access, use after free) or an undefined operation in the host code
ends the child with the sanitizer's report.  Leak checking is off: the interpreter itself keeps
allocations to its end.
"""

import os
import subprocess
import sys

import pytest

from myplacement_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_abi_under_address_and_undefined_behaviour_sanitizers():
    if not os.path.exists(build.HOSTSAN_LIB):
        pytest.skip("the host-sanitized library is not built (python -m myplacement_amd.build)")
    runtime = build.sanitizer_runtime()
    if not runtime:
        pytest.skip("no shared AddressSanitizer runtime next to hipcc")
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": runtime,
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=86",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "MP_GP_LIBRARY": build.HOSTSAN_LIB,
    })
    # the child really runs the instrumented library beside the sanitizer runtime
    probe = ("import ctypes, os; lib = ctypes.CDLL(os.environ['MP_GP_LIBRARY']); "
             "ctypes.CDLL(None)['__asan_report_load8']; print(lib.mp_gp_compiled())")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "1", r.stdout + r.stderr
    # the struct-size check compiles and runs a C program, which would inherit the preload
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
           "-k", "not struct_sizes", os.path.join(ROOT, "tests", "test_host.py")]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-6000:]
    assert r.returncode == 0, out[-6000:]
    assert " passed" in out, out[-3000:]
