"""CPU: the host plan layer (ops-fft_amd/lib) under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C ops-fft_amd sanitize` builds libcann_ops_fft_san.so (the plan layer compiled with
-fsanitize=address,undefined, the HIP kernel objects unchanged); the boundary and tuning-cache tests
then run in a child process against it (ACLFFT_LIBRARY) with the sanitizer runtimes preloaded.  Any
report fails the run (UBSan halts on error; ASan aborts on an invalid access)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ops-fft_amd")


def runtime(name):
    try:
        p = subprocess.run(["g++", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    except OSError:
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_plan_layer_under_asan_and_ubsan():
    asan, ubsan = runtime("libasan.so"), runtime("libubsan.so")
    if asan is None or ubsan is None:
        pytest.skip("sanitizer runtimes not available")
    if not os.path.exists(os.path.join(PKG, "libcann_ops_fft.so")):
        pytest.skip("library not built")
    subprocess.check_call(["make", "-s", "-C", PKG, "sanitize"])
    env = dict(os.environ, LD_PRELOAD=asan + ":" + ubsan,
               ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               ACLFFT_LIBRARY=os.path.join(PKG, "libcann_ops_fft_san.so"))
    out = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                          os.path.join(ROOT, "tests", "test_boundary.py")],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    print(out.stdout[-3000:], out.stderr[-3000:])
    assert out.returncode == 0
    assert "ERROR: AddressSanitizer" not in out.stderr and "runtime error" not in out.stderr
