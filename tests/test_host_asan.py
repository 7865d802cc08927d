"""The host code of libesc under AddressSanitizer and UndefinedBehavior-
Sanitizer (tests/asan: C ABI, parameter parser, lattices, FP64 small solves,
pass planners and kernel generator; HIP kernels built as usual).  Builds with
make and runs the driver on the CPU; any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

ASAN = os.path.join(ROOT, "tests", "asan")


@pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs make and hipcc")
def test_host_code_clean_under_asan_and_ubsan(tmp_path):
    b = subprocess.run(["make", "-s", "-j8", "-C", ASAN], capture_output=True, text=True, timeout=900)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", ESC_JIT_CACHE=str(tmp_path / "jit"))
    r = subprocess.run([os.path.join(ASAN, "_build", "host_check"), os.path.join(ROOT, "tests", "golden", "Midinput.inp")],
                       capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0 and "host checks passed" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
