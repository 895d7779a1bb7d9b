"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the CPU
oracle's whole pipeline and the library's host structure pass
(csrc/structure.cpp), built host-only with the ROCm clang (oracle/Makefile,
target sanitize) and run on small inputs.  The run must end cleanly (no
sanitizer report: UBSan is non-recoverable, ASan aborts on the first error)
and the library's structure and JDS must equal the oracle's."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANGXX = "/opt/rocm/lib/llvm/bin/clang++"


@pytest.mark.skipif(not os.path.exists(CLANGXX) or shutil.which("make") is None, reason="no ROCm clang")
def test_host_code_under_asan_and_ubsan():
    b = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                       timeout=300)
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "sanitize_check")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "sanitize check ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
