"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer:
`make -C pylblrtm_amd/csrc sanitize` builds pylblrtm_amd/csrc/host_check.cpp
with the library's sources (host pass instrumented, leak detection on,
every finding fatal) and runs it on the CPU: the .par reader on files that
cover its mmap/read and threaded paths and its malformed-line handling, and
the C ABI's argument validation.  No GPU is needed or touched."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc (ROCm) not installed")
@pytest.mark.timeout(900)
def test_host_code_is_clean_under_asan_and_ubsan():
    res = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "pylblrtm_amd", "csrc"),
                          "sanitize"], capture_output=True, text=True, timeout=880)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert "0 failed" in out, out[-4000:]
