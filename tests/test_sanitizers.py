"""Host-code sanitizer run (CPU only): the oracle and the host library's
GPU-independent code built with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/native/Makefile) and driven over random and
edge-case inputs (tests/native/sanitize_driver.cpp); any report fails."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_host_code_under_asan_and_ubsan(tmp_path):
    out = tmp_path / "build"
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "native"), "OUT=%s" % out],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(out / "sanitize_driver"), ROOT, str(scratch)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "sanitize driver: ok" in r.stdout
    assert "AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout
