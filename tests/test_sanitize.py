"""Host sanitizer run: the parsers/writers and the CPU oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile)
and driven over the fixture files, every truncation of them, and the whole
oracle path (tests/sanitize/driver.cpp).  Any sanitizer report fails."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_sanitized_parsers_and_oracle(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize")], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24", OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(HERE, "sanitize", "_build", "driver"), ROOT, str(tmp_path)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "sanitize driver ok" in r.stdout
    assert "runtime error" not in r.stdout and "AddressSanitizer" not in r.stdout
