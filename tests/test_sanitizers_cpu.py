"""The CPU oracle and the host client under AddressSanitizer and
UndefinedBehaviorSanitizer (with leak detection): tests/sanitize/driver.cpp runs
a whole answer at (4, 2) through them (keygen, query, encoding, expansion,
multiply, fold, decode; packed query, conversion, packing, compact decode;
transforms at q - 1 and coefficients equal to Q; the client's error paths).
Any sanitizer report ends the driver with a non-zero status."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")


def _have_asan() -> bool:
    if not shutil.which("gcc") or not shutil.which("g++"):
        return False
    out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return os.path.isabs(out) and os.path.exists(out)


@pytest.mark.skipif(not _have_asan(), reason="gcc with libasan not available")
def test_oracle_and_client_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", SAN], check=True, capture_output=True)
    # verify_asan_link_order=0: the environment may preload a library ahead of the sanitizer runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(SAN, "_build", "driver")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, f"status {r.returncode}\n{r.stdout}\n{r.stderr[-4000:]}"
    assert "sanitize: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
