"""The engine's host code under AddressSanitizer + UndefinedBehaviorSanitizer
and under ThreadSanitizer (CPU only): tests/sanitize/Makefile builds the host
sources with g++ and the sanitizer flags beside a test double of the device
entry points, and tests/sanitize/harness.cpp drives model I/O, prepare, the
verifier, the session protocol, checkpoint files, the serving loop's
threads under concurrent submitters and the certify-beside handoff of the
batched engine.  A sanitizer report aborts the harness."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("variant", ["asan", "tsan"])
def test_host_code_is_sanitizer_clean(variant, tmp_path):
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-s", "-f", "tests/sanitize/Makefile", f"VARIANT={variant}", f"-j{jobs}"], cwd=ROOT,
                   check=True, timeout=900)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1")
    out = subprocess.run([str(ROOT / "tests" / "sanitize" / "_build" / variant / "harness"),
                          str(ROOT / "tests" / "golden" / "netlib"), str(tmp_path)], env=env, capture_output=True,
                         text=True, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert "harness: 0 failure(s)" in out.stdout
    assert "Sanitizer" not in out.stderr, out.stderr[-4000:]
