"""Sanitizer gate (CPU): the oracle and the library's host code built with
AddressSanitizer + UndefinedBehaviorSanitizer and driven without a GPU.

* ``oracle/sanitize_check.c`` renders small scenes through the oracle (several
  threads, uneven resolutions, every diagnostic buffer) and checks a closed form;
* ``csrc/sanitize_check.cpp`` exercises the library's scratch planning (both
  orderings, culling, per-camera resolutions) and the argument validation of
  every entry point, all of which return before any device work.

A sanitizer report aborts the driver (halt_on_error), which fails the test.
Leak detection is off: the HIP runtime the library loads keeps process-lifetime
allocations.
"""

from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SANITIZE_ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
                    UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def build_and_run(directory: Path, binary: Path) -> str:
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-C", str(directory), f"-j{jobs}", "sanitize"], check=True, capture_output=True,
                   timeout=600)
    result = subprocess.run([str(binary)], env=SANITIZE_ENV, capture_output=True, text=True, timeout=300)
    output = result.stdout + result.stderr
    assert result.returncode == 0, output[-4000:]
    assert "AddressSanitizer" not in output and "runtime error" not in output, output[-4000:]
    return output


def test_oracle_under_address_and_undefined_sanitizers():
    output = build_and_run(ROOT / "oracle", ROOT / "_build" / "sanitize" / "oracle_sanitize_check")
    assert "oracle OK" in output


@pytest.mark.skipif(shutil.which("hipcc") is None and not Path("/opt/rocm/bin/hipcc").exists(),
                    reason="hipcc is needed to build the library")
def test_library_host_code_under_address_and_undefined_sanitizers():
    output = build_and_run(ROOT / "vectorized-gaussian-renderer-public_amd" / "csrc",
                           ROOT / "_build" / "sanitize" / "sanitize_check")
    assert "library host code OK" in output
