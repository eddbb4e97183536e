"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle's sources built with both (`make -C oracle asan`), then every CPU test
that uses it run again in a child process on that build (libasan preloaded,
leak checks off: the interpreter's own allocations are not the oracle's). A
heap or stack overrun, a use after free, a signed overflow or a misaligned
access anywhere in the oracle aborts the child with a report."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _runtime() -> str | None:
    try:
        p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_under_sanitizers():
    asan = _runtime()
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "asan"], check=True, timeout=600)
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               PSX_ORACLE_LIB=str(ROOT / "oracle" / "_asan" / "liboracle.so"))
    cmd = [sys.executable, "-m", "pytest", str(ROOT / "tests"), "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
           "--ignore", str(Path(__file__).resolve()), "--ignore", str(ROOT / "tests" / "test_abi_host.py")]
    p = subprocess.run(cmd, env=env, cwd=str(ROOT), capture_output=True, text=True, timeout=1500)
    out = p.stdout + p.stderr
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-6000:]
    assert p.returncode == 0, out[-6000:]
    assert " passed" in out
