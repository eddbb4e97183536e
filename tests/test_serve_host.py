"""The serving runtime's host logic on the CPU: the HBM arena and the
admission policy (process_amd/csrc/psx_arena.h) against a byte-map model,
built with g++ (the header has no device code)."""
from __future__ import annotations

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_arena_and_admission(tmp_path):
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "process_amd" / "csrc"),
                    str(ROOT / "tests" / "serve" / "arena_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "arena ok" in out.stdout
