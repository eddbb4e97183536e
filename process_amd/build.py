"""Builds libpsx.so (gfx950 HIP kernels + host runtime) in-tree with hipcc.

Run ``python -m process_amd.build``; the repository-level build() entry point calls it too.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libpsx.so"

SOURCES = ["psx_host.cpp", "psx_genome.hip", "psx_place.hip", "psx_seq.hip", "psx_serve.cpp", "psx_noalloc.hip", "psx_probe.hip",
           "psx_sam.cpp", "psx_context.hip", "psx_store.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
# hipcc flag: keep the compiler from emitting packed fp32 arithmetic.
NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src: str) -> Path:
    out = OBJ / (src + ".o")
    cmd = [HIPCC, *COMMON]
    if src.endswith(".hip"):
        cmd += ["-x", "hip", f"--offload-arch={ARCH}", *NO_PACKED_FP32]
    cmd += ["-c", str(CSRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build_lib(jobs: int = 4) -> Path:
    OBJ.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking libpsx.so failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    """Compiles the CPU oracle (test infrastructure) into oracle/liboracle.so."""
    r = subprocess.run(["make", "-s", "-B", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building the oracle failed:\n{r.stdout}\n{r.stderr}")
    return ROOT / "oracle" / "liboracle.so"


def main() -> None:
    print(build_lib())
    print(build_oracle())


if __name__ == "__main__":
    sys.exit(main())
