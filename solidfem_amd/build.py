"""Builds libsolidfem_hip.so in-tree with hipcc for gfx950, and the oracle's checker library.

Every source is compiled as HIP for gfx950 with -ffp-contract=off: the solve path repeats the
reference's IEEE roundings (no fused multiply-add anywhere, host setup included), and packed
FP32 instructions are disabled for the device code.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
OUT_DIR = PACKAGE / "_build"
LIBRARY = PACKAGE / "libsolidfem_hip.so"

SOURCES = ["sf_mesh.cpp", "sf_element.cpp", "sf_setup.cpp", "sf_device.hip", "sf_batch.hip", "sf_assembly.hip", "sf_setup_device.hip", "sf_serve.cpp", "sf_noalloc.hip", "sf_capi.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "--offload-arch=gfx950",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-parameter",
    "-I" + str(ROOT / "include"),
]


def _compile(source: str) -> Path:
    src = CSRC / source
    obj = OUT_DIR / (source + ".o")
    cmd = [HIPCC, "-x", "hip", "-c", str(src), "-o", str(obj)] + COMMON
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {source}:\n{proc.stderr[-6000:]}")
    if proc.stderr.strip():
        sys.stderr.write(proc.stderr)
    return obj


def build_library(jobs: int | None = None) -> Path:
    """Compiles every source and links the shared library. Always rebuilds."""
    OUT_DIR.mkdir(exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(_compile, SOURCES))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-o", str(tmp)] + [str(o) for o in objects] + [
        "--offload-arch=gfx950",
        "-fPIC",
        "-lpthread",
    ]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr[-6000:]}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


def build_oracle() -> Path:
    """Builds the oracle (test infrastructure only) with its own Makefile."""
    oracle = ROOT / "oracle"
    proc = subprocess.run(["make", "-s", "-C", str(oracle), "all"], capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{proc.stderr[-4000:]}")
    return oracle / "librefcpu.so"


if __name__ == "__main__":
    print(build_library())
    print(build_oracle())
