"""Build the gfx950 hot-path library (libssmd.so) in-tree with hipcc.

The oracle used by the tests is built by ``oracle/build.py``; this module
builds only the product library.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libssmd.so"
SOURCES = [CSRC / "ssmd.hip"]
DEPS = [CSRC / "lattice.hpp", CSRC / "kernels.hpp", ROOT / "include" / "ssmd.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SSMD_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3",
    "-std=c++20",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-shared",
    # packed FP32 (v_pk_*_f32) generation is disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-parameter",
]


def build_library(verbose: bool = True) -> Path:
    """Compile libssmd.so; always rebuilds (the build check compiles every call)."""
    LIBDIR.mkdir(parents=True, exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), *map(str, SOURCES), "-o", str(tmp)]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass reports the device-only target feature; that is expected
    noise = "not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build_library()
