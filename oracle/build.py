"""Build the CPU oracle (test infrastructure; see ssmd_oracle.cpp header).

Produces oracle/_build/libssmd_oracle.so with g++ (-O3, OpenMP) from
ssmd_oracle.cpp (the lattice step) and mlp_oracle.cpp (the closure network).  The
instruction set is pinned to x86-64-v3 so the library built here also runs
on the host CPUs of the GPU machines.

The reference itself (brenogemelgo/ssmdLBM) is not built here: its sources
need cuda_runtime.h, math_constants.h, builtin_types.h and cuda_fp16.h,
which this image does not have, so there is no oracle/_ref.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
OUT = HERE / "_build" / "libssmd_oracle.so"
CXX = os.environ.get("CXX", "g++")


def build_oracle(verbose: bool = True) -> Path:
    OUT.parent.mkdir(parents=True, exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [CXX, "-O3", "-march=x86-64-v3", "-fopenmp", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
           "-I", str(ROOT / "include"), str(HERE / "ssmd_oracle.cpp"), str(HERE / "mlp_oracle.cpp"), "-o", str(tmp)]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build_oracle()
