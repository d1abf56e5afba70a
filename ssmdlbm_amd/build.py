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
SOURCES = [CSRC / "ssmd.hip", CSRC / "io.cpp", CSRC / "closure.hip"] + sorted(CSRC.glob("lattice_q*.hip"))
DEPS = [CSRC / "lattice.hpp", CSRC / "kernels.hpp", CSRC / "launch.hpp", CSRC / "lattice_inst.hpp",
        ROOT / "include" / "ssmd.h"]
OBJDIR = LIBDIR / "obj"
DRIVER_SRC = CSRC / "host" / "ssmd_run.cpp"
DRIVER = LIBDIR / "ssmd_run"
PROBE_SRC = CSRC / "probe" / "hbm_probe.hip"
PROBE = LIBDIR / "libssmd_probe.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SSMD_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3",
    "-std=c++20",
    f"--offload-arch={ARCH}",
    "-fPIC",
    # contract a*b+c only within one expression: per-cell arithmetic is then
    # identical in every tile-shape instantiation (ensemble == solo, bit for bit)
    "-ffp-contract=on",
    # packed FP32 (v_pk_*_f32) generation is disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-parameter",
]


def _run(cmd, verbose):
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass reports the device-only target feature; that is expected
    noise = "not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}) on {cmd[-3]}:\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)


def build_library(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile libssmd.so (one object per source, in parallel); always rebuilds."""
    from concurrent.futures import ThreadPoolExecutor

    OBJDIR.mkdir(parents=True, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = OBJDIR / (src.stem + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", "-I", str(ROOT / "include"), str(src), "-o", str(obj)])
    jobs = jobs or max(1, min(len(cmds), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_driver(verbose: bool = True) -> Path:
    """Compile the native host driver ssmd_run against libssmd.so."""
    tmp = DRIVER.with_suffix(".tmp")
    cmd = [HIPCC, "-O2", "-std=c++20", "-I", str(ROOT / "include"), str(DRIVER_SRC), "-L", str(LIBDIR), "-lssmd",
           "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc (driver) failed ({res.returncode}):\n{res.stderr}")
    os.replace(tmp, DRIVER)
    return DRIVER


def build_probe(verbose: bool = True) -> Path:
    """Compile the HBM streaming probe (libssmd_probe.so) that bench.py measures the ceiling with."""
    LIBDIR.mkdir(parents=True, exist_ok=True)
    tmp = PROBE.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-shared", str(PROBE_SRC), "-o", str(tmp)], verbose)
    os.replace(tmp, PROBE)
    return PROBE


if __name__ == "__main__":
    build_library()
    build_driver()
    build_probe()
