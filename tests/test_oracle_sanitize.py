"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer
(host code only, CPU).

tests/sanitize/oracle_driver.cpp is built from the oracle's sources with
g++ -fsanitize=address,undefined (the ASan runtime linked statically, so
the driver needs no preloaded runtime) and run as a child process on small
cases that between them take every lattice, storage format, wrap mode, face
kind and closure of the oracle; a memory error, a leak or undefined
behaviour aborts it.  ThreadSanitizer is not used: libgomp is not built with
it, so it reports the OpenMP runtime's own synchronisation.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

from ssmdlbm_amd import abi, cases

ROOT = Path(__file__).resolve().parent.parent


def _cases():
    out = []
    out.append(cases.JET_D3Q19_SMALL.with_(nx=20, ny=18, nz=24, diameter=6).params())
    out.append(cases.JET_D3Q19_SMALL.with_(nx=16, ny=20, nz=22, diameter=6, q=27, pop_format=abi.POP_F16,
                                           wrap=abi.WRAP_TRUE).params())
    two = cases.JET_D3Q19_SMALL.with_(nx=18, ny=22, nz=20, diameter=6, pop_format=abi.POP_BF16).params()
    two.water.enabled = 1
    two.water.center_a, two.water.center_b, two.water.r2 = 8.5, 6.0, 4.0
    two.water.u, two.water.phi = 0.04, 0.0
    two.outflow_y = 1
    two.wrap_y = abi.WRAP_NONE
    out.append(two)
    smag = cases.closed_box(12, q=19, sigma=0.05)
    smag.closure, smag.cs = abi.CLOSURE_SMAGORINSKY, 0.1
    out.append(smag)
    learned = cases.closed_box(10, q=27, sigma=0.05)
    learned.closure, learned.tau_scale, learned.flux_coeff = abi.CLOSURE_LEARNED, 1.0, 1.5
    learned.wrap_z = abi.WRAP_REFERENCE
    out.append(learned)
    return out


@pytest.mark.skipif(shutil.which(os.environ.get("CXX", "g++")) is None, reason="no host C++ compiler")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "oracle_driver"
    cmd = [os.environ.get("CXX", "g++"), "-O1", "-g", "-fopenmp", "-std=c++17",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-static-libasan",
           "-I", str(ROOT / "include"), str(ROOT / "oracle" / "ssmd_oracle.cpp"), str(ROOT / "oracle" / "mlp_oracle.cpp"),
           str(ROOT / "tests" / "sanitize" / "oracle_driver.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=600)
    files = []
    for i, p in enumerate(_cases()):
        f = tmp_path / f"case{i}.bin"
        f.write_bytes(bytes(p))
        files.append(str(f))
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")
    r = subprocess.run([str(exe)] + files, capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, f"exit {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-4000:]}"
    assert "sanitized oracle run: ok" in r.stdout
    assert r.stdout.count("case ") == 2 * len(files)
