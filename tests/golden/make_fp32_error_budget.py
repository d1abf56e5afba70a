"""FP32 error budget of the reference arithmetic on the 64x64x128 D3Q19 jet.

Runs the oracle's restatement of the reference step in FP64 and in FP32 (the
reference's scalar_t) on the same inputs and records the relative L2 error
of rho, ux, uy, uz, phi of the FP32 run after 100, 300 and 1000 steps.  The
GPU acceptance test holds the FP32 kernels to errors no larger than these.
Writes tests/golden/fp32_error_budget_jet_d3q19_64x64x128.json.
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.oracle import Oracle  # noqa: E402
from ssmdlbm_amd import cases  # noqa: E402

CHECKPOINTS = (100, 300, 1000)


def main() -> None:
    p = cases.JET_D3Q19_SMALL.params()
    runs = {}
    for fp64 in (True, False):
        o = Oracle(p, fp64=fp64, threads=8)
        o.init()
        snaps = {}
        for s in CHECKPOINTS:
            o.step(s - o.step_index)
            snaps[s] = o.fields()
        runs[fp64] = snaps
    out = {"workload": cases.JET_D3Q19_SMALL.name, "checkpoints": {}}
    for s in CHECKPOINTS:
        a, b = runs[False][s], runs[True][s]
        out["checkpoints"][str(s)] = {k: float(np.linalg.norm(a[k] - b[k]) / np.linalg.norm(b[k])) for k in a}
    path = Path(__file__).resolve().parent / "fp32_error_budget_jet_d3q19_64x64x128.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
