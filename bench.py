"""Throughput of the gfx950 per-timestep path: MLUPS and % of the HBM roofline.

Default workload: D3Q27 + D3Q7 two-phase jet, FP32, 128x128x512 (inlet
diameter 16, u_z = 0.05, sigma = 0.05), W warm-up steps then K timed
steps replayed from HIP graphs (defaults 5000 + 5000: throughput over the
last 5000 of 10000 steps).  MLUPS counts every cell, ghosts included, as
the reference does (main.cu:199-203).

Also measured in the same run and reported in the same JSON line:
  * the fused lattice kernel's average duration from HIP events on the
    stream it runs on -> roofline.achieved (algorithmic bytes / time);
  * the 16-bit storage form against its FP32 twin (D3Q19 + D3Q7,
    256x256x1024), as the metric asks for FP16 storage vs FP32;
  * the CPU baseline: the oracle's restatement of the reference step in
    FP32 with OpenMP on a bounded sample of the same workload.
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 matrix peak
# closure network 8 -> 64 -> 64 -> 64 -> 1: 2 flop per multiply-add
MLP_FLOP_PER_CELL = 2 * (8 * 64 + 64 * 64 + 64 * 64 + 64 * 1)
METRIC = "MLUPS + % HBM roofline, D3Q19/27+D3Q7 jet; FP16-storage vs FP32"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measure(case, steps: int, warmup: int, kernel_steps: int, tune: bool = True):
    import torch

    from ssmdlbm_amd import cases
    from ssmdlbm_amd.solver import Simulation

    p = case.params()
    sim = Simulation(p, "cuda:0")
    sim.init()
    interior = (p.nx - 2) * (p.ny - 2) * (p.nz - 2)
    tile = (int(sim.p.tile), int(sim.p.lz))
    if tune:  # the autotuner picks the lattice tile shape and z-chunk on this state,
        # then checks the z-chunk on 200 whole steps against the automatic one
        sim.step(2)
        tile = sim.tune(reps=10, confirm_steps=200)
    if warmup:
        sim.graph_run(warmup)
    torch.cuda.synchronize()
    first_step = sim.step_index + 1  # steps are numbered from 1 after t = 0
    t0 = time.perf_counter()
    sim.graph_run(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    mlups = p.cells * steps / wall / 1e6

    # fused lattice kernel alone, bracketed by events on its own stream
    stream = sim.stream
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(kernel_steps)]
    for e0, e1 in ev:
        sim.launch_phase()
        e0.record(stream)
        sim.launch_lattice()
        e1.record(stream)
        sim.launch_boundaries()
        sim.swap()
    torch.cuda.synchronize()
    times = sorted(e0.elapsed_time(e1) for e0, e1 in ev)
    k_ms = sum(times) / len(times)
    # the state after all timed steps: one step writing every moment, then a finiteness check
    sim.step(1, macros=True)
    sim.synchronize()
    finite = bool(torch.isfinite(sim.t["uz"]).all()) and bool(torch.isfinite(sim.t["phi"]).all())
    max_uz = float(sim.t["uz"].abs().max())
    kb = cases.lattice_kernel_bytes_per_cell(case.q, case.pop_format)
    achieved = interior * kb / (k_ms * 1e-3) / 1e9
    b_alg = cases.algorithmic_bytes_per_cell(case.q, case.pop_format)
    step_bytes = cases.step_bytes_per_cell(case.q, case.pop_format)
    sim.close()
    del sim
    torch.cuda.empty_cache()
    return {
        "tile": list(tile),
        "first_timed_step": first_step,
        "mlups": mlups,
        "ms_per_step": wall / steps * 1e3,
        "kernel_ms": k_ms,
        "kernel_ms_median": times[len(times) // 2],
        "kernel_gbs": achieved,
        "kernel_bytes_per_cell": kb,
        "interior_cells": interior,
        "step_bytes_per_cell": step_bytes,
        "reference_step_bytes_per_cell": b_alg,
        "step_frac": mlups * 1e6 * b_alg / (HBM_PEAK_GBS * 1e9),
        "step_frac_own_bytes": mlups * 1e6 * step_bytes / (HBM_PEAK_GBS * 1e9),
        "state_finite": finite,
        "max_abs_uz": max_uz,
    }


def measure_closure(case, tile, steps: int, warmup: int, kernel_steps: int):
    """The workload with the LEARNED closure and the trained network
    (ssmdlbm_amd/data/closure_jet_d3q27): MLUPS of the whole step, and the
    network pass over every cell (ssmd_closure_update: 8 -> 64 -> 64 -> 64 -> 1
    on bf16 MFMA) timed alone with events on the simulation's stream."""
    import torch

    from ssmdlbm_amd import abi
    from ssmdlbm_amd.closure import Network
    from ssmdlbm_amd.solver import Simulation

    net = Network.load(ROOT / "ssmdlbm_amd" / "data" / "closure_jet_d3q27")
    p = case.params()
    p.closure = abi.CLOSURE_LEARNED
    p.tile, p.lz = tile
    net.scales.apply(p)
    sim = Simulation(p, "cuda:0")
    net.load_into(sim)
    sim.init()
    sim.graph_run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.graph_run(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stream = sim.stream
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(kernel_steps)]
    for e0, e1 in ev:  # the same features each time: tau_t is rewritten with the same values
        e0.record(stream)
        sim.closure_update()
        e1.record(stream)
    torch.cuda.synchronize()
    k_ms = sum(e0.elapsed_time(e1) for e0, e1 in ev) / len(ev)
    sim.step(1, macros=True)
    sim.synchronize()
    finite = bool(torch.isfinite(sim.t["uz"]).all())
    tau_t = sim.t["tau_t"]
    tau_mean, tau_max = float(tau_t.mean()), float(tau_t.max())
    sim.close()
    del sim, net
    torch.cuda.empty_cache()
    cells = p.cells
    flop = cells * MLP_FLOP_PER_CELL
    return {
        "workload": case.name + "_learned_closure",
        "mlups": round(cells * steps / wall / 1e6, 1),
        "ms_per_step": round(wall / steps * 1e3, 5),
        "network_kernel_ms": round(k_ms, 5),
        "network_flop_per_cell": MLP_FLOP_PER_CELL,
        "network_tflops": round(flop / (k_ms * 1e-3) / 1e12, 1),
        "network_frac_of_bf16_mfma_peak": round(flop / (k_ms * 1e-3) / 1e12 / BF16_MFMA_PEAK_TFLOPS, 4),
        "network_hbm_gbs": round(cells * 20 / (k_ms * 1e-3) / 1e9, 1),
        "tau_t_mean": tau_mean,
        "tau_t_max": tau_max,
        "state_finite": finite,
    }


class ReferenceCase:
    """The reference's compile-time default case (constants.cuh:70-101):
    128x256x256, water jet on y = 0 and oil jet on z = 0, outflow y and z."""

    def __init__(self, q: int, pop_format: int):
        from ssmdlbm_amd import abi, cases

        self.q, self.pop_format = q, pop_format
        self.name = f"reference_default_d3q{q}_{['f32', 'f16', 'bf16'][pop_format]}_128x256x256"
        self._p = cases.reference_default(q)
        self._p.pop_format = pop_format
        self.cells = self._p.cells
        self._abi = abi

    def params(self):
        return self._p.copy()


def measure_reference_cases(steps: int, warmup: int):
    """MLUPS of the reference's own default case, beside its published
    RTX 3050 figures (README.md:35-37: 440 D3Q19, 377 D3Q27)."""
    from ssmdlbm_amd import abi

    out = {}
    for q in (19, 27):
        for fmt in (abi.POP_F32, abi.POP_F16):
            c = ReferenceCase(q, fmt)
            m = measure(c, steps, warmup, 20)
            out[c.name] = {"mlups": round(m["mlups"], 1), "kernel_ms": round(m["kernel_ms"], 4),
                           "state_finite": m["state_finite"]}
    out["published_rtx3050_mlups"] = {"d3q19": 440, "d3q27": 377}
    return out


def measure_ensemble(steps: int, warmup: int, n_jets: int = 32):
    """The batching form: a seeded ensemble of jets advanced by one launch per
    kernel kind per step (graph-replayed, autotuned tile), against the same
    jets as separate simulations with their own graphs: (serial) all on one
    stream, so they run strictly one after another, and (streams) each on its
    own stream, so the GPU may overlap them."""
    import torch

    from ssmdlbm_amd import cases
    from ssmdlbm_amd.solver import Ensemble, Simulation

    members = cases.ensemble(n_jets, seed=4)
    ens = Ensemble([c.params() for c in members], "cuda:0")
    ens.init()
    ens.step(2)
    tile = ens.tune(reps=5)
    ens.graph_run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.graph_run(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = ens.cells
    ens.close()
    del ens
    torch.cuda.empty_cache()
    def separate(shared_stream):
        sims = [Simulation(c.params(), "cuda:0", stream=shared_stream) for c in members]
        for s in sims:
            s.init()
            s.graph_run(max(warmup // 10, 1))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for s in sims:
                s.graph_run(1)
        torch.cuda.synchronize()
        w = time.perf_counter() - t0
        for s in sims:
            s.close()
        del sims
        torch.cuda.empty_cache()
        return w

    wall_serial = separate(torch.cuda.Stream())
    wall_streams = separate(None)
    return {
        "jets": n_jets,
        "cells": cells,
        "tile": tile,
        "steps": steps,
        "mlups": round(cells * steps / wall / 1e6, 2),
        "ms_per_step": round(wall / steps * 1e3, 4),
        "serial_mlups": round(cells * steps / wall_serial / 1e6, 2),
        "batched_over_serial": round(wall_serial / wall, 3),
        "separate_streams_mlups": round(cells * steps / wall_streams / 1e6, 2),
        "batched_over_separate_streams": round(wall_streams / wall, 3),
    }


def cpu_info() -> dict:
    """The host CPU as lscpu reports it: model name, logical CPUs, the CPUs this process may use."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "allowed_cpus": allowed}


def _oracle_mlups(params, threads: int, budget_s: float, warmup: int = 2, max_steps: int = 200):
    """MLUPS of the FP32 oracle over a bounded number of steps (at least 3)."""
    from oracle.oracle import Oracle

    o = Oracle(params, fp64=False, threads=threads)
    o.init()
    t0 = time.perf_counter()
    o.step(warmup)
    per = (time.perf_counter() - t0) / warmup
    n = int(max(3, min(max_steps, budget_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    o.step(n)
    wall = time.perf_counter() - t0
    return params.cells * n / wall / 1e6, n, wall


def cpu_baseline(case, budget_s: float = 15.0, full_small: bool = True):
    """The oracle's FP32 restatement of the reference schedule (K1-K4 and the
    boundaries, OpenMP over z) on the host cores: a bounded sample of the
    bench workload on every allowed core (at most 16, the GPU machines' share)
    and on one, and the 64x64x128 jet's full 1000 steps."""
    from ssmdlbm_amd import cases

    # thread placement of the reference CPU baseline plan; read by the oracle's
    # OpenMP runtime when the oracle library is first loaded (here)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    info = cpu_info()
    threads = max(1, min(16, info["allowed_cpus"]))
    mlups, n, wall = _oracle_mlups(case.params(), threads, budget_s)
    # one thread on the same jet shortened to 64 planes (1 M cells), so that a
    # few steps take seconds; the CPU rate per cell hardly depends on nz
    short = case.with_(nz=64)
    one, n1, _ = _oracle_mlups(short.params(), 1, budget_s / 3, warmup=1, max_steps=20)
    out = {
        "value": round(mlups, 3),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{case.name}: {n} timed steps after 2 warm-up steps of the oracle's FP32 restatement of the "
                  f"reference schedule (K1-K4 + boundaries), OpenMP over z, {threads} threads",
        "cpu_model": info["model"],
        "nproc": info["nproc"],
        "allowed_cpus": info["allowed_cpus"],
        "omp_binding": {"OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND", "unset"),
                        "OMP_PLACES": os.environ.get("OMP_PLACES", "unset")},
        "single_thread_mlups": round(one, 3),
        "single_thread_sample": f"{case.name} shortened to {short.nx}x{short.ny}x{short.nz}: {n1} timed steps "
                                f"after 1 warm-up step, 1 thread",
        "gbs_at_B_alg": round(mlups * 1e6 * cases.algorithmic_bytes_per_cell(case.q, case.pop_format) / 1e9, 2),
    }
    if full_small:
        small = cases.JET_D3Q19_SMALL
        from oracle.oracle import Oracle

        o = Oracle(small.params(), fp64=False, threads=threads)
        o.init()
        o.step(10)  # warm-up, excluded
        t0 = time.perf_counter()
        o.step(1000)
        wall1 = time.perf_counter() - t0
        out[small.name + "_1000_steps"] = {"seconds": round(wall1, 3),
                                           "mlups": round(small.cells * 1000 / wall1 / 1e6, 3),
                                           "threads": threads}
    return out


def copy_bandwidth(reps: int = 10) -> float:
    """Device-to-device copy of 2^28 floats (1 GiB read + 1 GiB written), GB/s:
    the practical streaming ceiling the lattice kernel is compared with."""
    import torch

    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * src.numel() * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def profile_mode(case, steps: int, tune: bool, tile=None) -> int:
    """Tuned like the timed run (or with the given (tile, lz)), then `steps`
    eager steps (one dispatch per kernel, as counter collection wants them),
    then torch copies of 2^28 floats = 1 GiB read + 1 GiB written (the
    byte-count reference)."""
    import torch

    from ssmdlbm_amd.solver import Simulation

    sim = Simulation(case.params(), "cuda:0")
    sim.init()
    fixed = tile is not None
    if fixed:
        sim.p.tile, sim.p.lz = tile
    elif tune:
        sim.step(2)
        tile = sim.tune(reps=10)
    tile = (int(sim.p.tile), int(sim.p.lz))
    sim.step(steps)
    sim.synchronize()
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    how = "at the given tile" if fixed else "after tuning" if tune else "at the default tile"
    log(f"profile mode: tile {tile}, {steps} steps of {case.name} {how}, 3 copies of 1 GiB")
    return 0


def kernel_source_digest() -> str:
    """sha256 (16 hex digits) of the lattice kernel sources, which a traffic
    profile records so that a stale one can be told from a current one."""
    import hashlib

    h = hashlib.sha256()
    csrc = ROOT / "ssmdlbm_amd" / "csrc"
    for name in ("kernels.hpp", "lattice.hpp", "lattice_inst.hpp", "launch.hpp", "ssmd.hip"):
        h.update((csrc / name).read_bytes())
    return h.hexdigest()[:16]


def load_traffic(workload: str):
    """(HBM bytes per launch, provenance) from the committed counter profile."""
    path = ROOT / "profiles" / f"k_lattice_{workload}_traffic.json"
    if not path.exists():
        return None, None
    try:
        d = json.loads(path.read_text())
    except (ValueError, OSError):
        return None, None
    src = d.get("kernel_source_sha256_16")
    prov = {"profile": str(path.relative_to(ROOT)), "measured_with": d.get("commands"),
            "kernel_sources_match": src == kernel_source_digest() if src else None}
    return d.get("hbm_bytes_per_launch"), prov


def hbm_probe_mode(sizes_gib) -> int:
    """The streaming probe alone (for rocprofv3 --kernel-trace): one JSON line per configuration."""
    from ssmdlbm_amd import probe

    for r in probe.sweep(sizes_gib):
        print(json.dumps(r), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=5000)
    ap.add_argument("--workload", default="jet_d3q27_128x128x512")
    ap.add_argument("--kernel-steps", type=int, default=200)
    ap.add_argument("--form-steps", type=int, default=300)
    ap.add_argument("--no-forms", action="store_true", help="skip the FP16-vs-FP32 storage measurement")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default lattice tile shape")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only the HBM streaming probe (read / write / copy, cached and non-temporal, 1 and 4 GiB)")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="profiling mode for rocprofv3 --pmc: N eager steps of the workload, then a plain "
                         "1 GiB device copy (the byte-count reference), no timing output")
    ap.add_argument("--tile", type=int, nargs=2, metavar=("TILE", "LZ"),
                    help="profiling mode: run this lattice tile shape and z-chunk instead of tuning")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              f"(the multi-GPU z-slab decomposition is a design, not code)", file=sys.stderr)
        return 2

    import torch

    from ssmdlbm_amd import abi, cases

    if not torch.cuda.is_available():
        print("bench.py needs the GPU", file=sys.stderr)
        return 1
    abi.load()
    case = cases.WORKLOADS[args.workload]
    if args.hbm_probe:
        return hbm_probe_mode((1, 4))
    if args.profile_steps:
        return profile_mode(case, args.profile_steps, tune=not args.no_tune, tile=args.tile)
    log(f"workload {case.name}: warmup {args.warmup}, timed {args.steps}")
    main_m = measure(case, args.steps, args.warmup, args.kernel_steps, tune=not args.no_tune)
    log(f"main: {main_m}")
    copy_gbs = copy_bandwidth()
    log(f"device copy: {copy_gbs:.0f} GB/s")
    from ssmdlbm_amd import probe

    sweep = probe.sweep((1, 4))
    best = probe.ceiling(sweep)
    log(f"HBM probe best: { {k: v['gbs'] for k, v in best.items()} }")

    forms = None
    if not args.no_forms:
        forms = {}
        for c in (cases.JET_D3Q19_FP32_LARGE, cases.JET_D3Q19_FP16, cases.JET_D3Q19_BF16):
            m = measure(c, args.form_steps, 50, min(args.kernel_steps, 100), tune=not args.no_tune)
            forms[c.name] = {k: round(v, 4) if isinstance(v, float) else v for k, v in m.items()}
            log(f"form {c.name}: {m}")
        a = forms[cases.JET_D3Q19_FP16.name]["mlups"]
        b = forms[cases.JET_D3Q19_FP32_LARGE.name]["mlups"]
        forms["fp16_over_fp32_mlups"] = round(a / b, 4)
        forms["bf16_over_fp32_mlups"] = round(forms[cases.JET_D3Q19_BF16.name]["mlups"] / b, 4)
        forms["ensemble_32_jets_d3q19"] = measure_ensemble(args.form_steps, 50)
        log(f"ensemble: {forms['ensemble_32_jets_d3q19']}")
        forms["reference_default_case"] = measure_reference_cases(max(args.form_steps, 500), 200)
        log(f"reference default case: {forms['reference_default_case']}")
        forms["learned_closure"] = measure_closure(case, main_m["tile"], args.form_steps, 50,
                                                   min(args.kernel_steps, 100))
        forms["learned_closure"]["over_no_closure_mlups"] = round(forms["learned_closure"]["mlups"] /
                                                                  main_m["mlups"], 4)
        log(f"learned closure: {forms['learned_closure']}")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(case)
        log(f"cpu: {cpu}")

    kb = main_m["kernel_bytes_per_cell"]
    traffic, traffic_prov = load_traffic(case.name)
    ceiling_gbs = best["copy"]["gbs"]
    out = {
        "metric": METRIC,
        "value": round(main_m["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_m["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if case.pop_format == abi.POP_F32 else ("f16-storage/f32" if case.pop_format == abi.POP_F16
                                                                else "bf16-storage/f32"),
        "data": "synthetic",
        "config": {
            "workload": case.name,
            "lattice": f"D3Q{case.q}+D3Q7",
            "box": [case.nx, case.ny, case.nz],
            "tau": case.tau, "tau_g": case.tau_g, "sigma": case.sigma, "inlet_diameter": case.diameter,
            "u_inlet": case.u, "perturb": case.perturb,
            "timed_region": "HIP graph replay, 16-step graphs of (phase, fused lattice, boundary kernels)",
            "step_window": f"steps {main_m['first_timed_step']}..{main_m['first_timed_step'] + args.steps - 1} after "
                           f"t = 0 (2 steps before tuning, the tuner's step-level check of the z-chunk, "
                           f"{args.warmup} warm-up); throughput does not depend on the flow state",
            "lattice_tile": main_m["tile"],
            "mlups_cells": "all cells incl. ghost layer (reference convention)",
            "interior_mlups": round(main_m["mlups"] * main_m["interior_cells"] / case.cells, 2),
            "state_finite_after_run": main_m["state_finite"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(main_m["kernel_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(main_m["kernel_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_provenance": traffic_prov,
            "kernel": "k_lattice (fused computeForces+computeMoments+streamCollide)",
            "kernel_ms": round(main_m["kernel_ms"], 5),
            "algorithmic_bytes_per_launch": main_m["interior_cells"] * kb,
            "step_frac_at_B_alg": round(main_m["step_frac"], 4),
            "step_frac_at_own_step_bytes": round(main_m["step_frac_own_bytes"], 4),
            "device_copy_gbs": round(copy_gbs, 1),
            "frac_of_device_copy": round(main_m["kernel_gbs"] / copy_gbs, 4),
            "ceiling_gbs": ceiling_gbs,
            "frac_of_ceiling": round(main_m["kernel_gbs"] / ceiling_gbs, 4),
            "ceiling": {k: {kk: v[kk] for kk in ("gbs", "nontemporal", "grid", "gib")} for k, v in best.items()},
            "ceiling_note": "best rate of a hand-written dwordx4 stream (libssmd_probe.so) in this process; "
                            "copy counts read + write bytes, as the lattice kernel's achieved rate does",
        },
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
