"""Learned-closure study: training data from filtered fine-grid jets, training
on the GPU, and the a-posteriori comparison of coarse runs.

No counterpart in the reference (it has no subgrid model).  The fine twin
is the same physical jet at twice the resolution: 2n cells per axis,
interface width 2W, diameter 2D, centre 2c - 1/2, so coarse interior cell i
covers fine cells 2i-1, 2i (the box filter of ssmd_sgs_samples).  Time
scaling:

  diffusive (default)  dt_f = dt_c / 4: tau_f = tau_c, u_f = u_c / 2,
          sigma_f = sigma_c / 2; one coarse step is four fine steps.  Every
          physical parameter agrees, the interface mobility included (the
          D3Q7 update relaxes fully, so its mobility is 1/8 in lattice units
          on any grid, and dx^2 / dt is the same on both grids).
  acoustic  dt_f = dt_c / 2: tau_f = 1/2 + 2 (tau_c - 1/2), u_f = u_c,
          sigma_f = 2 sigma_c; one coarse step is two fine steps, but the fine
          grid's physical interface mobility is half the coarse grid's.

In both, the sharpening coefficient of the fine grid is that of width 2W.

Statistics, on the coarse grid for both (fine: Favre-filtered fields):
  centreline decay  time-averaged u_z on the four axis columns, per z;
  interface spectrum  time- and slice-averaged radial power spectrum of the
          interface-area density |grad phi| in horizontal slices.
Errors are relative L2 norms of the coarse statistic against the filtered
fine one.

Command line (one GPU):
  python -m ssmdlbm_amd.closure_study train --out DIR   fine runs -> network
  python -m ssmdlbm_amd.closure_study eval  --net DIR/closure --out DIR
  python -m ssmdlbm_amd.closure_study noise --out DIR   seed-to-seed spread
  python -m ssmdlbm_amd.closure_study sensitivity --report DIR/eval_report.json --out DIR
  python -m ssmdlbm_amd.closure_study calibrate --report DIR/eval_report.json --out DIR
  python -m ssmdlbm_amd.closure_study flux --report DIR/eval_report.json --out DIR
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np
import torch

from . import abi
from .cases import gamma_from_tau_g, viscosity_from_tau
from .closure import Network, Scales, offsets
from .solver import Simulation


@dataclass
class Geometry:
    nx: int = 128
    ny: int = 128
    nz: int = 512
    q: int = 27
    # Base relaxation time of the coarse grid.  The capillary-driven jets
    # diverge at tau 0.51 and 0.52 (with or without Smagorinsky, coarse and
    # fine alike); at 0.55 the fine twin (tau 0.6, sigma_f = 2 sigma) runs
    # 20000 steps for sigma <= 0.03 and diverges for sigma >= 0.06.
    tau: float = 0.55
    tau_g: float = 0.8
    width: float = 4.0
    diameter: int = 16
    perturb: float = 0.01
    cs: float = 0.1
    scaling: str = "diffusive"  # or "acoustic" (module docstring)

    @property
    def fine_steps(self) -> int:
        """Fine steps per coarse step."""
        return 4 if self.scaling == "diffusive" else 2


@dataclass
class Jet:
    u: float
    sigma: float
    seed: int


def coarse_params(g: Geometry, jet: Jet, closure: int = abi.CLOSURE_NONE,
                  scales: Scales | None = None) -> abi.Params:
    p = abi.Params()
    p.nx, p.ny, p.nz, p.q = g.nx, g.ny, g.nz, g.q
    p.pop_format = abi.POP_F32
    p.visc_water = p.visc_oil = viscosity_from_tau(g.tau)
    p.sigma = jet.sigma
    p.gamma = gamma_from_tau_g(g.tau_g, g.width)
    p.perturb, p.seed = g.perturb, jet.seed
    p.oil.enabled, p.oil.u, p.oil.phi = 1, jet.u, 1.0
    p.oil.center_a, p.oil.center_b = (g.nx - 1) * 0.5, (g.ny - 1) * 0.5
    r = g.diameter // 2
    p.oil.r2 = float(r * r)
    p.oil.diameter = g.diameter
    p.outflow_z = 1
    p.wrap_x = p.wrap_y = abi.WRAP_REFERENCE
    p.closure, p.cs = closure, g.cs
    (scales or Scales()).apply(p)
    return p


def fine_params(g: Geometry, jet: Jet) -> abi.Params:
    """The same jet at twice the resolution (g.scaling: diffusive or acoustic)."""
    c = coarse_params(g, jet)
    p = c.copy()
    p.nx, p.ny, p.nz = 2 * g.nx, 2 * g.ny, 2 * g.nz
    if g.scaling == "diffusive":
        p.visc_water = p.visc_oil = viscosity_from_tau(g.tau)
        p.oil.u = 0.5 * jet.u
        p.sigma = 0.5 * jet.sigma
    elif g.scaling == "acoustic":
        p.visc_water = p.visc_oil = viscosity_from_tau(0.5 + 2.0 * (g.tau - 0.5))
        p.sigma = 2.0 * jet.sigma
    else:
        raise ValueError(f"unknown scaling {g.scaling!r}")
    # the same physical interface width: twice the cells (the sharpening
    # coefficient halves; the D3Q7 mobility is fixed in lattice units)
    p.gamma = gamma_from_tau_g(g.tau_g, 2.0 * g.width)
    p.oil.center_a, p.oil.center_b = 2.0 * c.oil.center_a - 0.5, 2.0 * c.oil.center_b - 0.5
    r = g.diameter  # 2 * (D // 2) with D even
    p.oil.r2 = float(r * r)
    p.oil.diameter = 2 * g.diameter
    p.seed = jet.seed + 7919
    p.closure = abi.CLOSURE_NONE
    return p


def log(msg: str) -> None:
    print(f"[closure] {time.strftime('%H:%M:%S')} {msg}", flush=True)


class Filter:
    """ssmd_sgs_samples on device buffers for one (fine, coarse) pair."""

    def __init__(self, fine: abi.Params, coarse: abi.Params, device="cuda:0"):
        self.lib = abi.load()
        self.fine, self.coarse = fine.copy(), coarse.copy()
        nc = coarse.cells
        self.nc = nc
        self.work = torch.empty(int(self.lib.ssmd_sgs_work_bytes(C.byref(self.coarse))) // 4, device=device)
        self.feat = torch.empty(nc * 8, dtype=torch.int16, device=device)
        self.target = torch.empty(nc, device=device)
        self.weight = torch.empty(nc, device=device)

    def __call__(self, sim: Simulation) -> None:
        # torch reads these buffers on its own stream: no filter pass may
        # overlap a read of the previous one, and reads start after this one
        torch.cuda.synchronize(sim.device)
        abi.check(self.lib.ssmd_sgs_samples(C.byref(self.fine), C.byref(sim.d), C.byref(self.coarse),
                                            self.work.data_ptr(), self.feat.data_ptr(), self.target.data_ptr(),
                                            self.weight.data_ptr(), sim.stream_handle), "ssmd_sgs_samples")
        torch.cuda.synchronize(sim.device)

    def plane(self, k: int) -> torch.Tensor:
        """Filtered coarse plane k as (nz, ny, nx) (0 rho, 1-3 u, 4 phi, 5-10 S, 11-16 T, 17-19 J)."""
        return self.work[k * self.nc:(k + 1) * self.nc].view(self.coarse.shape)


class Stats:
    """Running centreline and interface-spectrum statistics on a coarse grid."""

    def __init__(self, p: abi.Params, z_range: tuple[int, int], kmax: int = 32, device="cuda:0"):
        self.nx, self.ny = int(p.nx), int(p.ny)
        cx, cy = (self.nx - 1) * 0.5, (self.ny - 1) * 0.5
        self.ax = [int(np.floor(cx)), int(np.ceil(cx))]
        self.ay = [int(np.floor(cy)), int(np.ceil(cy))]
        self.z0, self.z1 = z_range
        self.kmax = kmax
        # integer mode numbers of the (ny-2) x (nx-2) interior slice, binned
        # by rounded radius; the axes may differ in length
        mx, my = self.nx - 2, self.ny - 2
        kx = torch.fft.fftfreq(mx, d=1.0 / mx, device=device)
        ky = torch.fft.fftfreq(my, d=1.0 / my, device=device)
        kr = torch.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2).round().long()
        self.bins = kr
        self.centre = torch.zeros(int(p.nz), dtype=torch.float64, device=device)
        self.spec = torch.zeros(kmax + 1, dtype=torch.float64, device=device)
        self.area = torch.zeros(int(p.nz), dtype=torch.float64, device=device)
        self.samples = 0

    def add(self, uz: torch.Tensor, phi: torch.Tensor) -> None:
        """uz, phi: (nz, ny, nx) coarse fields."""
        col = uz[:, self.ay[0]:self.ay[1] + 1, self.ax[0]:self.ax[1] + 1].double().mean(dim=(1, 2))
        self.centre += col
        ph = phi.double()
        inner = ph[:, 1:-1, 1:-1]
        gx = torch.zeros_like(inner)
        gy = torch.zeros_like(inner)
        gz = torch.zeros_like(inner)
        gx[:, :, 1:-1] = 0.5 * (inner[:, :, 2:] - inner[:, :, :-2])
        gy[:, 1:-1, :] = 0.5 * (inner[:, 2:, :] - inner[:, :-2, :])
        gz[1:-1] = 0.5 * (inner[2:] - inner[:-2])
        a = torch.sqrt(gx ** 2 + gy ** 2 + gz ** 2)
        self.area += a.sum(dim=(1, 2))
        sl = a[self.z0:self.z1]
        f = torch.fft.fft2(sl - sl.mean(dim=(1, 2), keepdim=True))
        pw = (f.real ** 2 + f.imag ** 2).mean(dim=0)
        b = self.bins.clamp(max=self.kmax + 1)
        s = torch.zeros(self.kmax + 2, dtype=torch.float64, device=pw.device).scatter_add_(0, b.reshape(-1),
                                                                                           pw.reshape(-1))
        self.spec += s[:self.kmax + 1]
        self.samples += 1

    def result(self) -> dict:
        n = max(self.samples, 1)
        return {"centreline_uz": (self.centre / n).cpu().numpy().tolist(),
                "spectrum": (self.spec / n).cpu().numpy().tolist(),
                "area_per_plane": (self.area / n).cpu().numpy().tolist(), "samples": self.samples}


def rel_l2(a, b) -> float:
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def compare(coarse: dict, filt: dict, z_range: tuple[int, int], kmax: int) -> dict:
    z0, z1 = z_range
    return {
        "centreline_rel_l2": rel_l2(coarse["centreline_uz"][z0:z1], filt["centreline_uz"][z0:z1]),
        "spectrum_rel_l2": rel_l2(coarse["spectrum"][1:kmax + 1], filt["spectrum"][1:kmax + 1]),
        "interface_area_rel": float(abs(sum(coarse["area_per_plane"][z0:z1]) / sum(filt["area_per_plane"][z0:z1]) - 1)),
    }


# -- fine runs ----------------------------------------------------------------
def run_fine(g: Geometry, jet: Jet, coarse_steps: int, sample_every: int, stat_from: int,
             snapshot_steps: list[int], z_range, kmax, on_snapshot=None, tune: bool = True) -> dict:
    """Fine run of `g.fine_steps * coarse_steps` steps; statistics of the
    filtered fields every `sample_every` coarse steps from coarse step
    `stat_from`, and on_snapshot(filter) at the listed coarse steps."""
    fp, cp = fine_params(g, jet), coarse_params(g, jet)
    r = g.fine_steps
    sim = Simulation(fp)
    sim.init()
    if tune:
        sim.step(2)
        sim.tune(reps=4)
    flt = Filter(fp, cp)
    stats = Stats(cp, z_range, kmax)
    marks = sorted(set([s for s in range(stat_from, coarse_steps + 1, sample_every)] + list(snapshot_steps)))
    done = sim.step_index // r
    t0 = time.perf_counter()
    for m in marks:
        if r * m - 1 > sim.step_index:
            sim.graph_run(r * m - 1 - sim.step_index)
        sim.step(1, macros=True)
        flt(sim)
        if m >= stat_from and (m - stat_from) % sample_every == 0:
            stats.add(flt.plane(3), flt.plane(4))
        if m in snapshot_steps and on_snapshot:
            on_snapshot(flt)
        if m - done >= 1000:
            done = m
            log(f"fine u={jet.u:.3f} sigma={jet.sigma:.3f}: coarse step {m}/{coarse_steps} "
                f"({time.perf_counter() - t0:.0f} s)")
    sim.synchronize()
    if not torch.isfinite(sim.t["uz"]).all():
        raise RuntimeError("fine run diverged")
    sim.close()
    del sim
    torch.cuda.empty_cache()
    return stats.result()


def run_coarse(g: Geometry, jet: Jet, closure: int, net: Network | None, coarse_steps: int, sample_every: int,
               stat_from: int, z_range, kmax, tune: bool = True, edit=None) -> dict:
    """Coarse run and its statistics; `edit(params)` may change the
    parameters before the run (sensitivity study)."""
    p = coarse_params(g, jet, closure, net.scales if net else None)
    if edit is not None:
        edit(p)
    sim = Simulation(p)
    if closure == abi.CLOSURE_LEARNED:
        net.load_into(sim)
    sim.init()
    if tune:
        sim.step(2)
        sim.tune(reps=4)
    stats = Stats(p, z_range, kmax)
    t0 = time.perf_counter()
    for m in range(stat_from, coarse_steps + 1, sample_every):
        if m - 1 > sim.step_index:
            sim.graph_run(m - 1 - sim.step_index)
        sim.step(1, macros=True)
        torch.cuda.synchronize(sim.device)  # the statistics read the fields on torch's stream
        stats.add(sim.t["uz"].view(p.shape), sim.t["phi"].view(p.shape))
        torch.cuda.synchronize(sim.device)  # before the next steps overwrite them
    sim.synchronize()
    wall = time.perf_counter() - t0
    ok = bool(torch.isfinite(sim.t["uz"]).all())
    res = stats.result()
    res["finite"] = ok
    if closure == abi.CLOSURE_LEARNED:
        tt = sim.t["tau_t"]
        res["tau_t_mean"] = float(tt.mean())
        res["tau_t_max"] = float(tt.max())
    res["seconds"] = wall
    sim.close()
    del sim
    torch.cuda.empty_cache()
    return res


# -- training -----------------------------------------------------------------
class SampleSet:
    """Training samples collected on the device (cells with nonzero weight,
    subsampled uniformly to at most `per_snapshot` per snapshot)."""

    def __init__(self, per_snapshot: int, seed: int = 0):
        self.per = per_snapshot
        self.g = torch.Generator(device="cuda")
        self.g.manual_seed(seed)
        self.x, self.t, self.w = [], [], []

    def add(self, flt: Filter) -> None:
        keep = torch.nonzero(flt.weight > 0).squeeze(1)
        if keep.numel() > self.per:
            keep = keep[torch.randperm(keep.numel(), device=keep.device, generator=self.g)[:self.per]]
        self.x.append(flt.feat.view(-1, 8)[keep].clone())
        self.t.append(flt.target[keep].clone())
        self.w.append(flt.weight[keep].clone())

    def tensors(self):
        return torch.cat(self.x), torch.cat(self.t), torch.cat(self.w)


def fit_scales(g: Geometry, jets: list[Jet], probe_steps: int) -> Scales:
    """Input/output scales from a short fine run of the first jet: 1/rms of
    the filtered strain rate and of |grad phi|, rms of the target tau_t
    (weighted)."""
    jet = jets[0]
    fp, cp = fine_params(g, jet), coarse_params(g, jet, abi.CLOSURE_LEARNED, Scales(1.0, 1.0, 1.0))
    sim = Simulation(fp)
    sim.init()
    sim.graph_run(g.fine_steps * probe_steps - 1)
    sim.step(1, macros=True)
    flt = Filter(fp, cp)
    flt(sim)
    w = flt.weight
    m = w > 0
    feats = flt.feat.view(-1, 8)[m].view(torch.bfloat16).float()
    s_rms = float(torch.sqrt((feats[:, :6] ** 2).mean()))
    g_rms = float(torch.sqrt((feats[:, 7] ** 2).mean()))
    t = flt.target[m]
    t_rms = float(torch.sqrt((w[m] * t ** 2).sum() / w[m].sum()))
    sim.close()
    del sim
    torch.cuda.empty_cache()
    return Scales(strain_scale=1.0 / max(s_rms, 1e-12), grad_scale=1.0 / max(g_rms, 1e-12),
                  tau_scale=max(t_rms, 1e-6))


def train_main(args) -> int:
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    g = Geometry(nx=args.nx, ny=args.ny, nz=args.nz, q=args.q, tau=args.tau, scaling=args.scaling)
    rng = np.random.default_rng(args.seed)
    jets = [Jet(float(rng.uniform(*args.u_range)), float(rng.uniform(*args.sigma_range)), int(args.seed * 100 + j))
            for j in range(args.runs)]
    log(f"training jets: {[asdict(j) for j in jets]}")
    scales = fit_scales(g, jets, max(args.steps // 4, 200))
    log(f"scales: {asdict(scales)}")
    samples = SampleSet(args.per_snapshot, seed=args.seed)
    snaps = [int(s) for s in np.linspace(args.steps // 2, args.steps, args.snapshots)]
    for jet in jets:
        run_fine_scaled(g, jet, snaps, scales, samples.add)
    x, t, w = samples.tensors()
    w = w / w.mean()
    n = x.shape[0]
    log(f"{n} samples; target mean {float((w * t).sum() / w.sum()):.4f}")
    # hold out a tenth for validation
    perm = torch.randperm(n, device=x.device, generator=torch.Generator(device="cuda").manual_seed(1))
    nv = n // 10
    va, tr = perm[:nv], perm[nv:]
    net = Network(seed=args.seed, scales=scales)
    zero_loss = float((w[va] * t[va] ** 2).mean())
    losses = net.train(x[tr], t[tr], steps=args.train_steps, batch=args.batch, lr=args.lr, lr_final=args.lr / 30,
                       weight=w[tr], log_every=max(args.train_steps // 20, 1))
    net.gradient(x[va], t[va], w[va])
    val = float(net.loss.item())
    # Smagorinsky's eddy relaxation time from the same inputs (strain columns)
    s = x[va].view(torch.bfloat16).float()[:, :6] / scales.strain_scale
    ss = (s[:, :3] ** 2).sum(1) + 2 * (s[:, 3:] ** 2).sum(1)
    tau_smag = 3.0 * g.cs ** 2 * torch.sqrt(2 * ss) / scales.tau_scale
    smag_loss = float((w[va] * (tau_smag - t[va]) ** 2).mean())
    net.save(out / "closure")
    rep = {"jets": [asdict(j) for j in jets], "scales": asdict(scales), "samples": n, "train_losses": losses,
           "validation": {"learned": val, "no_closure": zero_loss, "smagorinsky": smag_loss}, "geometry": asdict(g),
           "snapshots": snaps}
    (out / "train_report.json").write_text(json.dumps(rep, indent=1))
    log(f"validation weighted MSE: learned {val:.4g}, no closure {zero_loss:.4g}, Smagorinsky {smag_loss:.4g}")
    return 0


def run_fine_scaled(g: Geometry, jet: Jet, snaps: list[int], scales: Scales, on_snap) -> None:
    """Fine run; at each listed coarse step the filter forms samples at the
    training scales (its coarse parameters carry them) and on_snap gets it."""
    fp = fine_params(g, jet)
    cp = coarse_params(g, jet, abi.CLOSURE_LEARNED, scales)
    r = g.fine_steps
    sim = Simulation(fp)
    sim.init()
    sim.step(2)
    sim.tune(reps=4)
    flt = Filter(fp, cp)
    t0 = time.perf_counter()
    for m in snaps:
        if r * m - 1 > sim.step_index:
            sim.graph_run(r * m - 1 - sim.step_index)
        sim.step(1, macros=True)
        flt(sim)
        on_snap(flt)
        log(f"fine u={jet.u:.3f} sigma={jet.sigma:.3f}: snapshot at coarse step {m} "
            f"({time.perf_counter() - t0:.0f} s)")
    sim.synchronize()
    if not torch.isfinite(sim.t["uz"]).all():
        raise RuntimeError("fine run diverged")
    sim.close()
    del sim
    torch.cuda.empty_cache()


def eval_main(args) -> int:
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    g = Geometry(nx=args.nx, ny=args.ny, nz=args.nz, q=args.q, tau=args.tau, scaling=args.scaling)
    jet = Jet(args.u, args.sigma, args.seed)
    net = Network.load(args.net)
    z_range = (1, args.nz // 2)
    kmax = args.kmax
    stat_from = args.steps // 2
    log(f"eval jet {asdict(jet)}; statistics every {args.every} coarse steps from {stat_from}")
    res = {"jet": asdict(jet), "geometry": asdict(g), "steps": args.steps, "stat_from": stat_from,
           "every": args.every, "z_range": z_range, "kmax": kmax}
    for name, closure in (("none", abi.CLOSURE_NONE), ("smagorinsky", abi.CLOSURE_SMAGORINSKY),
                          ("learned", abi.CLOSURE_LEARNED)):
        r = run_coarse(g, jet, closure, net, args.steps, args.every, stat_from, z_range, kmax)
        res[name] = r
        log(f"coarse {name}: finite {r['finite']} ({r['seconds']:.1f} s)")
    res["filtered_fine"] = run_fine(g, jet, args.steps, args.every, stat_from, [], z_range, kmax)
    res["errors"] = {k: compare(res[k], res["filtered_fine"], z_range, kmax)
                     for k in ("none", "smagorinsky", "learned")}
    fine_cells = (2 * g.nx) * (2 * g.ny) * (2 * g.nz) * g.fine_steps
    res["cell_update_ratio"] = fine_cells / (g.nx * g.ny * g.nz)
    (out / "eval_report.json").write_text(json.dumps(res))
    for k, v in res["errors"].items():
        log(f"{k:12s} centreline {v['centreline_rel_l2']:.4f}  spectrum {v['spectrum_rel_l2']:.4f}  "
            f"interface area {v['interface_area_rel']:.4f}")
    return 0


def noise_main(args) -> int:
    """The comparison's noise floor: the coarse no-closure jet of the
    evaluation run with `--seeds` different inlet-noise seeds, and the errors
    between every pair of them (same statistics, same window)."""
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    g = Geometry(nx=args.nx, ny=args.ny, nz=args.nz, q=args.q, tau=args.tau, scaling=args.scaling)
    z_range, kmax, stat_from = (1, args.nz // 2), args.kmax, args.steps // 2
    runs = []
    for k in range(args.seeds):
        jet = Jet(args.u, args.sigma, args.seed + 1000 * k)
        runs.append(run_coarse(g, jet, abi.CLOSURE_NONE, None, args.steps, args.every, stat_from, z_range, kmax))
        log(f"coarse none, seed {jet.seed}: finite {runs[-1]['finite']} ({runs[-1]['seconds']:.1f} s)")
    pairs = {f"{i}-{j}": compare(runs[i], runs[j], z_range, kmax)
             for i in range(len(runs)) for j in range(i + 1, len(runs))}
    keys = ("centreline_rel_l2", "spectrum_rel_l2", "interface_area_rel")
    summary = {k: {"mean": float(np.mean([v[k] for v in pairs.values()])),
                   "max": float(np.max([v[k] for v in pairs.values()]))} for k in keys}
    rep = {"u": args.u, "sigma": args.sigma, "seeds": [args.seed + 1000 * k for k in range(args.seeds)],
           "steps": args.steps, "stat_from": stat_from, "every": args.every, "geometry": asdict(g),
           "pairs": pairs, "summary": summary}
    (out / "noise_report.json").write_text(json.dumps(rep, indent=1))
    for k, v in summary.items():
        log(f"seed-to-seed {k}: mean {v['mean']:.4f}, max {v['max']:.4f}")
    return 0


def sensitivity_main(args) -> int:
    """Which coarse-grid parameter the a-posteriori errors respond to: the
    no-closure coarse jet of an evaluation report with its viscosity
    (tau - 1/2), surface tension or sharpening coefficient scaled by each
    factor, scored against that report's filtered fine statistics."""
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    ref = json.loads(Path(args.report).read_text())
    geo = dict(ref["geometry"])
    g = Geometry(**geo)
    jet = Jet(**ref["jet"])
    z_range, kmax, stat_from = tuple(ref["z_range"]), ref["kmax"], ref["stat_from"]
    fine = ref["filtered_fine"]
    tau0 = g.tau

    def scaled(name, f):
        def edit(p):
            if name == "viscosity":
                p.visc_water = p.visc_oil = viscosity_from_tau(0.5 + f * (tau0 - 0.5))
            elif name == "sigma":
                p.sigma = f * jet.sigma
            elif name == "gamma":
                p.gamma = f * p.gamma
        return edit

    res = {"report": str(args.report), "factors": args.factors, "runs": {}}
    for name in ("viscosity", "sigma", "gamma"):
        for f in args.factors:
            r = run_coarse(g, jet, abi.CLOSURE_NONE, None, ref["steps"], ref["every"], stat_from, z_range, kmax,
                           edit=scaled(name, f))
            e = compare(r, fine, z_range, kmax) if r["finite"] else None
            res["runs"][f"{name} x{f}"] = e
            log(f"{name} x{f}: {e}")
    (out / "sensitivity_report.json").write_text(json.dumps(res, indent=1))
    return 0


def interface_network(kappa: float, beta: float) -> np.ndarray:
    """Parameters of the network y = kappa * x7 + beta, x7 = |grad phi-hat| *
    grad_scale >= 0: hidden unit 0 of each layer carries x7 through the ReLUs
    (weights 1, exact in bf16), the output layer scales and offsets it."""
    p = np.zeros(abi.MLP_PARAMS, dtype=np.float32)
    o = offsets()
    p[o["W1"][0] + 0 * abi.MLP_INPUTS + 7] = 1.0
    p[o["W2"][0] + 0 * abi.MLP_HIDDEN + 0] = 1.0
    p[o["W3"][0] + 0 * abi.MLP_HIDDEN + 0] = 1.0
    p[o["W4"][0] + 0] = kappa
    p[o["b4"][0]] = beta
    return p


def calibrate_main(args) -> int:
    """A-posteriori calibration of a signed, interface-localised relaxation
    time tau_t = max(floor, 1e-3 (kappa x7 + beta)): Nelder-Mead on (kappa,
    beta) minimising the largest of the three statistic errors of a coarse
    LEARNED run of a training jet against that jet's filtered fine run; the
    result is scored on the held-out jet of an evaluation report.

    With --capillary a parameter s scales the coarse run's surface tension
    (sigma_coarse = s sigma), with --sharpening a parameter h its D3Q7
    sharpening coefficient (gamma_coarse = h gamma): coarse-grid closures of
    the capillary force and of the phase-field flux, the levers the interface
    statistics answer to and tau_t cannot reach.  With --flux a parameter f
    is the flux_coeff of the interface-localised tau_t (D_t = f tau_t / 3): a
    phase-field flux closure local to the interface."""
    from scipy.optimize import minimize

    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    ref = json.loads(Path(args.report).read_text())
    g = Geometry(**dict(ref["geometry"]))
    held = Jet(**ref["jet"])
    z_range, kmax, stat_from, steps, every = tuple(ref["z_range"]), ref["kmax"], ref["stat_from"], ref["steps"], ref["every"]
    train = Jet(args.train_u, args.train_sigma, args.seed)
    floor = -0.9 * 3.0 * viscosity_from_tau(g.tau) * 0.99
    log(f"training jet {asdict(train)}: filtered fine statistics")
    fine_train = run_fine(g, train, steps, every, stat_from, [], z_range, kmax)
    history = []

    def errors(jet, fine, kappa, beta, s=1.0, h=1.0, f=0.0):
        net = Network(interface_network(kappa, beta),
                      scales=Scales(strain_scale=1.0, grad_scale=args.grad_scale, tau_scale=1e-3, tau_floor=floor))

        def edit(p):
            p.sigma = p.sigma * s
            p.gamma = p.gamma * h
            p.flux_coeff = f

        r = run_coarse(g, jet, abi.CLOSURE_LEARNED, net, steps, every, stat_from, z_range, kmax, tune=False,
                       edit=edit if (s, h, f) != (1.0, 1.0, 0.0) else None)
        e = compare(r, fine, z_range, kmax) if r["finite"] else None
        return e, r

    # parameter vector: kappa, beta, then s (--capillary), h (--sharpening), f (--flux)
    names = (["kappa", "beta"] + (["s"] if args.capillary else []) + (["h"] if args.sharpening else []) +
             (["f"] if args.flux else []))

    def unpack(v):
        d = dict(zip(names, (float(a) for a in v)))
        return d["kappa"], d["beta"], d.get("s", 1.0), d.get("h", 1.0), d.get("f", 0.0)

    def objective(v):
        kappa, beta, s, h, f = unpack(v)
        if not (0.5 <= s <= 2.0 and 0.25 <= h <= 2.0 and abs(f) <= 100.0):  # keep to stable, studied ranges
            return 10.0
        e, r = errors(train, fine_train, kappa, beta, s, h, f)
        j = max(e.values()) if e else 10.0
        history.append({"kappa": kappa, "beta": beta, "sigma_scale": s, "gamma_scale": h, "flux_coeff": f,
                        "errors": e, "objective": j, "tau_t_mean": r.get("tau_t_mean"),
                        "tau_t_max": r.get("tau_t_max")})
        log(f"kappa {kappa:8.3f} beta {beta:8.3f} s {s:6.3f} h {h:6.3f} f {f:7.3f}: max error {j:.4f} {e}")
        return j

    x0 = ([args.kappa0, args.beta0] + ([1.0] if args.capillary else []) + ([args.h0] if args.sharpening else []) +
          ([args.flux0] if args.flux else []))
    steps0 = {"kappa": args.step, "beta": args.step, "s": args.sigma_step, "h": args.gamma_step, "f": args.flux_step}
    simplex = [list(x0)]
    for i, nm in enumerate(names):
        v = list(x0)
        v[i] += steps0[nm]
        simplex.append(v)
    res = minimize(objective, x0=np.array(x0), method="Nelder-Mead",
                   options={"maxfev": args.evals, "initial_simplex": np.array(simplex)})
    kappa, beta, s_best, h_best, f_best = unpack(res.x)
    held_fine = ref["filtered_fine"]
    e_best, _ = errors(held, held_fine, kappa, beta, s_best, h_best, f_best)
    e_none = ref["errors"]["none"]
    rep = {"train_jet": asdict(train), "held_out_jet": asdict(held), "floor": floor, "grad_scale": args.grad_scale,
           "history": history,
           "best": {"kappa": kappa, "beta": beta, "sigma_scale": s_best, "gamma_scale": h_best,
                    "flux_coeff": f_best, "train_objective": float(res.fun)},
           "held_out": {"calibrated": e_best, "none": e_none}}
    (out / "calibrate_report.json").write_text(json.dumps(rep, indent=1))
    log(f"best kappa {kappa:.3f} beta {beta:.3f} sigma scale {s_best:.3f} gamma scale {h_best:.3f} "
        f"flux_coeff {f_best:.3f}; "
        f"held-out calibrated {e_best}; none {e_none}")
    return 0


# -- phase-field flux closure -------------------------------------------------
class FluxFit:
    """A-priori least squares of the subgrid flux of the phase field, J =
    <phi u> - <phi><u> (filter planes 17-19), against gradient diffusion
    -D grad(phi-bar) (central differences on the coarse grid, over cells two
    or more layers inside): D = flux_coeff * nu_t per network (nu_t = tau_t / 3
    on the filter's inputs; the flux_coeff closure) and D = const.  Two tensor
    models are fitted beside them, J = C M with
      gradient  M_a = sum_b d_b phi-bar d_b u-bar_a (the leading term of the box
                filter's covariance for fields linear over a block: C = 1/16
                for the 2x filter in coarse cells);
      strain    M_a = sum_b S-bar_ab d_b phi-bar (the part of it a collision can
                form from its own non-equilibrium moments).
    Sums in FP64 over every snapshot added."""

    MODELS = ("gradient", "strain")

    def __init__(self, nets: dict[str, Network]):
        self.nets = nets
        self.s = {k: 0.0 for k in ("jj", "jg", "gg")}
        for m in self.MODELS:
            self.s[f"jm_{m}"] = 0.0
            self.s[f"mm_{m}"] = 0.0
        for n in nets:
            self.s[f"jg_{n}"] = 0.0
            self.s[f"gg_{n}"] = 0.0
        self.snapshots = 0

    def add(self, flts: dict[str, Filter]) -> None:
        flt = next(iter(flts.values()))
        phi = flt.plane(4).double()
        core = (slice(2, -2),) * 3
        gx = 0.5 * (phi[2:-2, 2:-2, 3:-1] - phi[2:-2, 2:-2, 1:-3])
        gy = 0.5 * (phi[2:-2, 3:-1, 2:-2] - phi[2:-2, 1:-3, 2:-2])
        gz = 0.5 * (phi[3:-1, 2:-2, 2:-2] - phi[1:-3, 2:-2, 2:-2])
        jx, jy, jz = (flt.plane(17 + a).double()[core] for a in range(3))
        mjg = -(jx * gx + jy * gy + jz * gz)  # -J . grad phi
        gg = gx * gx + gy * gy + gz * gz
        self.s["jj"] += float((jx * jx + jy * jy + jz * jz).sum())
        self.s["jg"] += float(mjg.sum())
        self.s["gg"] += float(gg.sum())

        def grad(f):
            return (0.5 * (f[2:-2, 2:-2, 3:-1] - f[2:-2, 2:-2, 1:-3]),
                    0.5 * (f[2:-2, 3:-1, 2:-2] - f[2:-2, 1:-3, 2:-2]),
                    0.5 * (f[3:-1, 2:-2, 2:-2] - f[1:-3, 2:-2, 2:-2]))
        g = (gx, gy, gz)
        du = [grad(flt.plane(1 + a).double()) for a in range(3)]  # du[a][b] = d_b u_a
        grad_m = [sum(g[b] * du[a][b] for b in range(3)) for a in range(3)]
        S = [flt.plane(5 + k).double()[core] for k in range(6)]  # xx yy zz xy xz yz
        Sm = ((S[0], S[3], S[4]), (S[3], S[1], S[5]), (S[4], S[5], S[2]))
        strain_m = [sum(Sm[a][b] * g[b] for b in range(3)) for a in range(3)]
        for name, M in (("gradient", grad_m), ("strain", strain_m)):
            self.s[f"jm_{name}"] += float((jx * M[0] + jy * M[1] + jz * M[2]).sum())
            self.s[f"mm_{name}"] += float((M[0] * M[0] + M[1] * M[1] + M[2] * M[2]).sum())
        for name, net in self.nets.items():
            f = flts[name]
            y = net.forward(f.feat.view(-1, 8)).double().view(f.coarse.shape)
            nu = torch.clamp(net.scales.tau_scale * y, min=net.scales.tau_floor)[core] / 3.0
            self.s[f"jg_{name}"] += float((nu * mjg).sum())
            self.s[f"gg_{name}"] += float((nu * nu * gg).sum())
        self.snapshots += 1

    def result(self) -> dict:
        s = self.s
        out = {"snapshots": self.snapshots, "flux_sq_sum": s["jj"],
               "constant": {"D": s["jg"] / s["gg"], "explained": s["jg"] ** 2 / (s["gg"] * s["jj"])}}
        for n in self.nets:
            a, b = s[f"jg_{n}"], s[f"gg_{n}"]
            out[n] = {"flux_coeff": a / b if b > 0 else 0.0, "explained": a * a / (b * s["jj"]) if b > 0 else 0.0}
        for m in self.MODELS:
            a, b = s[f"jm_{m}"], s[f"mm_{m}"]
            if b > 0:
                out[m] = {"C": a / b, "explained": a * a / (b * s["jj"])}
        return out


def flux_main(args) -> int:
    """The phase-field flux closure: flux_coeff fitted a priori (FluxFit) on
    the filtered fine runs of `--runs` training jets, for the trained network
    and for the calibrated interface network; then the held-out jet of an
    evaluation report run coarse with each network at flux_coeff 0 and at the
    fitted value, scored against that report's filtered fine statistics."""
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    ref = json.loads(Path(args.report).read_text())
    g = Geometry(**dict(ref["geometry"]))
    held = Jet(**ref["jet"])
    z_range, kmax, stat_from, steps, every = tuple(ref["z_range"]), ref["kmax"], ref["stat_from"], ref["steps"], ref["every"]
    nets = {"trained": Network.load(args.net), "interface": Network.load(args.interface_net)}
    rng = np.random.default_rng(args.seed)
    jets = [Jet(float(rng.uniform(*args.u_range)), float(rng.uniform(*args.sigma_range)), int(args.seed * 100 + j))
            for j in range(args.runs)]
    log(f"flux fit on jets {[asdict(j) for j in jets]}")
    fit = FluxFit(nets)
    snaps = [int(v) for v in np.linspace(steps // 2, steps, args.snapshots)]
    per_jet = []
    for jet in jets:
        fp = fine_params(g, jet)
        flts = {n: Filter(fp, coarse_params(g, jet, abi.CLOSURE_LEARNED, net.scales)) for n, net in nets.items()}
        one = FluxFit(nets)
        r = g.fine_steps
        sim = Simulation(fp)
        sim.init()
        sim.step(2)
        sim.tune(reps=4)
        t0 = time.perf_counter()
        for m in snaps:
            if r * m - 1 > sim.step_index:
                sim.graph_run(r * m - 1 - sim.step_index)
            sim.step(1, macros=True)
            for f in flts.values():
                f(sim)
            fit.add(flts)
            one.add(flts)
            log(f"fine u={jet.u:.3f} sigma={jet.sigma:.3f}: snapshot at coarse step {m} "
                f"({time.perf_counter() - t0:.0f} s)")
        sim.synchronize()
        if not torch.isfinite(sim.t["uz"]).all():
            raise RuntimeError("fine run diverged")
        sim.close()
        del sim, flts
        torch.cuda.empty_cache()
        per_jet.append({"jet": asdict(jet), "fit": one.result()})
        log(f"jet fit: {per_jet[-1]['fit']}")
    apriori = fit.result()
    log(f"a-priori fit over {len(jets)} jets: {apriori}")
    post = {}
    for name, net in nets.items():
        for coeff in (0.0, apriori[name]["flux_coeff"]):
            def edit(p, c=coeff):
                p.flux_coeff = c
            res = run_coarse(g, held, abi.CLOSURE_LEARNED, net, steps, every, stat_from, z_range, kmax, tune=False,
                             edit=edit)
            e = compare(res, ref["filtered_fine"], z_range, kmax) if res["finite"] else None
            post[f"{name}_flux{coeff:.4g}"] = {"network": name, "flux_coeff": coeff, "errors": e,
                                                "finite": res["finite"], "tau_t_mean": res.get("tau_t_mean")}
            log(f"held out, {name} network, flux_coeff {coeff:.4g}: {e}")
    rep = {"train_jets": [asdict(j) for j in jets], "snapshots": snaps, "apriori": apriori, "per_jet": per_jet,
           "held_out_jet": asdict(held), "a_posteriori": post, "none": ref["errors"]["none"]}
    (out / "flux_report.json").write_text(json.dumps(rep, indent=1))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    for name in ("train", "eval", "noise", "sensitivity", "calibrate", "flux"):
        s = sub.add_parser(name)
        s.add_argument("--out", required=True)
        s.add_argument("--nx", type=int, default=128)
        s.add_argument("--ny", type=int, default=128)
        s.add_argument("--nz", type=int, default=512)
        s.add_argument("--q", type=int, default=27)
        s.add_argument("--steps", type=int, default=10000, help="coarse steps (the fine runs take 4 or 2 per step)")
        s.add_argument("--scaling", choices=("diffusive", "acoustic"), default=Geometry.scaling,
                       help="time scaling of the fine twin")
        s.add_argument("--seed", type=int, default=5)
        s.add_argument("--tau", type=float, default=Geometry.tau, help="coarse base relaxation time")
    t = sub.choices["train"]
    t.add_argument("--runs", type=int, default=2)
    t.add_argument("--u-range", type=float, nargs=2, default=(0.03, 0.08), help="inlet u_z of the training jets")
    t.add_argument("--sigma-range", type=float, nargs=2, default=(0.01, 0.03), help="sigma of the training jets")
    t.add_argument("--snapshots", type=int, default=6)
    t.add_argument("--per-snapshot", type=int, default=1 << 20)
    t.add_argument("--train-steps", type=int, default=3000)
    t.add_argument("--batch", type=int, default=1 << 16)
    t.add_argument("--lr", type=float, default=3e-3)
    e = sub.choices["eval"]
    e.add_argument("--net", required=True)
    e.add_argument("--u", type=float, default=0.05)
    e.add_argument("--sigma", type=float, default=0.02)
    e.add_argument("--every", type=int, default=50)
    e.add_argument("--kmax", type=int, default=32)
    n = sub.choices["noise"]
    n.add_argument("--u", type=float, default=0.05)
    n.add_argument("--sigma", type=float, default=0.02)
    n.add_argument("--every", type=int, default=50)
    n.add_argument("--kmax", type=int, default=32)
    n.add_argument("--seeds", type=int, default=4)
    sn = sub.choices["sensitivity"]
    sn.add_argument("--report", required=True, help="an eval_report.json (its geometry, jet and fine statistics)")
    sn.add_argument("--factors", type=float, nargs="+", default=[0.5, 0.8, 1.25, 2.0])
    ca = sub.choices["calibrate"]
    ca.add_argument("--report", required=True, help="the held-out jet's eval_report.json")
    ca.add_argument("--train-u", type=float, default=0.06)
    ca.add_argument("--train-sigma", type=float, default=0.015)
    ca.add_argument("--grad-scale", type=float, default=4.0)
    ca.add_argument("--step", type=float, default=10.0, help="initial simplex step (units of 1e-3 in tau)")
    ca.add_argument("--evals", type=int, default=30)
    ca.add_argument("--kappa0", type=float, default=0.0, help="starting point")
    ca.add_argument("--beta0", type=float, default=0.0, help="starting point")
    ca.add_argument("--capillary", action="store_true",
                    help="also fit s, the coarse surface-tension scale (sigma_coarse = s sigma)")
    ca.add_argument("--sigma-step", type=float, default=0.25, help="initial simplex step of s")
    ca.add_argument("--sharpening", action="store_true",
                    help="also fit h, the coarse D3Q7 sharpening scale (gamma_coarse = h gamma)")
    ca.add_argument("--gamma-step", type=float, default=-0.25, help="initial simplex step of h")
    ca.add_argument("--h0", type=float, default=1.0, help="starting h (--sharpening)")
    ca.add_argument("--flux", action="store_true",
                    help="also fit f, the phase-field flux closure's flux_coeff (D_t = f tau_t / 3)")
    ca.add_argument("--flux0", type=float, default=0.0, help="starting f (--flux)")
    ca.add_argument("--flux-step", type=float, default=3.0, help="initial simplex step of f")
    fx = sub.choices["flux"]
    fx.add_argument("--report", required=True, help="the held-out jet's eval_report.json")
    fx.add_argument("--net", default=str(Path(__file__).resolve().parent / "data" / "closure_jet_d3q27"))
    fx.add_argument("--interface-net",
                    default=str(Path(__file__).resolve().parent / "data" / "closure_jet_d3q27_interface"))
    fx.add_argument("--runs", type=int, default=3)
    fx.add_argument("--u-range", type=float, nargs=2, default=(0.03, 0.08))
    fx.add_argument("--sigma-range", type=float, nargs=2, default=(0.01, 0.03))
    fx.add_argument("--snapshots", type=int, default=6)
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("the closure study runs on the GPU", file=sys.stderr)
        return 1
    return {"train": train_main, "eval": eval_main, "noise": noise_main, "sensitivity": sensitivity_main,
            "calibrate": calibrate_main, "flux": flux_main}[args.cmd](args)


if __name__ == "__main__":
    sys.exit(main())
