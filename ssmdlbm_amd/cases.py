"""Parameter blocks for the solver: the reference's compile-time case and
the jet workloads the hot path is measured and checked on.

The reference fixes everything at compile time (src/constants.cuh:70-101,
src/functions/globalFunctions.cuh:170-283); here the same quantities are a
runtime ``ssmd_params`` block with the same names and formulas.
"""
from __future__ import annotations

from dataclasses import dataclass

from . import abi

# D3Q19 / D3Q27 direction tables (reference velocitySet/D3Q19.cuh:82-131,
# D3Q27.cuh:92-140); D3Q7 is the first seven entries.
DIRECTIONS = (
    (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, -1), (0, 1, 1), (0, -1, -1), (1, -1, 0),
    (-1, 1, 0), (1, 0, -1), (-1, 0, 1), (0, 1, -1), (0, -1, 1), (1, 1, 1), (-1, -1, -1),
    (1, 1, -1), (-1, -1, 1), (1, -1, 1), (-1, 1, -1), (-1, 1, 1), (1, -1, -1),
)


def viscosity_from_tau(tau: float) -> float:
    """nu = cs2 (tau - 1/2) for the hydrodynamic lattices (cs2 = 1/3)."""
    return (tau - 0.5) / 3.0


def gamma_from_tau_g(tau_g: float, width: float) -> float:
    """physics::gamma of constants.cuh:95-100: kappa / cs2 with
    kappa = 4 cs2 (tau_g - 1/2) / width, i.e. 4 (tau_g - 1/2) / width."""
    cs2 = 0.25
    diff_int = cs2 * (tau_g - 0.5)
    kappa = 4.0 * diff_int / width
    return kappa / cs2


def reference_default(q: int = 19) -> abi.Params:
    """The reference's own case: 128x256x256, water jet on y=0, oil jet on z=0."""
    p = abi.Params()
    abi.check(abi.load().ssmd_params_reference_default(p, q), "ssmd_params_reference_default")
    return p


@dataclass(frozen=True)
class JetCase:
    """Density-matched single jet entering through the bottom face (z = 0).

    Ambient phi = 0, u = 0; a circular inlet of diameter `diameter` cells at
    the face centre injects phi = 1 at speed `u` with a seeded uniform
    velocity perturbation of relative amplitude `perturb`; extrapolation
    outflow at the top; periodic sides.
    """

    name: str
    nx: int
    ny: int
    nz: int
    q: int = 19
    tau: float = 0.55
    tau_g: float = 0.8
    width: float = 4.0
    sigma: float = 0.1
    diameter: int = 8
    u: float = 0.05
    perturb: float = 0.01
    seed: int = 1
    pop_format: int = abi.POP_F32
    wrap: int = abi.WRAP_REFERENCE

    def params(self) -> abi.Params:
        p = abi.Params()
        p.nx, p.ny, p.nz = self.nx, self.ny, self.nz
        p.q = self.q
        p.pop_format = self.pop_format
        nu = viscosity_from_tau(self.tau)
        p.visc_water = nu
        p.visc_oil = nu
        p.sigma = self.sigma
        p.gamma = gamma_from_tau_g(self.tau_g, self.width)
        p.perturb = self.perturb
        p.seed = self.seed
        p.water.enabled = 0
        p.oil.enabled = 1
        p.oil.center_a = (self.nx - 1) * 0.5
        p.oil.center_b = (self.ny - 1) * 0.5
        r = self.diameter // 2  # integer radius as mesh::radius_* (constants.cuh:80-81)
        p.oil.r2 = float(r * r)
        p.oil.diameter = self.diameter
        p.oil.u = self.u
        p.oil.phi = 1.0
        p.outflow_y = 0
        p.outflow_z = 1
        p.wrap_x = self.wrap
        p.wrap_y = self.wrap
        p.wrap_z = abi.WRAP_NONE
        return p

    def with_(self, **kw) -> "JetCase":
        d = dict(self.__dict__)
        d.update(kw)
        return JetCase(**d)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


# The workloads: a CPU-runnable parity case, the D3Q27 throughput case, and
# the FP16-storage case (with its FP32 twin for the storage-speedup ratio).
JET_D3Q19_SMALL = JetCase("jet_d3q19_64x64x128", 64, 64, 128, q=19, sigma=0.1, diameter=8, u=0.05, seed=1)
JET_D3Q27 = JetCase("jet_d3q27_128x128x512", 128, 128, 512, q=27, sigma=0.05, diameter=16, u=0.05, seed=2)
JET_D3Q19_FP16 = JetCase("jet_d3q19_fp16_256x256x1024", 256, 256, 1024, q=19, sigma=0.05, diameter=24,
                         u=0.04, seed=3, pop_format=abi.POP_F16)
JET_D3Q19_FP32_LARGE = JET_D3Q19_FP16.with_(name="jet_d3q19_fp32_256x256x1024", pop_format=abi.POP_F32)
JET_D3Q19_BF16 = JET_D3Q19_FP16.with_(name="jet_d3q19_bf16_256x256x1024", pop_format=abi.POP_BF16)

WORKLOADS = {c.name: c for c in (JET_D3Q19_SMALL, JET_D3Q27, JET_D3Q19_FP16, JET_D3Q19_FP32_LARGE, JET_D3Q19_BF16)}


def closed_box(n: int, q: int = 19, tau: float = 0.8, sigma: float = 0.01, gamma: float = 0.3,
               pop_format: int = abi.POP_F32) -> abi.Params:
    """Truly periodic cube without inlets: the mass-conservation case."""
    p = abi.Params()
    p.nx = p.ny = p.nz = n
    p.q = q
    p.pop_format = pop_format
    p.visc_water = p.visc_oil = viscosity_from_tau(tau)
    p.sigma = sigma
    p.gamma = gamma
    p.wrap_x = p.wrap_y = p.wrap_z = abi.WRAP_TRUE
    return p


def algorithmic_bytes_per_cell(q: int, pop_format: int) -> int:
    """Bytes one cell update of the reference's schedule moves (the whole-step
    yardstick): f read+write, seven g populations read+write, phi once
    written and once read = 2 Q s_f + 2*7*4 + 2*4."""
    return 2 * q * abi.POP_BYTES[pop_format] + 2 * 7 * 4 + 2 * 4


def step_bytes_per_cell(q: int, pop_format: int) -> int:
    """Bytes one cell update of this library's step moves: f read+write, the
    four g moment planes written (lattice kernel) and read (phase kernel), phi
    written and read = 2 Q s_f + 2*4*4 + 2*4."""
    return 2 * q * abi.POP_BYTES[pop_format] + 2 * 4 * 4 + 2 * 4


def lattice_kernel_bytes_per_cell(q: int, pop_format: int) -> int:
    """Algorithmic bytes of the fused lattice kernel alone: f read + f write,
    g moments (4 floats) written, phi read."""
    return 2 * q * abi.POP_BYTES[pop_format] + 4 * 4 + 4


def ensemble(n_jets: int = 32, seed: int = 4, q: int = 19, pop_format: int = abi.POP_F32) -> list[JetCase]:
    """Seeded ensemble of single jets with uneven boxes and parameters.

    Boxes: nx = ny uniform in {48, 56, ..., 128}, nz uniform in {96, 104, ..., 384};
    inlet u_z ~ U[0.02, 0.08], sigma ~ U[0.01, 0.1], diameter ~ U{6..16};
    the inlet-noise seed of jet j is seed * 1000 + j.
    """
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for j in range(n_jets):
        nxy = int(rng.integers(6, 17)) * 8
        nz = int(rng.integers(12, 49)) * 8
        u = float(rng.uniform(0.02, 0.08))
        sigma = float(rng.uniform(0.01, 0.1))
        d = int(rng.integers(6, 17))
        out.append(JetCase(f"ens{seed}_{j:02d}", nxy, nxy, nz, q=q, sigma=sigma, diameter=d, u=u,
                           seed=seed * 1000 + j, pop_format=pop_format))
    return out
