"""Host-side driver of the gfx950 hot path, mirroring the reference's
main.cu flow (allocate -> init -> per-step graph + boundaries -> output).

Device memory comes from torch (so fields are tensors that tests and tools
can read); every kernel runs inside libssmd.so through the C ABI on the
torch stream of the device.  There is no CPU or torch compute path.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Dict, Iterable, Optional

import numpy as np
import torch

from . import abi
from .cases import DIRECTIONS

_POP_DTYPE = {abi.POP_F32: torch.float32, abi.POP_F16: torch.float16, abi.POP_BF16: torch.bfloat16}


class Simulation:
    """One field set on one GPU (reference: the global LBMFields + FieldAllocate)."""

    def __init__(self, params: abi.Params, device: str | torch.device = "cuda:0",
                 stream: torch.cuda.Stream | None = None):
        """`stream`: run on this (non-null) stream instead of one of our own;
        simulations sharing a stream execute strictly one after another."""
        self.lib = abi.load()
        abi.check(self.lib.ssmd_validate(C.byref(params)), "ssmd_validate")
        if not torch.cuda.is_available():
            raise abi.SsmdError("no GPU: the hot path runs only on the device")
        self.p = params.copy()
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        # a stream of our own (default priority): graph capture cannot use the null stream
        self.stream = stream if stream is not None else torch.cuda.Stream(self.device)
        n = self.p.cells
        q = int(self.p.q)
        z = lambda *shape, dtype=torch.float32: torch.zeros(*shape, dtype=dtype, device=self.device)
        self.t: Dict[str, torch.Tensor] = {}
        for name in abi.SCALAR_FIELDS:
            if name in ("normx", "normy", "normz", "ind"):
                continue  # never allocated, as in the reference (main.cu:55-70)
            self.t[name] = z(n)
        pdt = _POP_DTYPE[int(self.p.pop_format)]
        # one g buffer for both g and g_next: its one reader, the phase update,
        # runs before the lattice kernel of the same step rewrites it
        g = z(4 * n)
        self._pops = {"fa": z(q * n, dtype=pdt), "fb": z(q * n, dtype=pdt), "ga": g, "gb": g}
        self.t["step_counter"] = z(1, dtype=torch.int64)
        self.t["reduce"] = z(abi.REDUCE_DOUBLES, dtype=torch.float64)
        self.d = abi.Fields()
        for name, ten in self.t.items():
            setattr(self.d, name, ten.data_ptr())
        self.d.f = self._pops["fa"].data_ptr()
        self.d.f_next = self._pops["fb"].data_ptr()
        self.d.g = self._pops["ga"].data_ptr()
        self.d.g_next = self._pops["gb"].data_ptr()
        self.d.owned = 0
        if int(self.p.closure) == abi.CLOSURE_LEARNED:
            # eddy relaxation time, network inputs (8 bf16 per cell), packed network (zero = tau_t 0)
            self.t["tau_t"] = z(n)
            self.t["feat"] = z(n * abi.MLP_INPUTS, dtype=torch.int16)
            self.t["mlp"] = z(int(self.lib.ssmd_mlp_packed_bytes()), dtype=torch.uint8)
            self.d.tau_t = self.t["tau_t"].data_ptr()
            self.d.feat = self.t["feat"].data_ptr()
            self.d.mlp = self.t["mlp"].data_ptr()
        self.step_index = 0
        self._graph: Optional[C.c_void_p] = None
        torch.cuda.synchronize(self.device)  # zero fills above ran on the current stream

    # -- stream -------------------------------------------------------------
    @property
    def stream_handle(self) -> int:
        return self.stream.cuda_stream

    def synchronize(self) -> None:
        self.stream.synchronize()

    # -- lifecycle ----------------------------------------------------------
    def init(self) -> None:
        abi.check(self.lib.ssmd_init(C.byref(self.p), C.byref(self.d), self.stream_handle), "ssmd_init")
        self.step_index = 0

    def step(self, n: int = 1, macros: bool = False) -> None:
        """n eager steps; with macros=True the last one writes every moment field."""
        for i in range(n):
            flags = abi.STEP_MACROS if (macros and i == n - 1) else 0
            abi.check(self.lib.ssmd_step(C.byref(self.p), C.byref(self.d), self.step_index, flags,
                                         self.stream_handle), "ssmd_step")
            self.step_index += 1

    # the three launches of a step, for timing them separately
    def launch_phase(self) -> None:
        abi.check(self.lib.ssmd_launch_phase(C.byref(self.p), C.byref(self.d), self.stream_handle), "phase")

    def launch_lattice(self, macros: bool = False) -> None:
        abi.check(self.lib.ssmd_launch_lattice(C.byref(self.p), C.byref(self.d), abi.STEP_MACROS if macros else 0,
                                               self.stream_handle), "lattice")

    def launch_boundaries(self, macros: bool = False) -> None:
        abi.check(self.lib.ssmd_launch_boundaries(C.byref(self.p), C.byref(self.d), self.step_index, 0,
                                                  abi.STEP_MACROS if macros else 0, self.stream_handle), "boundaries")

    def closure_update(self) -> None:
        """The LEARNED closure's network pass (the last launch of a step)."""
        abi.check(self.lib.ssmd_closure_update(C.byref(self.p), C.byref(self.d), self.stream_handle), "closure")

    def swap(self) -> None:
        abi.check(self.lib.ssmd_swap(C.byref(self.d)), "swap")
        self.step_index += 1

    def graph_run(self, n: int) -> None:
        """n steps replayed from captured HIP graphs (boundaries included)."""
        if self._graph is None:
            h = C.c_void_p()
            abi.check(self.lib.ssmd_graph_create(C.byref(self.p), C.byref(self.d), self.stream_handle, C.byref(h)),
                      "ssmd_graph_create")
            self._graph = h
        abi.check(self.lib.ssmd_graph_launch(self._graph, C.byref(self.d), self.step_index, n, self.stream_handle),
                  "ssmd_graph_launch")
        self.step_index += n

    # -- restart ------------------------------------------------------------
    def checkpoint(self, path: str) -> None:
        """Write the state the next step needs (ssmd_checkpoint_write)."""
        abi.check(self.lib.ssmd_checkpoint_write(C.byref(self.p), C.byref(self.d), self.step_index,
                                                 str(path).encode(), self.stream_handle), "checkpoint_write")

    def restore(self, path: str) -> None:
        """Load a checkpoint of the same case; the step index comes with it."""
        step = C.c_uint64()
        abi.check(self.lib.ssmd_checkpoint_read(C.byref(self.p), C.byref(self.d), C.byref(step), str(path).encode(),
                                                self.stream_handle), "checkpoint_read")
        self.step_index = int(step.value)

    def tune(self, reps: int = 20, confirm_steps: int = 0) -> tuple[int, int]:
        """Autotune the lattice tile shape and z-chunk; the choice is kept in self.p.

        ssmd_tune times the lattice kernel alone.  With confirm_steps > 0 the
        pick is then checked on whole steps: confirm_steps graph-replayed steps
        at the pick and at the same tile with the automatic z-chunk, keeping
        the faster (kernel-alone near-ties between z-chunks can go either way
        in the step).  That advances the simulation by 2 (16 + confirm_steps)
        steps; without it the state is left as ssmd_tune leaves it."""
        t, lz = C.c_int32(), C.c_int32()
        abi.check(self.lib.ssmd_tune(C.byref(self.p), C.byref(self.d), reps, self.stream_handle, C.byref(t),
                                     C.byref(lz)), "ssmd_tune")
        self._set_shape(t.value, lz.value)
        if confirm_steps > 0:
            auto = int(self.lib.ssmd_auto_lz(C.byref(self.p), t.value))
            abi.check(min(auto, 0), "ssmd_auto_lz")
            if auto != lz.value:
                timed = {}
                for cand in (lz.value, auto):
                    self._set_shape(t.value, cand)
                    self.graph_run(16)
                    self.synchronize()
                    t0 = time.perf_counter()
                    self.graph_run(confirm_steps)
                    self.synchronize()
                    timed[cand] = time.perf_counter() - t0
                self._set_shape(t.value, min(timed, key=timed.get))
        return int(self.p.tile), int(self.p.lz)

    def _set_shape(self, tile: int, lz: int) -> None:
        self.p.tile, self.p.lz = tile, lz
        if self._graph is not None:  # graphs hold the old launch shape
            self.lib.ssmd_graph_destroy(self._graph)
            self._graph = None

    def close(self) -> None:
        if self._graph is not None:
            self.lib.ssmd_graph_destroy(self._graph)
            self._graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- readback -------------------------------------------------------------
    def field(self, name: str) -> np.ndarray:
        """A scalar field as float64 numpy (nz, ny, nx)."""
        self.synchronize()
        return self.t[name].double().cpu().numpy().reshape(self.p.shape)

    def fields(self, names: Iterable[str] = ("rho", "ux", "uy", "uz", "phi")) -> Dict[str, np.ndarray]:
        return {n: self.field(n) for n in names}

    def _current(self, which: str) -> torch.Tensor:
        ptr = self.d.f if which == "f" else self.d.g
        for key in (("fa", "fb") if which == "f" else ("ga", "gb")):
            if self._pops[key].data_ptr() == ptr:
                return self._pops[key]
        raise abi.SsmdError("population buffer bookkeeping lost")

    def pops_source(self, which: str = "f") -> np.ndarray:
        """Populations the next step reads, source-indexed, float64 (Q, nz, ny, nx).

        g is stored as 4 moment planes (g_moments); its seven populations are
        formed here with the FP32 operations the phase kernel applies."""
        self.synchronize()
        if which == "g":
            return phase_populations(self.g_moments()).astype(np.float64)
        q = int(self.p.q)
        return self._current(which).double().cpu().numpy().reshape((q,) + self.p.shape)

    def g_moments(self) -> np.ndarray:
        """The stored phase state, float32 (4, nz, ny, nx): g_phi, o_x, o_y, o_z."""
        self.synchronize()
        return self._current("g").cpu().numpy().reshape((4,) + self.p.shape)

    def set_pops_source(self, which: str, values: np.ndarray) -> None:
        """Overwrite both buffers of f or g (source layout) with `values`.

        For g the seven populations of every cell must have the moment form
        w_k (g_phi + c_k . o); anything else raises ValueError."""
        self.synchronize()
        keys = ("fa", "fb") if which == "f" else ("ga", "gb")
        if which == "g":
            values = phase_moments(np.asarray(values, dtype=np.float64).reshape((7,) + self.p.shape))
        src = torch.from_numpy(np.ascontiguousarray(values, dtype=np.float64).reshape(-1))
        for k in keys:
            self._pops[k].copy_(src.to(self._pops[k].dtype).to(self.device))
        torch.cuda.synchronize(self.device)

    def pops_arrival(self, which: str = "f") -> np.ndarray:
        """Same populations in the reference's arrival-indexed layout (periodic
        images folded in); entries whose source lies outside the domain are NaN."""
        return source_to_arrival(self.pops_source(which), self.p)

    def phase_mass(self) -> np.ndarray:
        """FP64 [sum phi, sum rho ux, sum rho uy, sum rho uz] over interior cells."""
        out = (C.c_double * 4)()
        abi.check(self.lib.ssmd_phase_mass_f64(C.byref(self.p), C.byref(self.d), out, self.stream_handle),
                  "ssmd_phase_mass_f64")
        return np.array(out[:], dtype=np.float64)

    def check_finite(self) -> np.ndarray:
        """NaN/Inf sentinel (the reference has none): the FP64 sums of
        phase_mass() must be finite.  One non-finite population reaches rho and
        u of the cell that pulls it in the next step, and phi one step later;
        the momentum sums see it only when the last step ran with macros=True.
        Returns the sums; raises SsmdError naming the step otherwise."""
        s = self.phase_mass()
        if not np.isfinite(s).all():
            raise abi.SsmdError(f"non-finite state after step {self.step_index}: "
                                f"sum phi {s[0]!r}, sum rho u ({s[1]!r}, {s[2]!r}, {s[3]!r})")
        return s


def _source_axis(n: int, c: int, wrap: int):
    """Source coordinate of each arrival coordinate 0..n-1 along one axis for
    a direction component c, with the periodic image folded in; -1 = none."""
    s = np.arange(n) - c
    valid = (s >= 0) & (s < n)
    if wrap == abi.WRAP_REFERENCE:
        s[(s == 0) & (c > 0)] = n - 3
        s[(s == n - 1) & (c < 0)] = 2
    elif wrap == abi.WRAP_TRUE:
        s[(s == 0) & (c > 0)] = n - 2
        s[(s == n - 1) & (c < 0)] = 1
    s[~valid] = -1
    return s


W7 = (0.25, 0.125, 0.125, 0.125, 0.125, 0.125, 0.125)


def phase_populations(mom: np.ndarray) -> np.ndarray:
    """Seven D3Q7 populations (float32) from the 4 stored moment planes:
    k = 0: w_0 g_phi; k = 1 + 2a / 2 + 2a: w (g_phi +/- o_a), in FP32."""
    m = np.asarray(mom, dtype=np.float32)
    gp = m[0]
    out = [np.float32(W7[0]) * gp]
    for a in range(3):
        out.append(np.float32(W7[1]) * (gp + m[1 + a]))
        out.append(np.float32(W7[2]) * (gp - m[1 + a]))
    return np.stack(out)


def phase_moments(pops: np.ndarray, rtol: float = 1e-6) -> np.ndarray:
    """Inverse of phase_populations: (g_phi, o_x, o_y, o_z) from seven
    populations of the form w_k (g_phi + c_k . o); ValueError otherwise."""
    g = np.asarray(pops, dtype=np.float64)
    gp = g[0] / W7[0]
    mom = [gp] + [(g[1 + 2 * a] - g[2 + 2 * a]) / (2 * W7[1]) for a in range(3)]
    back = phase_populations(np.stack(mom)).astype(np.float64)
    scale = max(float(np.max(np.abs(g))), 1e-30)
    if np.max(np.abs(back - g)) > rtol * scale + 1e-12:
        raise ValueError("phase populations are not of the form w_k (g_phi + c_k . o)")
    return np.stack(mom)


def source_to_arrival(src: np.ndarray, params: abi.Params | None = None) -> np.ndarray:
    """arrival[i][a] = source[i][a - c_i] (NaN where a - c_i is outside).

    With `params`, a source slot on a periodic ghost face is replaced by its
    periodic image, as the kernels read it (the wrap is part of the pull)."""
    q = src.shape[0]
    out = np.full(src.shape, np.nan)
    nz, ny, nx = src.shape[1:]
    wx = int(params.wrap_x) if params is not None else 0
    wy = int(params.wrap_y) if params is not None else 0
    wz = int(params.wrap_z) if params is not None else 0
    for i in range(q):
        cx, cy, cz = DIRECTIONS[i]
        sx, sy, sz = _source_axis(nx, cx, wx), _source_axis(ny, cy, wy), _source_axis(nz, cz, wz)
        ax, ay, az = np.nonzero(sx >= 0)[0], np.nonzero(sy >= 0)[0], np.nonzero(sz >= 0)[0]
        out[i][np.ix_(az, ay, ax)] = src[i][np.ix_(sz[az], sy[ay], sx[ax])]
    return out


def arrival_to_source(arr: np.ndarray) -> np.ndarray:
    """Inverse of source_to_arrival without wraps (0 where the arrival is outside)."""
    q = arr.shape[0]
    out = np.zeros(arr.shape)
    nz, ny, nx = arr.shape[1:]
    for i in range(q):
        cx, cy, cz = DIRECTIONS[i]
        az = slice(max(cz, 0), nz + min(cz, 0))
        ay = slice(max(cy, 0), ny + min(cy, 0))
        ax = slice(max(cx, 0), nx + min(cx, 0))
        sz = slice(max(-cz, 0), nz + min(-cz, 0))
        sy = slice(max(-cy, 0), ny + min(-cy, 0))
        sx = slice(max(-cx, 0), nx + min(-cx, 0))
        out[i][sz, sy, sx] = arr[i][az, ay, ax]
    return out


class Ensemble:
    """Independent jets advanced together by single batched launches
    (persistent lattice kernel, batched phase and face kernels)."""

    def __init__(self, params_list, device: str | torch.device = "cuda:0"):
        self.lib = abi.load()
        self.members = [Simulation(p, device) for p in params_list]
        self.stream = torch.cuda.Stream(torch.device(device))
        for m in self.members:
            m.stream = self.stream
        self.n = len(self.members)
        self._p = (abi.Params * self.n)(*[m.p for m in self.members])
        self._h = None
        self.step_index = 0

    @property
    def stream_handle(self) -> int:
        return self.stream.cuda_stream

    def init(self) -> None:
        for m in self.members:
            m.init()
        torch.cuda.synchronize()
        if self._h is None:
            d = (abi.Fields * self.n)(*[m.d for m in self.members])
            h = C.c_void_p()
            abi.check(self.lib.ssmd_ensemble_create(self._p, d, self.n, C.byref(h)), "ssmd_ensemble_create")
            self._h = h
        self.step_index = 0

    def _sync_members(self) -> None:
        for j, m in enumerate(self.members):
            abi.check(self.lib.ssmd_ensemble_fields(self._h, j, C.byref(m.d)), "ssmd_ensemble_fields")
            m.step_index = self.step_index

    def step(self, n: int = 1, macros: bool = False) -> None:
        for i in range(n):
            flags = abi.STEP_MACROS if (macros and i == n - 1) else 0
            abi.check(self.lib.ssmd_ensemble_step(self._h, self.step_index, flags, self.stream_handle),
                      "ssmd_ensemble_step")
            self.step_index += 1
        self._sync_members()

    def graph_run(self, n: int) -> None:
        abi.check(self.lib.ssmd_ensemble_graph_launch(self._h, self.step_index, n, self.stream_handle),
                  "ssmd_ensemble_graph_launch")
        self.step_index += n
        self._sync_members()

    def tune(self, reps: int = 10) -> int:
        t = C.c_int32()
        abi.check(self.lib.ssmd_ensemble_tune(self._h, reps, self.stream_handle, C.byref(t)), "ssmd_ensemble_tune")
        return t.value

    def set_tile(self, tile: int) -> None:
        abi.check(self.lib.ssmd_ensemble_set_tile(self._h, tile), "ssmd_ensemble_set_tile")

    def synchronize(self) -> None:
        self.stream.synchronize()

    @property
    def cells(self) -> int:
        return sum(m.p.cells for m in self.members)

    def close(self) -> None:
        if self._h is not None:
            self.lib.ssmd_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


DERIVED_FIELDS = tuple(n for n, _ in abi.Derived._fields_)


class DerivedStats:
    """The reference's derived fields (derivedFields/) for one Simulation:
    running means of phi and u, Reynolds moments <u_i u_j>, vorticity."""

    def __init__(self, sim: Simulation, names: Iterable[str] = DERIVED_FIELDS):
        self.sim = sim
        self.t: Dict[str, torch.Tensor] = {}
        self.s = abi.Derived()
        for n in names:
            if n not in DERIVED_FIELDS:
                raise ValueError(f"unknown derived field {n}")
            self.t[n] = torch.zeros(sim.p.cells, dtype=torch.float32, device=sim.device)
            setattr(self.s, n, self.t[n].data_ptr())
        torch.cuda.synchronize(sim.device)

    def update(self) -> None:
        """After a step run with macros=True: the statistics of that step."""
        sim = self.sim
        abi.check(sim.lib.ssmd_derived_update(C.byref(sim.p), C.byref(sim.d), C.byref(self.s), sim.step_index - 1,
                                              sim.stream_handle), "derived_update")

    def field(self, name: str) -> np.ndarray:
        self.sim.synchronize()
        return self.t[name].double().cpu().numpy().reshape(self.sim.p.shape)
