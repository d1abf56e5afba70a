"""The learned subgrid closure: network parameters, training and use.

No counterpart in the reference (it has no subgrid model).  The network
(8 -> 64 -> 64 -> 64 -> 1, ReLU) runs on the gfx950 matrix cores through the
C ABI (ssmd_mlp_*); this module only owns its device buffers:

  params   SSMD_MLP_PARAMS float32 master weights (W1, b1, W2, b2, W3, b3, W4, b4)
  m, v     Adam moments
  packed   the bf16 operand layout the kernels read (ssmd_mlp_pack)

Inputs are bf16 bit patterns, 8 per sample: the strain rate (6, scaled),
phi and |grad phi-hat| (scaled), in the order the lattice kernel writes them.
"""
from __future__ import annotations

import ctypes as C
import json
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np
import torch

from . import abi

N_PARAMS = abi.MLP_PARAMS
N_IN, N_HID = abi.MLP_INPUTS, abi.MLP_HIDDEN


def offsets() -> dict:
    """Parameter slices (start, stop) of the flat parameter vector."""
    o, out = 0, {}
    for name, size in (("W1", N_HID * N_IN), ("b1", N_HID), ("W2", N_HID * N_HID), ("b2", N_HID),
                       ("W3", N_HID * N_HID), ("b3", N_HID), ("W4", N_HID), ("b4", 1)):
        out[name] = (o, o + size)
        o += size
    assert o == N_PARAMS
    return out


def bf16_bits(x: torch.Tensor) -> torch.Tensor:
    """float tensor (n, 8) -> contiguous int16 bit patterns of its bf16 rounding (RNE)."""
    return x.to(torch.bfloat16).contiguous().view(torch.int16)


def bits_to_float(b: torch.Tensor) -> torch.Tensor:
    return b.contiguous().view(torch.bfloat16).float()


@dataclass
class Scales:
    """Input/output scaling the lattice kernel applies (ssmd_params fields):
    tau_t = max(tau_floor, tau_scale * network); tau_floor < 0 admits backscatter."""
    strain_scale: float = 1.0
    grad_scale: float = 1.0
    tau_scale: float = 1.0
    tau_floor: float = 0.0

    def apply(self, p: abi.Params) -> None:
        p.strain_scale, p.grad_scale, p.tau_scale = self.strain_scale, self.grad_scale, self.tau_scale
        p.tau_floor = self.tau_floor


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, dtype=np.float32)
    o = offsets()
    for name, fan_in in (("W1", N_IN), ("W2", N_HID), ("W3", N_HID), ("W4", N_HID)):
        a, b = o[name]
        p[a:b] = rng.normal(0.0, np.sqrt(2.0 / fan_in), b - a)
    return p


class Network:
    """One network on one GPU; every call is asynchronous on torch's current stream."""

    def __init__(self, params: np.ndarray | None = None, device: str | torch.device = "cuda:0", seed: int = 0,
                 scales: Scales | None = None):
        self.lib = abi.load()
        if not torch.cuda.is_available():
            raise abi.SsmdError("no GPU: the network runs only on the device")
        self.device = torch.device(device)
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        if p.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters, got {p.size}")
        self.params = torch.tensor(p, device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.loss = torch.zeros(1, device=self.device)
        self.packed = torch.zeros(int(self.lib.ssmd_mlp_packed_bytes()), dtype=torch.uint8, device=self.device)
        self.scales = scales or Scales()
        self.adam_steps = 0
        self._work = torch.zeros(0, dtype=torch.uint8, device=self.device)
        self.pack()

    @property
    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def pack(self) -> None:
        abi.check(self.lib.ssmd_mlp_pack(self.params.data_ptr(), self.packed.data_ptr(), self._stream), "mlp_pack")

    def forward(self, x_bits: torch.Tensor) -> torch.Tensor:
        """Network output for (n, 8) int16 bf16 bit patterns."""
        x = self._inputs(x_bits)
        y = torch.empty(x.shape[0], dtype=torch.float32, device=self.device)
        abi.check(self.lib.ssmd_mlp_forward(self.packed.data_ptr(), x.data_ptr(), y.data_ptr(), x.shape[0],
                                            self._stream), "mlp_forward")
        return y

    def gradient(self, x_bits: torch.Tensor, target: torch.Tensor,
                 weight: torch.Tensor | None = None) -> tuple[torch.Tensor, torch.Tensor]:
        """(grad, loss) of the (weighted) mean squared error, into self.grad / self.loss."""
        x = self._inputs(x_bits)
        t = target.to(device=self.device, dtype=torch.float32).contiguous()
        n = x.shape[0]
        if t.numel() != n:
            raise ValueError("one target per sample")
        w = None
        if weight is not None:
            w = weight.to(device=self.device, dtype=torch.float32).contiguous()
            if w.numel() != n:
                raise ValueError("one weight per sample")
        need = int(self.lib.ssmd_mlp_grad_work_bytes(n))
        if self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.uint8, device=self.device)
        abi.check(self.lib.ssmd_mlp_grad(self.packed.data_ptr(), x.data_ptr(), t.data_ptr(),
                                         None if w is None else w.data_ptr(), n, self.grad.data_ptr(),
                                         self.loss.data_ptr(), self._work.data_ptr(), self._work.numel(),
                                         self._stream), "mlp_grad")
        return self.grad, self.loss

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        """One Adam step from self.grad, then repack the bf16 operands."""
        self.adam_steps += 1
        abi.check(self.lib.ssmd_mlp_adam(self.params.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(),
                                         self.v.data_ptr(), self.adam_steps, lr, beta1, beta2, eps, self._stream),
                  "mlp_adam")
        self.pack()

    def train(self, x_bits: torch.Tensor, target: torch.Tensor, steps: int, batch: int = 65536, lr: float = 1e-3,
              seed: int = 0, log_every: int = 0, weight: torch.Tensor | None = None,
              lr_final: float | None = None) -> list[float]:
        """Minibatch Adam on (x, target[, weight]) resident on the device, the
        learning rate decaying geometrically to lr_final (constant when None);
        returns the loss of every `log_every`-th step (all steps when 0)."""
        x = self._inputs(x_bits)
        t = target.to(device=self.device, dtype=torch.float32)
        w = None if weight is None else weight.to(device=self.device, dtype=torch.float32)
        n = x.shape[0]
        g = torch.Generator(device=self.device)
        g.manual_seed(seed)
        losses = []
        for s in range(steps):
            if batch >= n:
                xb, tb, wb = x, t, w
            else:
                idx = torch.randint(0, n, (batch,), device=self.device, generator=g)
                xb, tb = x[idx].contiguous(), t[idx].contiguous()
                wb = None if w is None else w[idx].contiguous()
            self.gradient(xb, tb, wb)
            rate = lr if lr_final is None or steps < 2 else lr * (lr_final / lr) ** (s / (steps - 1))
            self.adam(rate)
            if not log_every or s % log_every == 0 or s == steps - 1:
                losses.append(float(self.loss.item()))
        return losses

    def load_into(self, sim) -> None:
        """Install this network (and its scales) in a Simulation run with the LEARNED closure."""
        if int(sim.p.closure) != abi.CLOSURE_LEARNED:
            raise abi.SsmdError("the simulation does not use the learned closure")
        torch.cuda.current_stream(self.device).synchronize()
        sim.t["mlp"].copy_(self.packed)
        self.scales.apply(sim.p)
        torch.cuda.synchronize(self.device)

    def numpy_params(self) -> np.ndarray:
        return self.params.detach().cpu().numpy().astype(np.float32)

    def save(self, path: str | Path) -> None:
        path = Path(path)
        np.save(path.with_suffix(".npy"), self.numpy_params())
        path.with_suffix(".json").write_text(json.dumps({"scales": asdict(self.scales), "params": N_PARAMS,
                                                         "adam_steps": self.adam_steps}, indent=1))

    @classmethod
    def load(cls, path: str | Path, device: str | torch.device = "cuda:0") -> "Network":
        path = Path(path)
        meta = json.loads(path.with_suffix(".json").read_text())
        net = cls(np.load(path.with_suffix(".npy"), allow_pickle=False), device=device, scales=Scales(**meta["scales"]))
        return net

    def _inputs(self, x_bits: torch.Tensor) -> torch.Tensor:
        x = x_bits.to(self.device)
        if x.dtype != torch.int16 or x.dim() != 2 or x.shape[1] != N_IN:
            raise ValueError("inputs are (n, 8) int16 bf16 bit patterns")
        x = x.contiguous()
        if x.data_ptr() % 16:
            x = x.clone()
        return x
