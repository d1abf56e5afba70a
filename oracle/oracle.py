"""ctypes wrapper of the CPU oracle (oracle/ssmd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from ssmdlbm_amd import abi

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libssmd_oracle.so"

FIELD_IDS = {n: i for i, n in enumerate(
    ("rho", "ux", "uy", "uz", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz", "phi", "fsx", "fsy", "fsz", "tau_t")
    + tuple(f"feat{j}" for j in range(8)))}

_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run python oracle/build.py")
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER(abi.Params)
    h = C.c_void_p
    d = C.POINTER(C.c_double)
    sig = {
        "oracle_create": (h, [P, C.c_int]),
        "oracle_destroy": (None, [h]),
        "oracle_set_threads": (None, [h, C.c_int]),
        "oracle_init": (None, [h]),
        "oracle_step": (None, [h, C.c_uint64]),
        "oracle_run": (None, [h, C.c_uint64, C.c_uint64]),
        "oracle_k_phase": (None, [h]),
        "oracle_k_forces": (None, [h]),
        "oracle_k_moments": (None, [h]),
        "oracle_k_collide": (None, [h]),
        "oracle_k_boundaries": (None, [h, C.c_uint64]),
        "oracle_get_field": (C.c_int, [h, C.c_int, d]),
        "oracle_set_field": (C.c_int, [h, C.c_int, d]),
        "oracle_get_pops": (C.c_int, [h, C.c_int, d]),
        "oracle_set_pops": (C.c_int, [h, C.c_int, d]),
        "oracle_sums": (None, [h, d]),
        "oracle_w": (C.c_double, [C.c_int, C.c_int]),
        "oracle_c": (C.c_int, [C.c_int, C.c_int]),
        "oracle_feq": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
        "oracle_fneq": (C.c_double, [C.c_int, C.c_int, d, C.c_double, C.c_double, C.c_double]),
        "oracle_force": (C.c_double, [C.c_int, C.c_int] + [C.c_double] * 6),
        "oracle_geq": (C.c_double, [C.c_int] + [C.c_double] * 4),
        "oracle_antidiff": (C.c_double, [C.c_int] + [C.c_double] * 4),
        "oracle_f32_to_f16": (C.c_uint16, [C.c_float]),
        "oracle_f16_to_f32": (C.c_float, [C.c_uint16]),
        "oracle_f32_to_bf16": (C.c_uint16, [C.c_float]),
        "oracle_bf16_to_f32": (C.c_float, [C.c_uint16]),
        "oracle_splitmix64": (C.c_uint64, [C.c_uint64]),
        "oracle_noise": (C.c_float, [C.c_uint64] * 4),
        "oracle_smag_tau": (C.c_double, [C.c_double, d, C.c_double, C.c_double]),
        "oracle_mlp_params": (C.c_int, []),
        "oracle_mlp_forward": (None, [d, d, C.c_int64, C.c_int, d]),
        "oracle_mlp_grad": (None, [d, d, d, d, C.c_int64, C.c_int, d, d]),
        "oracle_bf16_decode": (None, [C.POINTER(C.c_uint16), C.c_int64, d]),
        "oracle_adam": (None, [d, d, d, d, C.c_int64, C.c_int64] + [C.c_double] * 4),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib.oracle_params_size.restype = C.c_size_t
    if lib.oracle_params_size() != C.sizeof(abi.Params):
        raise RuntimeError("oracle built against a different ssmd_params layout; rebuild with oracle/build.py")
    _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """The reference step restated on the CPU (FP64 by default)."""

    def __init__(self, params: abi.Params, fp64: bool = True, threads: int | None = None):
        self.lib = load()
        self.p = params.copy()
        self.h = self.lib.oracle_create(C.byref(self.p), 1 if fp64 else 0)
        if not self.h:
            raise ValueError("oracle_create rejected the parameters")
        if threads:
            self.lib.oracle_set_threads(self.h, int(threads))
        self.q = int(self.p.q)
        self.shape = self.p.shape
        self.step_index = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_destroy(self.h)
            self.h = None

    def init(self):
        self.lib.oracle_init(self.h)
        self.step_index = 0

    def step(self, n: int = 1):
        self.lib.oracle_run(self.h, self.step_index, n)
        self.step_index += n

    # kernel-level entry points (reference kernel order)
    def k_phase(self):
        self.lib.oracle_k_phase(self.h)

    def k_forces(self):
        self.lib.oracle_k_forces(self.h)

    def k_moments(self):
        self.lib.oracle_k_moments(self.h)

    def k_collide(self):
        self.lib.oracle_k_collide(self.h)

    def k_boundaries(self, step: int):
        self.lib.oracle_k_boundaries(self.h, step)

    def field(self, name: str) -> np.ndarray:
        out = np.empty(self.p.cells)
        rc = self.lib.oracle_get_field(self.h, FIELD_IDS[name], _dptr(out))
        assert rc == 0
        return out.reshape(self.shape)

    def fields(self, names=("rho", "ux", "uy", "uz", "phi")):
        return {n: self.field(n) for n in names}

    def set_field(self, name: str, values: np.ndarray):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        assert a.size == self.p.cells
        assert self.lib.oracle_set_field(self.h, FIELD_IDS[name], _dptr(a)) == 0

    def pops(self, which: str = "f") -> np.ndarray:
        q = self.q if which == "f" else 7
        out = np.empty(q * self.p.cells)
        assert self.lib.oracle_get_pops(self.h, 0 if which == "f" else 1, _dptr(out)) == 0
        return out.reshape((q,) + self.shape)

    def set_pops(self, which: str, values: np.ndarray):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        assert self.lib.oracle_set_pops(self.h, 0 if which == "f" else 1, _dptr(a)) == 0

    def sums(self) -> np.ndarray:
        out = np.zeros(4)
        self.lib.oracle_sums(self.h, _dptr(out))
        return out


# -- closure network ----------------------------------------------------------
def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def mlp_forward(params, x, quant: bool = True) -> np.ndarray:
    """Network output per sample; x is (n, 8)."""
    lib = load()
    p, xx = _f64(params), _f64(x).reshape(-1, 8)
    y = np.empty(xx.shape[0])
    lib.oracle_mlp_forward(_dptr(p), _dptr(xx), xx.shape[0], int(quant), _dptr(y))
    return y


def mlp_grad(params, x, t, quant: bool = True, w=None):
    """(gradient of the (weighted) mean squared error, loss)."""
    lib = load()
    p, xx, tt = _f64(params), _f64(x).reshape(-1, 8), _f64(t)
    ww = None if w is None else _f64(w)
    g = np.empty(lib.oracle_mlp_params())
    loss = C.c_double()
    lib.oracle_mlp_grad(_dptr(p), _dptr(xx), _dptr(tt), None if ww is None else _dptr(ww), xx.shape[0], int(quant),
                        _dptr(g), C.byref(loss))
    return g, loss.value


def bf16_decode(bits) -> np.ndarray:
    lib = load()
    b = np.ascontiguousarray(bits, dtype=np.uint16)
    out = np.empty(b.size)
    lib.oracle_bf16_decode(b.ctypes.data_as(C.POINTER(C.c_uint16)), b.size, _dptr(out))
    return out.reshape(b.shape)


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in place on float64 arrays."""
    load().oracle_adam(_dptr(params), _dptr(_f64(grad)), _dptr(m), _dptr(v), params.size, step, lr, b1, b2, eps)
