"""ctypes binding of the C ABI declared in include/ssmd.h (libssmd.so).

The structures mirror the header field for field.  Loading fails loudly
when the library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libssmd.so"

POP_F32, POP_F16, POP_BF16 = 0, 1, 2
WRAP_NONE, WRAP_REFERENCE, WRAP_TRUE = 0, 1, 2
CLOSURE_NONE, CLOSURE_SMAGORINSKY, CLOSURE_LEARNED = 0, 1, 2
MLP_INPUTS, MLP_HIDDEN, MLP_PARAMS = 8, 64, 8961
MLP_GRAD_NOALLOC_MAX = 65536
OK, EINVAL, EUNSUPPORTED, ENOTREADY = 0, -1, -2, -3
ABI_VERSION = 6
STEP_MACROS = 1
REDUCE_DOUBLES = 8192
POP_BYTES = {POP_F32: 4, POP_F16: 2, POP_BF16: 2}


class Jet(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("diameter", C.c_int32),
        ("center_a", C.c_double),
        ("center_b", C.c_double),
        ("r2", C.c_double),
        ("u", C.c_double),
        ("phi", C.c_double),
    ]


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32),
        ("ny", C.c_uint32),
        ("nz", C.c_uint32),
        ("q", C.c_int32),
        ("pop_format", C.c_int32),
        ("outflow_y", C.c_int32),
        ("outflow_z", C.c_int32),
        ("wrap_x", C.c_int32),
        ("wrap_y", C.c_int32),
        ("wrap_z", C.c_int32),
        ("tile", C.c_int32),
        ("lz", C.c_int32),
        ("visc_water", C.c_double),
        ("visc_oil", C.c_double),
        ("sigma", C.c_double),
        ("gamma", C.c_double),
        ("perturb", C.c_double),
        ("seed", C.c_uint64),
        ("water", Jet),
        ("oil", Jet),
        ("closure", C.c_int32),
        ("pad1", C.c_int32),
        ("cs", C.c_double),
        ("strain_scale", C.c_double),
        ("grad_scale", C.c_double),
        ("tau_scale", C.c_double),
        ("tau_floor", C.c_double),
        ("flux_coeff", C.c_double),
    ]

    @property
    def cells(self) -> int:
        return int(self.nx) * int(self.ny) * int(self.nz)

    @property
    def shape(self) -> tuple[int, int, int]:
        """(nz, ny, nx): numpy view order of a scalar field."""
        return (int(self.nz), int(self.ny), int(self.nx))

    def copy(self) -> "Params":
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p


SCALAR_FIELDS = ("rho", "ux", "uy", "uz", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz", "phi",
                 "normx", "normy", "normz", "ind", "fsx", "fsy", "fsz")


class Fields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in SCALAR_FIELDS] + [
        ("f", C.c_void_p),
        ("g", C.c_void_p),
        ("f_next", C.c_void_p),
        ("g_next", C.c_void_p),
        ("step_counter", C.c_void_p),
        ("reduce", C.c_void_p),
        ("owned", C.c_uint64),
        ("tau_t", C.c_void_p),
        ("feat", C.c_void_p),
        ("mlp", C.c_void_p),
    ]


class Derived(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "avg_phi", "avg_ux", "avg_uy", "avg_uz", "avg_uxux", "avg_uyuy", "avg_uzuz", "avg_uxuy", "avg_uxuz",
        "avg_uyuz", "vort_x", "vort_y", "vort_z", "vort_mag")]


# name -> (restype, argtypes); every function the header declares
SIGNATURES = {
    "ssmd_abi_version": (C.c_int, []),
    "ssmd_error_string": (C.c_char_p, [C.c_int]),
    "ssmd_params_reference_default": (C.c_int, [C.POINTER(Params), C.c_int32]),
    "ssmd_validate": (C.c_int, [C.POINTER(Params)]),
    "ssmd_bytes": (C.c_size_t, [C.POINTER(Params), C.c_int32]),
    "ssmd_alloc": (C.c_int, [C.POINTER(Params), C.POINTER(Fields)]),
    "ssmd_free": (C.c_int, [C.POINTER(Fields)]),
    "ssmd_init": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_void_p]),
    "ssmd_step": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_uint64, C.c_uint32, C.c_void_p]),
    "ssmd_launch_phase": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_void_p]),
    "ssmd_launch_lattice": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_uint32, C.c_void_p]),
    "ssmd_launch_boundaries": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_uint64, C.c_int32, C.c_uint32,
                                         C.c_void_p]),
    "ssmd_swap": (C.c_int, [C.POINTER(Fields)]),
    "ssmd_run": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_uint64, C.c_uint64, C.c_void_p]),
    "ssmd_graph_create": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_void_p, C.POINTER(C.c_void_p)]),
    "ssmd_graph_launch": (C.c_int, [C.c_void_p, C.POINTER(Fields), C.c_uint64, C.c_uint64, C.c_void_p]),
    "ssmd_graph_destroy": (C.c_int, [C.c_void_p]),
    "ssmd_phase_mass_f64": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.POINTER(C.c_double), C.c_void_p]),
    "ssmd_populations_to_reference": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_void_p, C.c_void_p,
                                                C.c_void_p]),
    "ssmd_quantize": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "ssmd_tune": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_int32, C.c_void_p, C.POINTER(C.c_int32),
                            C.POINTER(C.c_int32)]),
    "ssmd_tile_count": (C.c_int, []),
    "ssmd_auto_lz": (C.c_int32, [C.POINTER(Params), C.c_int32]),
    "ssmd_tile_shape": (C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ssmd_ensemble_create": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_int32, C.POINTER(C.c_void_p)]),
    "ssmd_ensemble_step": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p]),
    "ssmd_ensemble_graph_launch": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p]),
    "ssmd_ensemble_tune": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_int32)]),
    "ssmd_ensemble_set_tile": (C.c_int, [C.c_void_p, C.c_int32]),
    "ssmd_ensemble_fields": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Fields)]),
    "ssmd_ensemble_destroy": (C.c_int, [C.c_void_p]),
    "ssmd_write_bin": (C.c_int, [C.POINTER(Params), C.c_void_p, C.c_char_p, C.c_char_p, C.c_uint64, C.c_void_p]),
    "ssmd_write_vti": (C.c_int, [C.POINTER(Params), C.c_char_p, C.c_uint64, C.POINTER(C.c_char_p), C.c_int32]),
    "ssmd_write_info": (C.c_int, [C.POINTER(Params), C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint64, C.c_uint64]),
    "ssmd_derived_update": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.POINTER(Derived), C.c_uint64,
                                      C.c_void_p]),
    "ssmd_checkpoint_write": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_uint64, C.c_char_p, C.c_void_p]),
    "ssmd_checkpoint_read": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.POINTER(C.c_uint64), C.c_char_p,
                                       C.c_void_p]),
    "ssmd_mlp_packed_bytes": (C.c_size_t, []),
    "ssmd_mlp_pack": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "ssmd_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "ssmd_mlp_grad_work_bytes": (C.c_size_t, [C.c_uint32]),
    "ssmd_mlp_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_size_t, C.c_void_p]),
    "ssmd_mlp_grad_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                        C.c_void_p, C.c_void_p]),
    "ssmd_mlp_adam": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_float,
                                C.c_float, C.c_float, C.c_void_p]),
    "ssmd_closure_update": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.c_void_p]),
    "ssmd_sgs_work_bytes": (C.c_size_t, [C.POINTER(Params)]),
    "ssmd_sgs_samples": (C.c_int, [C.POINTER(Params), C.POINTER(Fields), C.POINTER(Params), C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p]),
}


class SsmdError(RuntimeError):
    pass


_lib = None


def load() -> C.CDLL:
    """Load libssmd.so (built by ssmdlbm_amd.build); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise SsmdError(f"{LIB_PATH} is not built; run __graft_entry__.build() or python -m ssmdlbm_amd.build")
    # torch loads its own HIP runtime; loading it first makes libssmd bind to
    # that same runtime (a second copy sees no device and shares no memory).
    import torch  # noqa: F401
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.ssmd_abi_version() != ABI_VERSION:
        raise SsmdError("libssmd ABI version mismatch")
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().ssmd_error_string(rc).decode()
        raise SsmdError(f"{what}: error {rc} ({msg})")
