"""The HBM streaming probe (csrc/probe/hbm_probe.hip, libssmd_probe.so):
read-only, write-only and copy streams of 16-byte accesses, cached or
non-temporal, over one pass or a persistent grid.  bench.py uses it to state
the practical bandwidth ceiling beside the 8 TB/s peak, measured in the same
process as the solver workloads."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libssmd_probe.so"
OPS = {"read": 0, "write": 1, "copy": 2}

_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        lib = C.CDLL(str(LIB_PATH))
        lib.ssmd_probe_max_blocks.restype = C.c_uint32
        lib.ssmd_probe_max_blocks.argtypes = [C.c_uint64]
        lib.ssmd_probe_stream.restype = C.c_int
        lib.ssmd_probe_stream.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int,
                                          C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]
        _lib = lib
    return _lib


def sweep(sizes_gib=(1, 4), reps: int = 10, persistent_blocks: int = 2048, device: str = "cuda:0") -> list[dict]:
    """GB/s of every (op, non-temporal, grid, size) combination.  Bytes moved
    count reads and writes: a copy of B bytes moves 2B."""
    import torch

    lib = load()
    out = []
    stream = torch.cuda.current_stream(device)
    for gib in sizes_gib:
        nbytes = int(gib) << 30
        src = torch.ones(nbytes // 4, dtype=torch.int32, device=device)
        dst = torch.empty_like(src)
        partial = torch.zeros(int(lib.ssmd_probe_max_blocks(nbytes)), dtype=torch.int32, device=device)
        torch.cuda.synchronize()
        for op, code in OPS.items():
            for nt in (0, 1):
                for blocks in (0, persistent_blocks):
                    ms = C.c_float()
                    rc = lib.ssmd_probe_stream(code, nt, blocks, src.data_ptr(), dst.data_ptr(), nbytes, reps,
                                               partial.data_ptr(), stream.cuda_stream, C.byref(ms))
                    if rc != 0:
                        raise RuntimeError(f"probe {op} nt={nt} blocks={blocks}: error {rc}")
                    moved = nbytes * (2 if op == "copy" else 1)
                    out.append({"op": op, "nontemporal": bool(nt), "grid": "persistent" if blocks else "one_pass",
                                "gib": gib, "ms": round(ms.value, 4),
                                "gbs": round(moved / (ms.value * 1e-3) / 1e9, 1)})
        del src, dst, partial
        torch.cuda.empty_cache()
    return out


def ceiling(results: list[dict]) -> dict:
    """Best rate per op over the sweep."""
    best = {}
    for r in results:
        if r["op"] not in best or r["gbs"] > best[r["op"]]["gbs"]:
            best[r["op"]] = r
    return best
