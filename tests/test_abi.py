"""C ABI checks that need no GPU: exports, struct layout, host-side logic."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ssmdlbm_amd import abi, cases

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "ssmd.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(ssmd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = abi.load()
    names = declared_functions()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n
    # the binding covers the whole header
    assert sorted(abi.SIGNATURES) == names


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "ssmd.h"
int main(void) {
  printf("%zu %zu %zu\\n", sizeof(ssmd_jet), sizeof(ssmd_params), sizeof(ssmd_fields));
  printf("%zu %zu %zu %zu\\n", offsetof(ssmd_params, visc_water), offsetof(ssmd_params, seed),
         offsetof(ssmd_params, water), offsetof(ssmd_params, oil));
  printf("%zu %zu %zu\\n", offsetof(ssmd_fields, f), offsetof(ssmd_fields, step_counter), offsetof(ssmd_fields, owned));
  printf("%zu %zu %zu %zu\\n", offsetof(ssmd_params, closure), offsetof(ssmd_params, cs),
         offsetof(ssmd_params, tau_floor), offsetof(ssmd_fields, mlp));
  printf("%d %d %d\\n", SSMD_ABI_VERSION, SSMD_MLP_PARAMS, SSMD_CLOSURE_LEARNED);
  printf("%zu\\n", offsetof(ssmd_params, flux_coeff));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = [list(map(int, l.split())) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                               check=True).stdout.splitlines()]
    assert out[0] == [C.sizeof(abi.Jet), C.sizeof(abi.Params), C.sizeof(abi.Fields)]
    assert out[1] == [abi.Params.visc_water.offset, abi.Params.seed.offset, abi.Params.water.offset,
                      abi.Params.oil.offset]
    assert out[2] == [abi.Fields.f.offset, abi.Fields.step_counter.offset, abi.Fields.owned.offset]
    assert out[3] == [abi.Params.closure.offset, abi.Params.cs.offset, abi.Params.tau_floor.offset,
                      abi.Fields.mlp.offset]
    assert out[4] == [abi.ABI_VERSION, abi.MLP_PARAMS, abi.CLOSURE_LEARNED]
    assert out[5] == [abi.Params.flux_coeff.offset]


def test_reference_default_constants():
    # constants.cuh:70-101, globalFunctions.cuh:170-283
    p = cases.reference_default(19)
    assert (p.nx, p.ny, p.nz) == (128, 256, 256)
    f32 = np.float32
    assert p.visc_water == pytest.approx(float(f32(0.05)) * 13 / 1400, rel=1e-15)
    assert p.visc_oil == pytest.approx(float(f32(0.05)) * 13 / 450, rel=1e-15)
    assert p.sigma == float((f32(0.05) * f32(0.05) * f32(13)) / f32(500))
    assert p.gamma == 0.5
    assert p.water.r2 == 36.0 and p.oil.r2 == 36.0  # diameter 13 -> radius 6
    assert p.water.center_a == 63.5 and p.oil.center_a == 63.5
    assert p.water.center_b == float(f32(0.7) * f32(127.5))
    assert p.oil.center_b == float(f32(0.5) * f32(127.5))
    assert (p.water.phi, p.oil.phi) == (0.0, 1.0)
    assert p.outflow_y == 1 and p.outflow_z == 1 and p.wrap_x == abi.WRAP_REFERENCE and p.wrap_y == 0
    assert abi.load().ssmd_validate(p) == 0


def test_jet_case_constants():
    p = cases.JET_D3Q19_SMALL.params()
    assert p.visc_water == p.visc_oil == pytest.approx(1 / 60)
    assert p.gamma == pytest.approx(0.3)  # 4 (0.8 - 1/2) / 4
    assert p.oil.r2 == 16.0 and p.oil.center_a == 31.5 and p.oil.center_b == 31.5
    assert cases.JET_D3Q27.params().oil.r2 == 64.0
    assert cases.JET_D3Q19_FP16.params().oil.r2 == 144.0
    assert cases.algorithmic_bytes_per_cell(19, abi.POP_F32) == 216
    assert cases.algorithmic_bytes_per_cell(27, abi.POP_F32) == 280
    assert cases.algorithmic_bytes_per_cell(19, abi.POP_F16) == 140
    assert cases.lattice_kernel_bytes_per_cell(27, abi.POP_F32) == 236
    assert cases.step_bytes_per_cell(27, abi.POP_F32) == 256
    assert cases.step_bytes_per_cell(19, abi.POP_F16) == 116


def test_validation_rejects_bad_parameters():
    lib = abi.load()
    p = cases.JET_D3Q19_SMALL.params()
    for field, val in (("q", 15), ("pop_format", 3), ("nx", 4), ("wrap_x", 3)):
        b = p.copy()
        setattr(b, field, val)
        assert lib.ssmd_validate(b) == -1, field
    b = p.copy()
    b.wrap_z = abi.WRAP_TRUE  # periodic z with an inlet on z = 0
    assert lib.ssmd_validate(b) == -1
    b = p.copy()
    b.nx, b.ny, b.nz = 2048, 2048, 1100  # beyond 32-bit cell indices
    assert lib.ssmd_validate(b) == -2
    b.nx, b.ny, b.nz = 1024, 1024, 1024  # 4 GiB float planes: beyond 32-bit byte offsets
    assert lib.ssmd_validate(b) == -2
    b.nz = 1020  # the largest cube-like box that fits
    assert lib.ssmd_validate(b) == 0

    assert lib.ssmd_bytes(p, 1) == 19 * 4 * p.cells
    b = p.copy()
    b.pop_format = abi.POP_F16
    assert lib.ssmd_bytes(b, 1) == 19 * 2 * p.cells and lib.ssmd_bytes(b, 2) == 4 * 4 * p.cells
    # the LEARNED floor keeps tau = 1/2 + 3 nu + tau_t above 1/2 with a 10% margin
    margin = 0.9 * 3 * min(p.visc_water, p.visc_oil)
    for floor, rc in ((0.0, 0), (-0.99 * margin, 0), (-1.01 * margin, -1), (1e-3, -1), (float("nan"), -1)):
        b = p.copy()
        b.tau_floor = floor
        assert lib.ssmd_validate(b) == rc, floor
    # the phase-field flux closure rides on the LEARNED closure's tau_t
    for closure, coeff, rc in ((abi.CLOSURE_NONE, 0.0, 0), (abi.CLOSURE_NONE, 1.0, -1),
                               (abi.CLOSURE_SMAGORINSKY, 1.0, -1), (abi.CLOSURE_LEARNED, 1.0, 0),
                               (abi.CLOSURE_LEARNED, -3.0, 0), (abi.CLOSURE_LEARNED, float("inf"), -1),
                               (abi.CLOSURE_LEARNED, float("nan"), -1)):
        b = p.copy()
        b.closure, b.flux_coeff = closure, coeff
        if closure == abi.CLOSURE_SMAGORINSKY:
            b.cs = 0.1
        assert lib.ssmd_validate(b) == rc, (closure, coeff)


def test_tile_shapes():
    lib = abi.load()
    shapes = []
    for t in range(lib.ssmd_tile_count()):
        tx, ty = C.c_int32(), C.c_int32()
        assert lib.ssmd_tile_shape(t, C.byref(tx), C.byref(ty)) == 0
        shapes.append((tx.value, ty.value))
    assert shapes == [(64, 4), (128, 2), (64, 2), (128, 4), (128, 4)]
    for case in (cases.JET_D3Q19_SMALL, cases.JET_D3Q27):
        for t in range(len(shapes)):
            b = case.params().copy()
            b.tile = t
            assert lib.ssmd_validate(b) == 0, (case.q, t)
    b = cases.JET_D3Q19_SMALL.params().copy()
    b.tile = len(shapes)
    assert lib.ssmd_validate(b) == -1
    assert lib.ssmd_tile_shape(len(shapes), C.byref(C.c_int32()), C.byref(C.c_int32())) == -1


def test_automatic_z_chunk():
    """lz = 0 halves the interior (rounding up) while the launch has fewer
    than 2048 blocks and the chunk exceeds 16 planes: 16 for the 128x2 tile
    on 128x128x512 (64 columns x 32 chunks), 64 for 128x4 on 256x256x1024
    (128 x 16), and 22 -> 11 planes on a box too small to reach 2048."""
    lib = abi.load()
    assert lib.ssmd_auto_lz(cases.JET_D3Q27.params(), 1) == 16
    assert lib.ssmd_auto_lz(cases.JET_D3Q19_FP16.params(), 3) == 64
    assert lib.ssmd_auto_lz(cases.JET_D3Q19_FP16.params(), 0) == 128  # 64x4: 256 columns x 8 chunks
    small = cases.JET_D3Q19_SMALL.with_(nx=20, ny=18, nz=24).params()
    assert lib.ssmd_auto_lz(small, 0) == 11  # 1 x 5 columns, interior 22: one halving
    assert lib.ssmd_auto_lz(cases.JET_D3Q27.params(), lib.ssmd_tile_count()) == abi.EINVAL
    bad = cases.JET_D3Q27.params()
    bad.nx = 2
    assert lib.ssmd_auto_lz(bad, 1) == abi.EINVAL


def test_error_paths_without_device():
    lib = abi.load()
    p = cases.JET_D3Q19_SMALL.params()
    d = abi.Fields()  # no buffers
    assert lib.ssmd_step(p, d, 0, 0, None) == -3
    assert lib.ssmd_init(p, d, None) == -3
    assert lib.ssmd_graph_launch(None, d, 0, 1, None) == -1
    assert lib.ssmd_error_string(-3) == b"fields or graph not ready"
    with pytest.raises(abi.SsmdError):
        abi.check(-2, "x")


def test_bin_skips_a_null_field(tmp_path):
    """PostProcess::bin skips a field whose device pointer is null (normx ..
    ind are declared but never allocated; PostProcess.cuh:64-73): no file, no
    error.  A missing directory name or parameter block is still an error."""
    lib = abi.load()
    p = cases.JET_D3Q19_SMALL.params()
    d = str(tmp_path).encode()
    assert lib.ssmd_write_bin(p, None, d, b"normx", 7, None) == 0
    assert list(tmp_path.iterdir()) == []
    assert lib.ssmd_write_bin(p, None, None, b"normx", 7, None) == abi.EINVAL
    assert lib.ssmd_write_bin(None, None, d, b"normx", 7, None) == abi.EINVAL


def test_info_file_reference_lines(tmp_path):
    """The info file of the reference's own case carries its metadata lines
    (hostFunctions.cuh:63-76): Reynolds 1400 / 450 and Weber 500 come back
    from nu = u D / Re and sigma = u_oil^2 D_oil / We, diameters 13."""
    lib = abi.load()
    p = cases.reference_default(19)
    assert (p.water.diameter, p.oil.diameter) == (13, 13)
    assert lib.ssmd_write_info(p, str(tmp_path).encode(), b"007", b"D3Q19", 100000, 1000) == 0
    lines = (tmp_path / "007_info.txt").read_text().splitlines()
    want = {
        "ID:": "007", "Velocity set:": "D3Q19", "Water velocity:": "0.05", "Oil velocity:": "0.05",
        "Water Reynolds:": "1400", "Oil Reynolds:": "450", "Weber number:": "500",
        "Domain size:": "NX=128, NY=256, NZ=256", "Water diameter:": "D=13", "Oil diameter:": "D=13",
        "Timesteps:": "100000", "Output interval:": "1000",
    }
    got = {}
    for line in lines:
        for key in want:
            if line.startswith(key):
                got[key] = line[len(key):].strip()
    assert got == want
    assert lines[0].startswith("---") and lines[-1].startswith("---")
    # a single oil jet: no water jet (Reynolds 0), Re = u D / nu = 0.05 * 8 / (1/60) = 24
    j = cases.JET_D3Q19_SMALL.params()
    assert lib.ssmd_write_info(j, str(tmp_path).encode(), b"jet", b"D3Q19", 10, 5) == 0
    text = (tmp_path / "jet_info.txt").read_text()
    assert "Water Reynolds:     0\n" in text and "Oil Reynolds:       24\n" in text
    assert "Weber number:       0.2\n" in text and "Oil diameter:       D=8\n" in text  # 0.05^2 * 8 / 0.1
