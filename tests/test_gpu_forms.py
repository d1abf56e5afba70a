"""GPU checks of the batching form (ensemble, autotuner, graphs) and output."""
import ctypes as C
import struct

import numpy as np
import pytest

from ssmdlbm_amd import abi, cases
from ssmdlbm_amd.solver import Ensemble, Simulation

pytestmark = pytest.mark.gpu

MACRO = ("rho", "ux", "uy", "uz", "phi")


def small_ensemble(n=5, q=19, pop=abi.POP_F32):
    # uneven boxes and parameters, shrunk so solo twins run quickly
    out = []
    for j, c in enumerate(cases.ensemble(n, seed=4, q=q, pop_format=pop)):
        out.append(c.with_(nx=24 + 8 * (j % 3), ny=24 + 8 * (j % 3), nz=40 + 8 * j, diameter=6 + j % 4))
    return out


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_members_bitwise_equal_solo(gpu, q, graph):
    members = small_ensemble(q=q)
    ens = Ensemble([c.params() for c in members])
    ens.init()
    steps = 30
    if graph:
        ens.graph_run(steps - 1)
    else:
        ens.step(steps - 1)
    ens.step(1, macros=True)
    ens.synchronize()
    for c, m in zip(members, ens.members):
        solo = Simulation(c.params())
        solo.init()
        solo.step(steps - 1)
        solo.step(1, macros=True)
        for name in MACRO:
            np.testing.assert_array_equal(m.field(name), solo.field(name), err_msg=f"{c.name} {name}")
        np.testing.assert_array_equal(m.pops_source("f"), solo.pops_source("f"))
        np.testing.assert_array_equal(m.pops_source("g"), solo.pops_source("g"))


def test_ensemble_tile_shapes_agree(gpu):
    members = small_ensemble(n=3)
    res = []
    for tile in range(abi.load().ssmd_tile_count()):
        ens = Ensemble([c.params() for c in members])
        ens.init()
        ens.set_tile(tile)
        ens.graph_run(12)
        ens.synchronize()
        res.append([m.field("uz") for m in ens.members])
        ens.close()
    for r in res[1:]:
        for a, b in zip(res[0], r):
            np.testing.assert_array_equal(a, b)


def test_tuner_is_side_effect_free(gpu):
    p = cases.JET_D3Q19_SMALL.with_(nx=40, ny=32, nz=48).params()
    a = Simulation(p)
    a.init()
    a.step(5)
    t, lz = a.tune(reps=3)
    assert 0 <= t < abi.load().ssmd_tile_count() and lz >= 1
    a.step(10, macros=True)
    b = Simulation(p)
    b.init()
    b.step(15, macros=True)
    for name in MACRO:
        np.testing.assert_array_equal(a.field(name), b.field(name))


def test_ensemble_tune(gpu):
    ens = Ensemble([c.params() for c in small_ensemble(n=3)])
    ens.init()
    t = ens.tune(reps=2)
    assert 0 <= t < abi.load().ssmd_tile_count()
    ens.graph_run(3)


def test_vti_and_bin_output(gpu, tmp_path):
    p = cases.JET_D3Q19_SMALL.with_(nx=20, ny=18, nz=24).params()
    sim = Simulation(p)
    sim.init()
    sim.step(3, macros=True)
    sim.synchronize()
    lib = abi.load()
    d = str(tmp_path).encode()
    for name in ("phi", "rho"):
        abi.check(lib.ssmd_write_bin(C.byref(p), sim.t[name].data_ptr(), d, name.encode(), 3, sim.stream_handle))
    names = (C.c_char_p * 3)(b"phi", b"rho", b"missing")
    abi.check(lib.ssmd_write_vti(C.byref(p), d, 3, names, 3))
    raw = np.fromfile(tmp_path / "phi000003.bin", dtype=np.float32).reshape(p.shape)
    np.testing.assert_array_equal(raw, sim.field("phi").astype(np.float32))
    blob = (tmp_path / "step_000003.vti").read_bytes()
    marker = b'<AppendedData encoding="raw">\n'
    off = blob.index(b"_", blob.index(marker) + len(marker)) + 1
    head = blob[:off].decode()
    assert 'WholeExtent="0 19 0 17 0 23"' in head and 'Name="phi"' in head and 'Name="rho"' in head
    assert 'Name="missing"' not in head
    nbytes = p.cells * 4
    assert struct.unpack("<I", blob[off:off + 4])[0] == nbytes
    np.testing.assert_array_equal(np.frombuffer(blob[off + 4:off + 4 + nbytes], dtype=np.float32), raw.reshape(-1))
    rho = np.frombuffer(blob[off + 8 + nbytes:off + 8 + 2 * nbytes], dtype=np.float32)
    np.testing.assert_array_equal(rho, sim.field("rho").astype(np.float32).reshape(-1))
    assert blob.endswith(b"\n  </AppendedData>\n</VTKFile>\n")
