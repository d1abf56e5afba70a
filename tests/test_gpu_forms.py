"""GPU checks of the batching form (ensemble, autotuner, graphs) and output."""
import ctypes as C
import struct

import numpy as np
import pytest
import torch

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


@pytest.mark.parametrize("q,pop,odd", [(19, abi.POP_F32, False), (27, abi.POP_F32, False), (19, abi.POP_F16, False),
                                       (27, abi.POP_BF16, False), (19, abi.POP_F32, True)])
@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_members_bitwise_equal_solo(gpu, q, pop, odd, graph):
    # 16-bit D3Q27 runs the persistent kernel with two planes of pulls in flight;
    # rows of a multiple of 4 in every member run the quad phase kernel, one
    # odd row length (`odd`) the one-cell batched phase kernel
    members = small_ensemble(q=q, pop=pop)
    if odd:
        members[1] = members[1].with_(nx=members[1].nx + 3)
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


@pytest.mark.slow
def test_seeded_32_jet_ensemble_at_its_own_boxes(gpu):
    """The batching form's workload itself: the 32 seeded D3Q19 jets (boxes
    48x48x96 .. 128x128x384, 68.6 M cells; u, sigma, diameter per jet), tile
    chosen by the autotuner, steps replayed from the ensemble's graphs.
    After 20 steps the smallest jet is within 1e-5 relative L2 of the FP64
    oracle; after 50 the smallest, the largest and the most elongated jet
    are bit-identical to their solo runs (fields and both population sets)."""
    from conftest import cpu_threads
    from oracle.oracle import Oracle

    jets = cases.ensemble(32, seed=4)
    ens = Ensemble([c.params() for c in jets])
    ens.init()
    tile = ens.tune(reps=2)
    print(f"32-jet ensemble, {ens.cells / 1e6:.1f} M cells, tile {tile}", flush=True)
    ens.graph_run(19)
    ens.step(1, macros=True)
    ens.synchronize()
    small = min(range(len(jets)), key=lambda i: jets[i].cells)
    large = max(range(len(jets)), key=lambda i: jets[i].cells)
    elongated = max(range(len(jets)), key=lambda i: jets[i].nz / jets[i].nx)
    assert len({small, large, elongated}) == 3
    o = Oracle(jets[small].params(), fp64=True, threads=cpu_threads())
    o.init()
    o.step(20)
    m = ens.members[small]
    for name in MACRO:
        b = o.field(name)
        e = np.linalg.norm(m.field(name) - b) / np.linalg.norm(b)
        assert e < 1e-5, (jets[small].name, name, e)
    ens.graph_run(29)
    ens.step(1, macros=True)
    ens.synchronize()
    for j in (small, large, elongated):
        c = jets[j]
        solo = Simulation(c.params())
        solo.init()
        solo.step(19)
        solo.step(1, macros=True)
        solo.step(29)
        solo.step(1, macros=True)
        m = ens.members[j]
        for name in MACRO:
            np.testing.assert_array_equal(m.field(name), solo.field(name), err_msg=f"{c.name} {name}")
        np.testing.assert_array_equal(m.pops_source("f"), solo.pops_source("f"), err_msg=c.name)
        np.testing.assert_array_equal(m.g_moments(), solo.g_moments(), err_msg=c.name)
        solo.close()
    ens.close()


@pytest.mark.parametrize("q,pop,nx", [(19, abi.POP_F32, 136), (27, abi.POP_F32, 136), (27, abi.POP_F16, 136),
                                      (19, abi.POP_F16, 136), (19, abi.POP_BF16, 136), (19, abi.POP_F16, 135)])
def test_tile_shapes_agree(gpu, q, pop, nx):
    # two x tiles of 128 (one partial; an odd row length too), partial y
    # tiles for every shape; for 16-bit storage also z chunks of odd length
    # and a partial last chunk (the D3Q27 16-bit kernels keep two planes of
    # pulls in flight, the z loop unrolled by two)
    c = cases.JET_D3Q19_SMALL.with_(nx=nx, ny=30, nz=40, q=q, diameter=8, pop_format=pop)
    lib = abi.load()
    res = []
    runs = [(t, 0) for t in range(lib.ssmd_tile_count())]
    if pop != abi.POP_F32:
        runs += [(3, 5), (0, 7), (3, 1)]
    for tile, lz in runs:
        p = c.params()
        p.tile, p.lz = tile, lz
        sim = Simulation(p)
        sim.init()
        sim.step(7)
        sim.step(1, macros=True)
        res.append((f"tile {tile} lz {lz}", {name: sim.field(name) for name in MACRO}, sim.pops_source("f")))
        sim.close()
    assert len(res) == len(runs)
    for what, fields, f in res[1:]:
        for name in MACRO:
            np.testing.assert_array_equal(fields[name], res[0][1][name], err_msg=f"{what} {name}")
        np.testing.assert_array_equal(f, res[0][2], err_msg=f"{what} f")


def test_ensemble_tile_shapes_agree(gpu):
    members = small_ensemble(n=3)
    res = []
    # the last index is the ensemble-only mixed tiling (64 x 8 tiles for the
    # members of at most 64 cells per row: all of these)
    for tile in range(abi.load().ssmd_tile_count() + 1):
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


def test_tuner_step_check_advances_like_steps(gpu):
    """Simulation.tune(confirm_steps) times whole graph steps at the pick and
    at the automatic z-chunk, advancing the simulation; the state it leaves is
    that of the same number of plain steps (launch shapes change no result)."""
    p = cases.JET_D3Q19_SMALL.with_(nx=40, ny=32, nz=96).params()
    lib = abi.load()
    auto = lib.ssmd_auto_lz(p, 1)
    assert auto != 8

    class Picks8:  # the library, with ssmd_tune answering tile 1, lz 8
        def __getattr__(self, name):
            return getattr(lib, name)

        def ssmd_tune(self, p_, d_, reps, stream, tile, lz):
            tile._obj.value, lz._obj.value = 1, 8
            return 0

    a = Simulation(p)
    a.init()
    a.step(5)
    a.lib = Picks8()
    t, lz = a.tune(reps=1, confirm_steps=8)
    a.lib = lib
    assert t == 1 and lz in (8, auto)
    assert a.step_index == 5 + 2 * (16 + 8)
    a.step(3, macros=True)
    b = Simulation(p)
    b.init()
    b.step(a.step_index - 1)
    b.step(1, macros=True)
    for name in MACRO:
        np.testing.assert_array_equal(a.field(name), b.field(name), err_msg=name)
    np.testing.assert_array_equal(a.pops_source("f"), b.pops_source("f"))
    a.close()
    b.close()


def test_ensemble_tune(gpu):
    ens = Ensemble([c.params() for c in small_ensemble(n=3)])
    ens.init()
    t = ens.tune(reps=2)
    assert 0 <= t <= abi.load().ssmd_tile_count()  # the last: the ensemble-only mixed tiling
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


# -- restart -------------------------------------------------------------------
@pytest.mark.parametrize("closure", [abi.CLOSURE_NONE, abi.CLOSURE_LEARNED])
def test_checkpoint_resume_is_bitwise(gpu, tmp_path, closure):
    from ssmdlbm_amd.closure import Network, Scales, init_params
    p = cases.JET_D3Q19_SMALL.with_(nx=24, ny=22, nz=30, diameter=8).params()
    p.closure = closure
    net = None
    if closure == abi.CLOSURE_LEARNED:
        p.strain_scale, p.grad_scale, p.tau_scale = 200.0, 4.0, 0.01
        net = Network(init_params(3), scales=Scales(200.0, 4.0, 0.01))

    def fresh():
        s = Simulation(p)
        if net is not None:
            net.load_into(s)
        s.init()
        return s

    a = fresh()
    a.step(30)
    a.checkpoint(tmp_path / "state.ckpt")
    a.graph_run(20)
    a.step(1, macros=True)
    b = Simulation(p)  # no network loaded: a LEARNED checkpoint carries it
    b.init()
    b.step(3)  # the restore must not depend on what the fields held before
    b.restore(tmp_path / "state.ckpt")
    assert b.step_index == 30
    b.graph_run(20)
    b.step(1, macros=True)
    for name in ("rho", "ux", "uy", "uz", "phi") + (("tau_t",) if net is not None else ()):
        assert np.array_equal(a.field(name), b.field(name)), name
    assert np.array_equal(a.pops_source("f"), b.pops_source("f"))
    assert np.array_equal(a.pops_source("g"), b.pops_source("g"))


def test_checkpoint_rejects_another_case(gpu, tmp_path):
    p = cases.JET_D3Q19_SMALL.with_(nx=24, ny=22, nz=30, diameter=8).params()
    a = Simulation(p)
    a.init()
    a.checkpoint(tmp_path / "s.ckpt")
    q = p.copy()
    q.nz = 32
    b = Simulation(q)
    with pytest.raises(abi.SsmdError):
        b.restore(tmp_path / "s.ckpt")
    # same box, another physical parameter: rejected; another launch shape: accepted
    for field, value in (("sigma", p.sigma * 2), ("visc_oil", p.visc_oil * 1.1), ("seed", p.seed + 1)):
        r = p.copy()
        setattr(r, field, value)
        with pytest.raises(abi.SsmdError):
            Simulation(r).restore(tmp_path / "s.ckpt")
    r = p.copy()
    r.tile, r.lz = 3, 8
    Simulation(r).restore(tmp_path / "s.ckpt")
    # a LEARNED checkpoint: another flux closure coefficient is another case
    pl = p.copy()
    pl.closure, pl.tau_scale = abi.CLOSURE_LEARNED, 0.01
    a = Simulation(pl)
    a.init()
    a.checkpoint(tmp_path / "l.ckpt")
    r = pl.copy()
    r.flux_coeff = 2.0
    with pytest.raises(abi.SsmdError):
        Simulation(r).restore(tmp_path / "l.ckpt")
    Simulation(pl).restore(tmp_path / "l.ckpt")


# -- derived statistics ---------------------------------------------------------
def test_derived_statistics_closed_forms(gpu):
    """Linear velocity u = s_t (a y, b z, c x) + (0, 0, w) and phi = s_t x over
    t = 1..5 steps: the running means are mean(s_t) times the fields, the
    Reynolds moments mean(s_t^2)-weighted products, and the vorticity of the
    last step s_5 (-b, -c, -a) on interior cells, 0 on the ghost layer."""
    from ssmdlbm_amd.solver import DerivedStats
    p = cases.JET_D3Q19_SMALL.with_(nx=14, ny=12, nz=10, diameter=4).params()
    sim = Simulation(p)
    sim.init()
    names = ("avg_phi", "avg_ux", "avg_uz", "avg_uxux", "avg_uxuz", "avg_uyuz", "vort_x", "vort_y", "vort_z",
             "vort_mag")
    ds = DerivedStats(sim, names)
    nz, ny, nx = p.shape
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    a, b, c, w = 1e-3, 2e-3, -1.5e-3, 0.02
    svals = [1.0, 0.5, 1.25, 0.75, 1.1]
    for t, s in enumerate(svals):
        fields = {"ux": s * a * Y, "uy": s * b * Z, "uz": s * c * X + w, "phi": s * X.astype(float)}
        for k, v in fields.items():
            sim.t[k].copy_(torch.tensor(v.reshape(-1), dtype=torch.float32))
        torch.cuda.synchronize()
        sim.step_index = t + 1  # the statistics of step t
        ds.update()
    m1, m2 = np.mean(svals), np.mean(np.square(svals))
    ux, uy, uz = a * Y, b * Z, c * X
    np.testing.assert_allclose(ds.field("avg_phi"), m1 * X, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(ds.field("avg_ux"), m1 * ux, rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(ds.field("avg_uz"), m1 * uz + w, rtol=1e-5, atol=1e-8)
    np.testing.assert_allclose(ds.field("avg_uxux"), m2 * ux * ux, rtol=1e-5, atol=1e-12)
    expect_uxuz = np.mean([s * ux * (s * uz + w) for s in svals], axis=0)
    np.testing.assert_allclose(ds.field("avg_uxuz"), expect_uxuz, rtol=1e-5, atol=1e-11)
    np.testing.assert_allclose(ds.field("avg_uyuz"), np.mean([s * uy * (s * uz + w) for s in svals], axis=0),
                               rtol=1e-5, atol=1e-11)
    inner = np.zeros(p.shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    s = svals[-1]
    for name, val in (("vort_x", -s * b), ("vort_y", -s * c), ("vort_z", -s * a)):
        f = ds.field(name)
        np.testing.assert_allclose(f[inner], val, rtol=1e-4)
        assert not f[~inner].any(), name
    np.testing.assert_allclose(ds.field("vort_mag")[inner], s * np.sqrt(a * a + b * b + c * c), rtol=1e-4)


@pytest.mark.parametrize("q", [19, 27])
def test_f16_store_is_rne_of_the_fp32_result(gpu, q):
    """FP16 population storage rounds each post-collision FP32 value to
    nearest-even, as the reference's __float2half of its float result
    (utils.cuh:77-83): from one f16-representable state, the f16 step's
    stored populations equal RNE(the FP32 step's), bit for bit, at every slot
    (collision and boundary kernels alike)."""
    base = cases.JET_D3Q19_SMALL.with_(nx=32, ny=32, nz=48, diameter=8, q=q, perturb=0.01).params()
    p16 = base.copy()
    p16.pop_format = abi.POP_F16
    s16 = Simulation(p16)
    s16.init()
    s16.step(40)
    s32 = Simulation(base)
    s32.init()
    s16.synchronize()
    s32._current("f").copy_(s16._current("f").float())
    s32._current("g").copy_(s16._current("g"))
    for name, t in s16.t.items():
        if t.dtype == torch.float32 and name in s32.t:
            s32.t[name].copy_(t)
    s32.step_index = s16.step_index
    torch.cuda.synchronize()
    s16.step(1)
    s32.step(1)
    s16.synchronize()
    s32.synchronize()
    got = s16._current("f").view(torch.int16).cpu().numpy()
    want = s32._current("f").half().view(torch.int16).cpu().numpy()  # torch: IEEE RNE
    bad = np.count_nonzero(got != want)
    assert bad == 0, f"{bad} of {got.size} f16 slots differ from RNE of the FP32 step"
    s16.close()
    s32.close()


@pytest.mark.parametrize("fmt", [abi.POP_F16, abi.POP_BF16], ids=["f16", "bf16"])
def test_full_size_16bit_step_is_rne_of_fp32(gpu, fmt):
    """The 16-bit storage case at its own size (256x256x1024, 67 M cells, the
    128x4 tile and 32-plane chunks the tuner picks there), so the paths only
    large boxes take run: non-temporal g stores, 4-plane phase chunks with
    non-temporal phi stores.  From one random 16-bit-representable state in
    every slot (ghosts included), one 16-bit step stores exactly RNE of the
    FP32 step's populations, and the two steps' g moments and phi, computed
    in FP32 from the same values, are bit-identical."""
    p16 = cases.JET_D3Q19_FP16.with_(pop_format=fmt).params()
    p16.tile, p16.lz = 3, 32
    p32 = p16.copy()
    p32.pop_format = abi.POP_F32
    n = p16.cells
    print(f"256x256x1024 format {fmt}: random state, one step each", flush=True)
    s16 = Simulation(p16)
    s16.init()
    s16.synchronize()
    gen = torch.Generator(device=gpu).manual_seed(11)
    f16 = s16._current("f")
    f16.copy_(torch.randn(f16.numel(), generator=gen, device=gpu) * 2e-3)  # cast: RNE
    g = s16._current("g")
    g[:n].copy_(torch.rand(n, generator=gen, device=gpu))
    g[n:].copy_((torch.rand(3 * n, generator=gen, device=gpu) - 0.5) * 0.02)
    s16.t["phi"].copy_(torch.rand(n, generator=gen, device=gpu))
    s32 = Simulation(p32)
    s32.init()
    s32.synchronize()
    s32._current("f").copy_(f16.float())
    s32._current("g").copy_(g)
    for name, t in s16.t.items():
        if t.dtype == torch.float32 and name in s32.t:
            s32.t[name].copy_(t)
    torch.cuda.synchronize()
    s16.step(1)
    s32.step(1)
    s16.synchronize()
    s32.synchronize()
    got = s16._current("f").view(torch.int16)
    want = s32._current("f").to(f16.dtype).view(torch.int16)  # torch: IEEE RNE
    assert int(torch.count_nonzero(got)) > 0.9 * got.numel()
    bad = int((got != want).sum())
    assert bad == 0, f"{bad} of {got.numel()} 16-bit slots differ from RNE of the FP32 step"
    assert torch.equal(s16._current("g"), s32._current("g"))
    assert torch.equal(s16.t["phi"], s32.t["phi"])
    s16.close()
    s32.close()
    del s16, s32, got, want, f16, g
    torch.cuda.empty_cache()
