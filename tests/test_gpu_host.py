"""GPU checks of the native host side: library-owned fields (ssmd_alloc /
ssmd_run / ssmd_free, the reference's FieldAllocate and time loop) and the
C++ driver ssmd_run (the reference's main.cu flow) run as a child process."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

from ssmdlbm_amd import abi, cases
from ssmdlbm_amd.solver import Simulation

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "ssmdlbm_amd" / "_lib" / "ssmd_run"


def small_jet(closure=abi.CLOSURE_NONE, q=19):
    p = cases.JET_D3Q19_SMALL.with_(nx=24, ny=22, nz=30, diameter=8, q=q).params()
    p.closure = closure
    if closure == abi.CLOSURE_LEARNED:
        p.strain_scale, p.grad_scale, p.tau_scale = 200.0, 4.0, 0.01
    return p


@pytest.mark.parametrize("closure,q", [(abi.CLOSURE_NONE, 19), (abi.CLOSURE_NONE, 27), (abi.CLOSURE_LEARNED, 27)])
def test_library_owned_fields_match_the_torch_driver(gpu, tmp_path, closure, q):
    """Fields allocated by ssmd_alloc (zero-filled, LEARNED buffers included),
    initialized by ssmd_init and advanced by ssmd_run + one macro ssmd_step,
    hold the same state, bit for bit, as the torch-allocated Simulation after
    the same 12 steps: their checkpoint files are byte-identical.  Then
    ssmd_free releases them and clears the descriptor; a second free is a no-op."""
    from ssmdlbm_amd.closure import Network, Scales, init_params
    lib = abi.load()
    p = small_jet(closure, q)
    net = None
    if closure == abi.CLOSURE_LEARNED:
        net = Network(init_params(5), scales=Scales(200.0, 4.0, 0.01))
    stream = torch.cuda.Stream()
    s = stream.cuda_stream
    d = abi.Fields()
    abi.check(lib.ssmd_alloc(C.byref(p), C.byref(d)), "ssmd_alloc")
    try:
        assert d.owned and d.rho and d.f and d.g_next and d.step_counter and d.reduce
        assert bool(d.tau_t and d.feat and d.mlp) == (closure == abi.CLOSURE_LEARNED)
        if net is not None:
            torch.cuda.synchronize()
            abi.check(lib.ssmd_mlp_pack(net.params.data_ptr(), d.mlp, s), "ssmd_mlp_pack")
        abi.check(lib.ssmd_init(C.byref(p), C.byref(d), s), "ssmd_init")
        abi.check(lib.ssmd_run(C.byref(p), C.byref(d), 0, 11, s), "ssmd_run")
        abi.check(lib.ssmd_step(C.byref(p), C.byref(d), 11, abi.STEP_MACROS, s), "ssmd_step")
        abi.check(lib.ssmd_checkpoint_write(C.byref(p), C.byref(d), 12, str(tmp_path / "owned.ckpt").encode(), s),
                  "checkpoint")
        mass = (C.c_double * 4)()
        abi.check(lib.ssmd_phase_mass_f64(C.byref(p), C.byref(d), mass, s), "phase_mass")

        sim = Simulation(p)
        if net is not None:
            net.load_into(sim)
        sim.init()
        sim.step(11)
        sim.step(1, macros=True)
        sim.checkpoint(tmp_path / "torch.ckpt")
        a = (tmp_path / "owned.ckpt").read_bytes()
        b = (tmp_path / "torch.ckpt").read_bytes()
        assert len(a) == len(b) and a == b
        np.testing.assert_array_equal(np.array(mass[:]), sim.phase_mass())
        if net is not None:
            assert np.abs(sim.field("tau_t")).max() > 0  # the closure ran on both
        sim.close()
    finally:
        assert lib.ssmd_free(C.byref(d)) == 0
    assert not d.owned and not d.rho and not d.f and not d.mlp
    assert lib.ssmd_free(C.byref(d)) == 0


def test_native_driver_writes_the_reference_file_set(gpu, tmp_path):
    """ssmd_run D3Q19 <ID> --case jet_small --steps 20 --save-every 10: the
    reference's main.cu flow (info file, .bin + .vti at steps 0, 10, 20, the
    timing and MLUPS banner over NSTEPS+1 steps), through the C ABI on
    library-owned fields with graph replay between saves.  phi of the last
    save equals the Python driver's state after the same 21 steps."""
    assert DRIVER.exists(), "the driver is built by __graft_entry__.build()"
    out = subprocess.run([str(DRIVER), "D3Q19", "007", "--case", "jet_small", "--steps", "20", "--save-every", "10",
                          "--out", str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    sim_dir = tmp_path / "D3Q19" / "007"
    info = (sim_dir / "007_info.txt").read_text()
    assert "Velocity set:       D3Q19" in info and "NX=64, NY=64, NZ=128" in info
    assert "Oil Reynolds:       24" in info and "Timesteps:          20" in info
    p = cases.JET_D3Q19_SMALL.params()
    for step in (0, 10, 20):
        for name in ("phi", "rho"):
            assert (sim_dir / f"{name}{step:06d}.bin").stat().st_size == p.cells * 4
        assert (sim_dir / f"step_{step:06d}.vti").stat().st_size > 2 * p.cells * 4
    mlups = re.search(r"Performance\s*:\s*([0-9.eE+-]+) MLUPS", out.stdout)
    assert mlups and float(mlups.group(1)) > 0
    assert re.search(r"Total execution time\s*:\s*[0-9.eE+-]+ s", out.stdout)
    mass = float(re.search(r"Phase mass \(FP64\)\s*:\s*([0-9.eE+-]+)", out.stdout).group(1))
    # the same run through the Python driver: the save at step s follows step index s
    sim = Simulation(p)
    sim.init()
    for s in range(21):
        sim.step(1, macros=s in (0, 10, 20))
    phi = np.fromfile(sim_dir / "phi000020.bin", dtype=np.float32).reshape(p.shape)
    np.testing.assert_array_equal(phi, sim.field("phi").astype(np.float32))
    assert mass == pytest.approx(sim.phase_mass()[0], rel=1e-5)  # printed with 6 digits
    sim.close()
