"""GPU parity of the subgrid closures against the CPU oracle (through the C ABI).

Tolerances:
  * network forward (BF16 matrix cores, FP32 accumulation) vs the oracle's
    FP64 network with the same bf16 rounding points: max |dy| <= 1e-2 rms(y)
    (an activation may round to the neighbouring bf16 value when FP32 and
    FP64 sums straddle a rounding boundary), median <= 1e-4 rms(y);
  * gradient of the mean squared error: relative L2 <= 1e-2 over the whole
    vector and <= 2e-2 per parameter group; loss within 1e-4;
  * Adam step: update within 1e-5 relative (FP32 vs FP64 arithmetic);
  * Smagorinsky jet, 200 steps: rho, u, phi within 1e-5 relative L2 of FP64;
  * learned closure: the network inputs the step writes are within one bf16
    step (2^-8 relative) of the oracle's unrounded inputs; the step's fields
    within 1e-5 relative L2; tau_t within the forward tolerance.
"""
import numpy as np
import pytest
import torch

from conftest import cpu_threads
from oracle import oracle as O
from oracle.oracle import Oracle
from ssmdlbm_amd import abi, cases
from ssmdlbm_amd.closure import Network, Scales, bf16_bits, bits_to_float, init_params, offsets
from ssmdlbm_amd.solver import Simulation

pytestmark = pytest.mark.gpu

MACRO = ("rho", "ux", "uy", "uz", "phi")
GROUPS = ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def random_params(seed, bias=0.05):
    p = init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    for g in ("b1", "b2", "b3", "b4"):
        a, b = offsets()[g]
        p[a:b] = rng.normal(0, bias, b - a)
    return p.astype(np.float32)


def inputs(n, seed):
    g = torch.Generator().manual_seed(seed)
    return bf16_bits(torch.randn(n, 8, generator=g))


@pytest.mark.parametrize("n", [1, 31, 1000, 40000])
def test_forward_matches_oracle(gpu, n):
    p = random_params(1)
    net = Network(p)
    xb = inputs(n, 2)
    y = net.forward(xb).cpu().numpy().astype(np.float64)
    yo = O.mlp_forward(p.astype(np.float64), bits_to_float(xb).double().numpy(), quant=True)
    rms = np.sqrt(np.mean(yo ** 2))
    err = np.abs(y - yo)
    print(f"n={n}: max {err.max() / rms:.2e} median {np.median(err) / rms:.2e} of rms {rms:.3f}")
    assert err.max() <= 1e-2 * rms
    assert np.median(err) <= 1e-4 * rms


def test_gradient_matches_oracle(gpu):
    p = random_params(3)
    net = Network(p)
    n = 1000
    xb = inputs(n, 4)
    t = torch.randn(n, generator=torch.Generator().manual_seed(5))
    g, loss = net.gradient(xb, t)
    g = g.cpu().numpy().astype(np.float64)
    go, lo = O.mlp_grad(p.astype(np.float64), bits_to_float(xb).double().numpy(), t.double().numpy(), quant=True)
    assert float(loss.item()) == pytest.approx(lo, rel=1e-4)
    total = rel_l2(g, go)
    per = {k: rel_l2(g[a:b], go[a:b]) for k, (a, b) in offsets().items()}
    print(f"gradient rel L2 {total:.2e}; per group {per}")
    assert total <= 1e-2
    assert all(v <= 2e-2 for v in per.values()), per


def test_weighted_gradient_matches_oracle(gpu):
    p = random_params(13)
    net = Network(p)
    n = 777
    xb = inputs(n, 14)
    gen = torch.Generator().manual_seed(15)
    t = torch.randn(n, generator=gen)
    w = torch.rand(n, generator=gen) * 3.0
    g, loss = net.gradient(xb, t, w)
    g = g.cpu().numpy().astype(np.float64)
    go, lo = O.mlp_grad(p.astype(np.float64), bits_to_float(xb).double().numpy(), t.double().numpy(), quant=True,
                        w=w.double().numpy())
    assert float(loss.item()) == pytest.approx(lo, rel=1e-4)
    assert rel_l2(g, go) <= 1e-2


def test_gradient_is_deterministic(gpu):
    net = Network(random_params(6))
    xb = inputs(5000, 7)
    t = torch.randn(5000, generator=torch.Generator().manual_seed(8))
    g1 = net.gradient(xb, t)[0].clone()
    g2 = net.gradient(xb, t)[0].clone()
    assert torch.equal(g1, g2)


def test_adam_matches_oracle(gpu):
    rng = np.random.default_rng(9)
    p = rng.normal(0, 1, abi.MLP_PARAMS).astype(np.float32)
    grad = rng.normal(0, 1, abi.MLP_PARAMS).astype(np.float32)
    m = rng.normal(0, 0.1, abi.MLP_PARAMS).astype(np.float32)
    v = rng.uniform(0, 0.1, abi.MLP_PARAMS).astype(np.float32)
    net = Network(p)
    net.grad.copy_(torch.tensor(grad))
    net.m.copy_(torch.tensor(m))
    net.v.copy_(torch.tensor(v))
    net.adam_steps = 2
    net.adam(lr=1e-3)
    got = net.params.cpu().numpy().astype(np.float64)
    po, mo, vo = p.astype(np.float64), m.astype(np.float64), v.astype(np.float64)
    O.adam(po, grad.astype(np.float64), mo, vo, 3, 1e-3)
    upd, upd_o = got - p, po - p
    assert np.max(np.abs(upd - upd_o)) <= 1e-5 * np.max(np.abs(upd_o)) + 2e-7 * np.max(np.abs(p))


def test_training_fits_a_smooth_target(gpu):
    n = 1 << 15
    xb = inputs(n, 11)
    x = bits_to_float(xb)
    t = 0.5 * torch.relu(x[:, 0] - x[:, 1]) + 0.1 * x[:, 2] ** 2 + 0.05 * x[:, 6]
    net = Network(seed=12)
    losses = net.train(xb, t, steps=400, batch=4096, lr=3e-3, log_every=50)
    print("losses", losses)
    assert losses[-1] < 0.05 * losses[0]
    # the reported loss is the mean squared error of the network as packed
    net.gradient(xb[:4096], t[:4096])
    y = net.forward(xb[:4096]).cpu()
    assert float(net.loss.item()) == pytest.approx(float(((y - t[:4096]) ** 2).mean()), rel=1e-4)


# -- closures inside the lattice step ---------------------------------------
def closure_jet(kind, **kw):
    p = cases.JET_D3Q19_SMALL.with_(nx=24, ny=22, nz=30, diameter=8).params()
    nu = (0.51 - 0.5) / 3.0
    p.visc_water = p.visc_oil = nu
    p.closure = kind
    p.cs = 0.1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("q", [19, 27])
def test_smagorinsky_jet_matches_oracle(gpu, q):
    p = closure_jet(abi.CLOSURE_SMAGORINSKY)
    p.q = q
    steps = 200
    sim = Simulation(p)
    sim.init()
    sim.step(steps, macros=True)
    o = Oracle(p, fp64=True, threads=cpu_threads())
    o.init()
    o.step(steps)
    errs = {k: rel_l2(sim.field(k), o.field(k)) for k in MACRO}
    print(f"D3Q{q} Smagorinsky {steps} steps: {errs}")
    assert max(errs.values()) < 1e-5, errs
    # the closure is active: the same jet without it ends elsewhere
    p0 = closure_jet(abi.CLOSURE_NONE)
    p0.q = q
    plain = Simulation(p0)
    plain.init()
    plain.step(steps, macros=True)
    assert rel_l2(plain.field("uz"), sim.field("uz")) > 1e-4


def learned_case(q=19):
    p = closure_jet(abi.CLOSURE_LEARNED, strain_scale=200.0, grad_scale=4.0, tau_scale=0.02)
    p.q = q
    return p


@pytest.mark.parametrize("q", [19, 27])
def test_learned_closure_inputs_and_update(gpu, q):
    """The LEARNED step on D3Q19 and on D3Q27 (the coarse lattice of the
    learned-closure workload, whose third-order collision reads the same tau)."""
    p = learned_case(q)
    params = random_params(21, bias=0.2)
    net = Network(params, scales=Scales(200.0, 4.0, 0.02))
    sim = Simulation(p)
    sim.init()
    o = Oracle(p, fp64=True, threads=cpu_threads())
    o.init()
    # a few steps on both with tau_t = 0 (the zero network the fields start
    # with), then the trained-like network and a shared random tau_t
    sim.step(20)
    o.step(20)
    sim.synchronize()
    assert not sim.t["tau_t"].any()
    net.load_into(sim)
    rng = np.random.default_rng(22)
    tau_t = rng.uniform(0.0, 0.02, p.cells)
    sim.t["tau_t"].copy_(torch.tensor(tau_t, dtype=torch.float32))
    o.set_field("tau_t", tau_t.astype(np.float32).astype(np.float64))
    torch.cuda.synchronize()
    # one step without the network pass: the step reads tau_t and writes the inputs
    sim.launch_phase()
    sim.launch_lattice(macros=True)
    sim.launch_boundaries(macros=True)
    sim.swap()
    o.step(1)
    sim.synchronize()
    errs = {k: rel_l2(sim.field(k), o.field(k)) for k in MACRO}
    assert max(errs.values()) < 1e-5, errs
    feat = bits_to_float(sim.t["feat"].view(p.cells, 8).cpu()).double().numpy()
    inner = np.zeros(p.shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    for j in range(8):
        ref = o.field(f"feat{j}").reshape(-1)[inner]
        got = feat[inner, j]
        scale = np.abs(ref).max()
        bad = np.abs(got - ref) > 2.0 ** -8 * np.abs(ref) + 1e-5 * scale
        assert scale > 0 and not bad.any(), (j, int(bad.sum()), scale)
    # the network pass: tau_t = max(tau_floor, tau_scale * network(inputs)), floor 0
    sim.closure_update()
    got = sim.field("tau_t").reshape(-1)
    yo = O.mlp_forward(params.astype(np.float64), feat, quant=True)
    want = 0.02 * np.maximum(yo, 0.0)
    rms = np.sqrt(np.mean(want ** 2))
    assert rms > 0
    assert np.abs(got - want).max() <= 1e-2 * rms + 1e-9
    # a negative floor admits backscatter down to the floor
    assert (yo < 0).any()
    sim.p.tau_floor = -0.003
    sim.closure_update()
    got = sim.field("tau_t").reshape(-1)
    want = np.maximum(0.02 * yo, -0.003)
    assert got.min() < 0 and got.min() >= np.float32(-0.003)  # the floor as the kernel holds it
    assert np.abs(got - want).max() <= 1e-2 * rms + 1e-9


@pytest.mark.parametrize("q", [19, 27])
def test_learned_flux_closure_matches_oracle(gpu, q):
    """The phase-field flux closure (flux_coeff = 30: D_t = max(10 tau_t,
    -0.1)) on a LEARNED jet with a shared random tau_t in [-0.015, 0.03],
    so about a tenth of the cells sit at the -0.1 clamp: ten steps without
    the network pass against the FP64 oracle (1e-5 relative L2 on every
    moment and phi).  The same steps with flux_coeff = 0 move phi by ~5%:
    the closure is exercised, not a no-op."""
    p = learned_case(q)
    p.flux_coeff = 30.0
    sim = Simulation(p)
    sim.init()
    o = Oracle(p, fp64=True, threads=cpu_threads())
    o.init()
    sim.step(20)
    o.step(20)
    sim.synchronize()
    tau_t = np.random.default_rng(23).uniform(-0.015, 0.03, p.cells).astype(np.float32)
    sim.t["tau_t"].copy_(torch.tensor(tau_t))
    o.set_field("tau_t", tau_t.astype(np.float64))
    torch.cuda.synchronize()
    for s in range(10):
        last = s == 9
        sim.launch_phase()
        sim.launch_lattice(macros=last)
        sim.launch_boundaries(macros=last)
        sim.swap()
    o.step(10)
    sim.synchronize()
    errs = {k: rel_l2(sim.field(k), o.field(k)) for k in MACRO}
    assert max(errs.values()) < 1e-5, errs
    p0 = p.copy()
    p0.flux_coeff = 0.0
    o0 = Oracle(p0, fp64=True, threads=cpu_threads())
    o0.init()
    o0.step(20)
    o0.set_field("tau_t", tau_t.astype(np.float64))
    o0.step(10)
    assert rel_l2(o0.field("phi"), o.field("phi")) > 1e-2


@pytest.mark.parametrize("flux_coeff", [0.0, 30.0])
def test_learned_closure_graph_replay_matches_eager(gpu, flux_coeff):
    p = learned_case()
    p.flux_coeff = flux_coeff
    net = Network(random_params(31, bias=0.2), scales=Scales(200.0, 4.0, 0.02))
    runs = []
    for graph in (False, True):
        sim = Simulation(p)
        net.load_into(sim)
        sim.init()
        if graph:
            sim.graph_run(25)
        else:
            sim.step(25)
        sim.synchronize()
        runs.append({k: sim.field(k) for k in ("phi", "uz", "tau_t")})
    for k in runs[0]:
        assert np.array_equal(runs[0][k], runs[1][k]), k
    assert np.abs(runs[0]["tau_t"]).max() > 0


# -- training data from a filtered fine field set -----------------------------
@pytest.mark.parametrize("scaling", ["acoustic", "diffusive"])
def test_sgs_samples_analytic(gpu, scaling):
    """Linear fields on the fine grid give closed forms after the 2x box filter:
    u_x = a y, u_z = c x, phi = b x, Pi = diag(p0, -p0, 0), rho = 1; in coarse
    lattice units, with rt = dt_c / dt_f = 4 nu_c / nu_f (2 acoustic, 4
    diffusive) and rv = rt / 2:
      u~_x = rv a (2 cy - 1/2), u~_z = rv c (2 cx - 1/2), phi~ = b (2 cx - 1/2),
      S = rt (-3 Pi / (2 tau_f)),
      T_xx = rv^2 a^2 / 4 (variance of y over a block), T_zz = rv^2 c^2 / 4,
      other T = 0 (x and y vary independently over a block),
      J_z = <phi u_z> - <phi> <u_z> = rv b c / 4 (covariance of x and x), J_x = J_y = 0,
      nu_t = -(dev T : S) / (2 S:S) = rv^2 a^2 / (16 s) with s = 1.5 rt p0 / tau_f
      (T_zz meets S_zz = 0)."""
    from ssmdlbm_amd.closure_study import Filter
    coarse = cases.JET_D3Q19_SMALL.with_(nx=12, ny=10, nz=14, diameter=4).params()
    coarse.oil.enabled = 0  # only the geometry matters here; no inlet near the periodic faces
    coarse.closure = abi.CLOSURE_LEARNED
    coarse.strain_scale, coarse.grad_scale, coarse.tau_scale = 1000.0, 25.0, 0.01
    fine = cases.JET_D3Q19_SMALL.with_(nx=22, ny=18, nz=26, diameter=4).params()
    nu = 2.0 * coarse.visc_water if scaling == "acoustic" else coarse.visc_water
    rt = 2.0 if scaling == "acoustic" else 4.0
    rv = rt / 2
    fine.visc_water = fine.visc_oil = nu
    tau_f = 0.5 + 3 * nu
    a, b, c, p0 = 0.01, 0.02, 0.03, 1e-4
    sim = Simulation(fine)
    nz, ny, nx = fine.shape
    zz, yy, xx = torch.meshgrid(torch.arange(nz), torch.arange(ny), torch.arange(nx), indexing="ij")
    dev = sim.t["rho"].device
    vals = {"rho": torch.ones(fine.shape), "ux": a * yy.float(), "uy": torch.zeros(fine.shape),
            "uz": c * xx.float(), "phi": b * xx.float(), "pxx": torch.full(fine.shape, p0),
            "pyy": torch.full(fine.shape, -p0)}
    for k in ("pzz", "pxy", "pxz", "pyz"):
        vals[k] = torch.zeros(fine.shape)
    for k, v in vals.items():
        sim.t[k].copy_(v.reshape(-1).to(dev))
    torch.cuda.synchronize()
    flt = Filter(fine, coarse)
    flt(sim)
    torch.cuda.synchronize()
    cz, cy, cx = coarse.shape
    s = 1.5 * rt * p0 / tau_f
    inner = (slice(1, cz - 1), slice(1, cy - 1), slice(1, cx - 1))
    Z, Y, X = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
    pl = lambda k: flt.plane(k).cpu().double().numpy()
    np.testing.assert_allclose(pl(0)[inner], 1.0, rtol=1e-6)
    np.testing.assert_allclose(pl(1)[inner], (rv * a * (2 * Y - 0.5))[inner], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(pl(3)[inner], (rv * c * (2 * X - 0.5))[inner], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(pl(4)[inner], (b * (2 * X - 0.5))[inner], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(pl(5)[inner], -s, rtol=1e-5)
    np.testing.assert_allclose(pl(6)[inner], s, rtol=1e-5)
    np.testing.assert_allclose(pl(11)[inner], rv * rv * a * a / 4, rtol=2e-3)  # nearby squares differ in FP32
    np.testing.assert_allclose(pl(13)[inner], rv * rv * c * c / 4, rtol=2e-3)
    np.testing.assert_allclose(pl(19)[inner], rv * b * c / 4, rtol=2e-3)
    for k in (7, 8, 9, 10, 12, 14):
        assert np.abs(pl(k)[inner]).max() < 1e-5 * rv * rv * a * a, k
    # T_xz, T_yz: FP32 differences of block means of u_x u_z ~ a c x y
    for k in (15, 16):
        assert np.abs(pl(k)[inner]).max() < 1e-6 * rv * rv * a * c * fine.nx * fine.ny, k
    for k in (17, 18):
        assert np.abs(pl(k)[inner]).max() < 1e-4 * rv * b * c, k
    target = flt.target.cpu().double().numpy().reshape(coarse.shape)
    weight = flt.weight.cpu().double().numpy().reshape(coarse.shape)
    np.testing.assert_allclose(target[inner], 3 * rv * rv * a * a / (16 * s) / 0.01, rtol=3e-3)
    np.testing.assert_allclose(weight[inner], 2 * s * s, rtol=1e-5)
    ghost = np.ones(coarse.shape, dtype=bool)
    ghost[inner] = False
    assert not weight[ghost].any() and not target[ghost].any()
    feat = bits_to_float(flt.feat.view(-1, 8).cpu()).double().numpy().reshape(cz, cy, cx, 8)
    np.testing.assert_allclose(feat[inner][..., 0], -s * 1000.0, rtol=2 ** -8)
    np.testing.assert_allclose(feat[inner][..., 1], s * 1000.0, rtol=2 ** -8)
    assert not feat[inner][..., 2:6].any()
    np.testing.assert_allclose(feat[inner][..., 6], (b * (2 * X - 0.5))[inner], rtol=2 ** -8)
    # |grad phi-hat| of a linear phi-bar is its slope (2 b per coarse cell) away from the ghost layer
    deep = (slice(2, cz - 2), slice(2, cy - 2), slice(2, cx - 2))
    np.testing.assert_allclose(feat[deep][..., 7], 2 * b * 25.0, rtol=2 ** -8)


@pytest.mark.parametrize("scaling", ["diffusive", "acoustic"])
def test_learned_closure_beats_no_closure_and_smagorinsky_a_priori(gpu, scaling):
    """Filtered 64x64x128 D3Q19 jet -> 32x32x64 samples: a network trained on
    earlier snapshots predicts the subgrid stress of later (held-out)
    snapshots better than no closure (tau_t = 0) and than Smagorinsky
    (tau_t = 3 Cs^2 |S|, Cs = 0.1), in the weighted squared error whose
    minimum is the least-squares fit of the subgrid stress."""
    from ssmdlbm_amd.closure_study import Filter, Geometry, Jet, SampleSet, coarse_params, fine_params
    g = Geometry(nx=32, ny=32, nz=64, q=19, diameter=8, tau=0.55, scaling=scaling)
    jet = Jet(u=0.05, sigma=0.05, seed=3)
    sc = Scales(strain_scale=100.0, grad_scale=4.0, tau_scale=0.01)
    fp, cp = fine_params(g, jet), coarse_params(g, jet, abi.CLOSURE_LEARNED, sc)
    sim = Simulation(fp)
    sim.init()
    flt = Filter(fp, cp)
    train, held = SampleSet(1 << 20, seed=1), SampleSet(1 << 20, seed=2)
    for k, m in enumerate(range(600, 1601, 100)):
        sim.graph_run(g.fine_steps * m - 1 - sim.step_index)
        sim.step(1, macros=True)
        flt(sim)
        (train if k < 8 else held).add(flt)
    x, t, w = train.tensors()
    xh, th, wh = held.tensors()
    wm = w.mean()
    w, wh = w / wm, wh / wm
    net = Network(seed=4, scales=sc)
    net.train(x, t, steps=600, batch=16384, lr=3e-3, lr_final=1e-4, weight=w)
    net.gradient(xh, th, wh)
    learned = float(net.loss.item())
    none = float((wh * th ** 2).mean())
    s = bits_to_float(xh)[:, :6].to(th.device) / sc.strain_scale
    ss = (s[:, :3] ** 2).sum(1) + 2 * (s[:, 3:] ** 2).sum(1)
    smag = 3.0 * 0.1 ** 2 * torch.sqrt(2 * ss) / sc.tau_scale
    smag_err = float((wh * (smag - th) ** 2).mean())
    print(f"held-out weighted MSE: learned {learned:.4g}, no closure {none:.4g}, Smagorinsky {smag_err:.4g}; "
          f"{x.shape[0]} training samples, target mean {float((w * t).mean()):.3f}")
    assert learned < 0.9 * none
    assert learned < 0.9 * smag_err


@pytest.mark.slow
def test_learned_closure_at_the_coarse_box(gpu):
    """The learned-closure workload's coarse run: D3Q27 + D3Q7 on 128x128x512
    at its stated tau = 0.51 + tau_t, with the shipped network and its scales.
    Over 20 steps the fields stay finite, tau_t stays at or above its floor
    (0), and the graph replay
    equals the eager steps bit for bit (fields, tau_t and both population sets)."""
    from pathlib import Path
    import ssmdlbm_amd
    from ssmdlbm_amd.closure_study import Geometry, Jet, coarse_params
    net = Network.load(Path(ssmdlbm_amd.__file__).parent / "data" / "closure_jet_d3q27")
    g = Geometry(tau=0.51)
    p = coarse_params(g, Jet(0.05, 0.05, 5), abi.CLOSURE_LEARNED, net.scales)
    assert (p.nx, p.ny, p.nz, p.q) == (128, 128, 512, 27)
    assert p.visc_water == pytest.approx((0.51 - 0.5) / 3)
    runs = []
    for graph in (False, True):
        sim = Simulation(p)
        net.load_into(sim)
        sim.init()
        print(f"128x128x512 LEARNED, graph={graph}", flush=True)
        if graph:
            sim.graph_run(19)
        else:
            sim.step(19)
        sim.step(1, macros=True)
        sim.synchronize()
        runs.append({k: sim.field(k) for k in ("rho", "ux", "uy", "uz", "phi", "tau_t")})
        runs[-1]["f"] = sim.pops_source("f")
        runs[-1]["g"] = sim.g_moments()
        sim.close()
        del sim
    a, b = runs
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    for k in ("rho", "ux", "uy", "uz", "phi", "tau_t"):
        assert np.isfinite(a[k]).all(), k
    assert a["tau_t"].min() >= 0.0
    print(f"coarse LEARNED: tau_t mean {a['tau_t'].mean():.3e}, max {a['tau_t'].max():.3e}")


def test_shipped_network_runs_a_learned_jet(gpu):
    """The trained network in ssmdlbm_amd/data: loads with its scales,
    installs into a D3Q27 LEARNED jet, and the run stays finite with
    tau_t >= 0 (the closure clamps the network output at 0)."""
    from pathlib import Path
    import ssmdlbm_amd
    from ssmdlbm_amd.closure_study import Geometry, Jet, coarse_params
    path = Path(ssmdlbm_amd.__file__).parent / "data" / "closure_jet_d3q27"
    net = Network.load(path)
    assert net.scales.tau_scale > 0 and net.scales.strain_scale > 0
    g = Geometry(nx=48, ny=48, nz=96, q=27, diameter=12)
    p = coarse_params(g, Jet(0.05, 0.02, 9), abi.CLOSURE_LEARNED, net.scales)
    sim = Simulation(p)
    net.load_into(sim)
    sim.init()
    sim.graph_run(400)
    sim.step(1, macros=True)
    sim.synchronize()
    tau_t = sim.field("tau_t")
    assert np.isfinite(sim.field("uz")).all() and np.isfinite(tau_t).all()
    assert tau_t.min() >= 0.0
    print(f"shipped network: tau_t mean {tau_t.mean():.3e}, max {tau_t.max():.3e}")
