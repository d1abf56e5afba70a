"""CPU pins of the closure oracle (oracle/mlp_oracle.cpp, smag_tau in
oracle/ssmd_oracle.cpp).  The reference has no subgrid model; these functions
are pinned by the mathematics that defines them."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

NP = 8961


def mse(params, x, t, quant=False):
    return float(np.mean((O.mlp_forward(params, x, quant) - t) ** 2))


@pytest.fixture(scope="module")
def problem():
    rng = np.random.default_rng(7)
    p = rng.normal(0, 0.25, NP)
    x = rng.normal(0, 1, (37, 8))
    t = rng.normal(0, 1, 37)
    return p, x, t


def test_parameter_count(oracle_lib):
    assert O.load().oracle_mlp_params() == NP


def test_unquantized_gradient_matches_central_differences(problem):
    p, x, t = problem
    g, loss = O.mlp_grad(p, x, t, quant=False)
    assert loss == pytest.approx(mse(p, x, t), rel=1e-12)
    rng = np.random.default_rng(1)
    # every parameter group: W1, b1, W2, b2, W3, b3, W4, b4
    groups = [(0, 512), (512, 576), (576, 4672), (4672, 4736), (4736, 8832), (8832, 8896), (8896, 8960), (8960, 8961)]
    for a, b in groups:
        for i in rng.choice(np.arange(a, b), size=min(6, b - a), replace=False):
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (mse(pp, x, t) - mse(pm, x, t)) / (2 * h)
            assert g[i] == pytest.approx(fd, rel=2e-5, abs=1e-9), (i, g[i], fd)


def test_output_bias_gradient_is_twice_the_mean_residual():
    # y depends on b4 with slope 1, so d/db4 of mean (y - t)^2 = 2 mean(y - t)
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.2, NP)
    x = rng.normal(0, 1, (64, 8))
    t = rng.normal(0, 1, 64)
    g, _ = O.mlp_grad(p, x, t, quant=False)
    y = O.mlp_forward(p, x, False)
    assert g[8960] == pytest.approx(2 * np.mean(y - t), rel=1e-12)


def test_quantized_network_equals_exact_one_on_bf16_exact_values():
    # Small dyadic weights and inputs keep every product and sum exactly
    # representable in bf16: the rounding points then change nothing.
    rng = np.random.default_rng(5)
    p = np.zeros(NP)
    p[0:512] = rng.integers(-2, 3, 512) * 0.125       # W1
    p[576:4672] = rng.integers(-1, 2, 4096) * 0.0625  # W2
    p[4736:8832] = rng.integers(-1, 2, 4096) * 0.0625  # W3
    p[8896:8960] = rng.integers(-2, 3, 64) * 0.25     # W4
    x = rng.integers(-4, 5, (16, 8)) * 0.5
    t = np.zeros(16)
    assert np.array_equal(O.mlp_forward(p, x, True), O.mlp_forward(p, x, False))
    g1, _ = O.mlp_grad(p, x, t, quant=True)
    g0, _ = O.mlp_grad(p, x, t, quant=False)
    # deltas are not dyadic-exact in general (2/n * residual); compare loosely,
    # the exact-values forward above is the pin of the forward rounding points
    assert np.allclose(g1, g0, rtol=2e-2, atol=1e-3 * np.abs(g0).max())


def test_quantized_forward_rounds_weights_and_inputs():
    rng = np.random.default_rng(9)
    p = rng.normal(0, 0.3, NP)
    x = rng.normal(0, 1, (40, 8))
    yq, y = O.mlp_forward(p, x, True), O.mlp_forward(p, x, False)
    # bf16 keeps 8 significand bits: the quantized network differs, at that level
    rel = np.abs(yq - y).max() / np.abs(y).max()
    assert 1e-5 < rel < 5e-2
    # inputs already bf16-exact and weights rounded beforehand: only the hidden
    # activations are still rounded, so the difference shrinks
    pr = np.array([O.load().oracle_bf16_to_f32(O.load().oracle_f32_to_bf16(v)) for v in p], dtype=np.float64)
    xr = O.bf16_decode(np.array([O.load().oracle_f32_to_bf16(v) for v in x.ravel()], dtype=np.uint16)).reshape(x.shape)
    rel2 = np.abs(O.mlp_forward(pr, xr, True) - O.mlp_forward(pr, xr, False)).max() / np.abs(y).max()
    assert rel2 < rel


def test_smagorinsky_relaxation_time_satisfies_its_definition():
    lib = O.load()
    rng = np.random.default_rng(11)
    for _ in range(50):
        pi = rng.normal(0, 1e-3, 6)
        pi[:3] -= pi[:3].mean()  # traceless, as the kernels' Pi
        rho, tau0, cs = rng.uniform(0.9, 1.1), rng.uniform(0.5005, 0.8), rng.uniform(0.05, 0.2)
        tau = lib.oracle_smag_tau(tau0, pi.ctypes.data_as(C.POINTER(C.c_double)), rho, cs)
        # strain rate from the non-equilibrium moments at the total tau
        s = -pi / (2 * rho * tau / 3)
        s_norm = np.sqrt(2 * (s[0] ** 2 + s[1] ** 2 + s[2] ** 2 + 2 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2)))
        nu_t = (tau - tau0) / 3  # cs^2 (tau - tau0)
        assert nu_t == pytest.approx(cs ** 2 * s_norm, rel=1e-10)
    zero = np.zeros(6)
    assert lib.oracle_smag_tau(0.6, zero.ctypes.data_as(C.POINTER(C.c_double)), 1.0, 0.1) == pytest.approx(0.6)
    big = np.array([1e-3, -1e-3, 0, 2e-3, 0, 0])
    assert lib.oracle_smag_tau(0.6, big.ctypes.data_as(C.POINTER(C.c_double)), 1.0, 0.0) == pytest.approx(0.6)


def test_adam_first_step_is_a_signed_step_of_lr():
    # after bias correction the first Adam step is -lr * g / (|g| + eps)
    rng = np.random.default_rng(2)
    p = rng.normal(0, 1, 100)
    g = rng.normal(0, 1, 100)
    m, v = np.zeros(100), np.zeros(100)
    p1 = p.copy()
    O.adam(p1, g, m, v, 1, 1e-2, eps=1e-12)
    assert np.allclose(p1 - p, -1e-2 * np.sign(g), rtol=1e-9, atol=1e-12)
    # a constant gradient keeps the step at lr (moments track it exactly)
    for s in range(2, 6):
        p2 = p1.copy()
        O.adam(p1, g, m, v, s, 1e-2, eps=1e-12)
        assert np.allclose(p1 - p2, -1e-2 * np.sign(g), rtol=1e-9, atol=1e-12)


def test_weighted_loss_is_linear_in_the_weights(problem):
    p, x, t = problem
    g0, l0 = O.mlp_grad(p, x, t, quant=False)
    g1, l1 = O.mlp_grad(p, x, t, quant=False, w=np.ones(len(t)))
    assert np.array_equal(g0, g1) and l0 == l1
    rng = np.random.default_rng(4)
    wa, wb = rng.uniform(0, 2, len(t)), rng.uniform(0, 2, len(t))
    ga, la = O.mlp_grad(p, x, t, quant=False, w=wa)
    gb, lb = O.mlp_grad(p, x, t, quant=False, w=wb)
    gs, ls = O.mlp_grad(p, x, t, quant=False, w=wa + 2 * wb)
    assert ls == pytest.approx(la + 2 * lb, rel=1e-12)
    assert np.allclose(gs, ga + 2 * gb, rtol=1e-10, atol=1e-14)
    # a zero weight removes the sample
    w = np.ones(len(t))
    w[5] = 0.0
    gz, lz = O.mlp_grad(p, x, t, quant=False, w=w)
    keep = np.arange(len(t)) != 5
    gk, lk = O.mlp_grad(p, x[keep], t[keep], quant=False)
    n, m = len(t), int(keep.sum())
    assert lz * n == pytest.approx(lk * m, rel=1e-12)
    assert np.allclose(gz * n, gk * m, rtol=1e-10, atol=1e-14)


def test_interface_network_is_affine_in_the_gradient_feature():
    """closure_study.interface_network(kappa, beta) is the network
    y = kappa * x7 + beta (x7 = |grad phi-hat| * grad_scale >= 0), exactly,
    through the bf16 quantization of the matrix-core kernels."""
    from oracle.oracle import mlp_forward
    from ssmdlbm_amd.closure_study import interface_network
    rng = np.random.default_rng(3)
    x = rng.standard_normal((64, 8))
    # bf16-representable gradient features, as the lattice kernel writes them
    x[:, 7] = (np.abs(x[:, 7]).astype(np.float32).view(np.uint32) & 0xFFFF0000).view(np.float32)
    for kappa, beta in ((12.5, -3.0), (-10.62, -1.52), (0.0, 0.25)):
        y = mlp_forward(interface_network(kappa, beta).astype(np.float64), x, quant=True)
        k32, b32 = float(np.float32(kappa)), float(np.float32(beta))  # the parameters are float32
        np.testing.assert_allclose(y, k32 * x[:, 7] + b32, rtol=1e-12, atol=1e-12)


def test_calibrate_command_line_accepts_the_closure_parameters(tmp_path, capsys):
    """The calibration's options parse on any machine; without a GPU the
    study stops right after parsing (it runs only on the device)."""
    import torch

    from ssmdlbm_amd import closure_study
    argv = ["calibrate", "--out", str(tmp_path), "--report", str(tmp_path / "eval_report.json"), "--capillary",
            "--sharpening", "--kappa0", "-8.2", "--beta0", "-2.0", "--sigma-step", "0.2", "--gamma-step", "-0.2",
            "--evals", "5"]
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the command would run the study")
    assert closure_study.main(argv) == 1
    assert "runs on the GPU" in capsys.readouterr().err
    with pytest.raises(SystemExit):
        closure_study.main(["calibrate", "--out", str(tmp_path), "--report", "r.json", "--capillary-scale", "2"])


# -- phase-field flux closure --------------------------------------------------
def _flux_box(q, coeff):
    """Periodic 32 x 4 x 4 interior, no flow, no surface tension, no
    sharpening: only the D3Q7 diffusion and the closure flux act on phi."""
    from ssmdlbm_amd import abi, cases
    p = cases.closed_box(6, q=q, sigma=0.0, gamma=0.0)
    p.nx = 34
    p.closure = abi.CLOSURE_LEARNED
    p.tau_scale = 1.0
    p.flux_coeff = coeff
    return p


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("coeff,tau_t", [(0.0, 0.2), (1.5, 0.1), (2.0, -0.12), (30.0, -0.03)])
def test_flux_closure_is_gradient_diffusion(q, coeff, tau_t):
    """With u = 0 and gamma = 0 the D3Q7 update (populations w_k (phi +
    c_k . M) leaving each cell, w_0 = 1/4, w_k = 1/8) with M = -4 D grad phi
    and the isotropic gradient (phi(x+1) - phi(x-1)) / 2 maps the periodic
    mode sin(k x) to lambda sin(k x), lambda = 1 - (1 - cos k) / 4 - D sin^2 k:
    diffusion with mobility 1/8 + D for long waves.  The closure's
    D = max(flux_coeff tau_t / 3, -0.1) is read off the decay over 200
    steps (tau_t held uniform; the last case is clamped at -0.1).  The ghost
    layer's phi is refreshed after each phase update so the stencil sees the
    periodic image of the current field (the step itself refreshes it one
    kernel later, after the collision)."""
    p = _flux_box(q, coeff)
    o = O.Oracle(p, fp64=True, threads=2)
    o.init()
    nz, ny, nx = p.shape
    L = nx - 2
    k = 2.0 * np.pi / L
    xi = np.arange(nx) - 1.0  # interior x' = 0..L-1; ghosts at -1 and L are periodic images
    mode = np.sin(k * xi)[None, None, :] * np.ones(p.shape)
    phi0, amp0 = 0.4, 0.1
    phi = phi0 + amp0 * mode
    w7 = np.array([O.load().oracle_w(7, i) for i in range(7)])
    o.set_pops("g", w7[:, None, None, None] * phi[None])  # arrivals that sum to phi at every cell
    o.set_field("tau_t", np.full(p.shape, tau_t))

    def wrap_ghosts(f):
        f[:, :, 0], f[:, :, -1] = f[:, :, -2], f[:, :, 1]
        f[:, 0, :], f[:, -1, :] = f[:, -2, :], f[:, 1, :]
        f[0], f[-1] = f[-2], f[1]
        return f

    steps = 200
    for s in range(steps):
        o.k_phase()
        o.set_field("phi", wrap_ghosts(o.field("phi")))
        o.k_forces()
        o.k_moments()
        o.k_collide()
        o.k_boundaries(s)
    o.k_phase()
    got = o.field("phi")[1:-1, 1:-1, 1:-1]
    sines = np.sin(k * xi[1:-1])
    amp = np.mean((got - phi0) * sines[None, None, :]) / np.mean(sines ** 2)
    D = max(coeff * tau_t / 3.0, -0.1)
    lam = 1.0 - (1.0 - np.cos(k)) / 4.0 - D * np.sin(k) ** 2
    # the flux enters through the normal grad/(|grad| + 1e-9): relative 1e-9 / |grad| ~ 1e-7 of it here
    assert amp / amp0 == pytest.approx(lam ** steps, rel=1e-6)
    # the mode stays a pure mode, mass is conserved
    np.testing.assert_allclose(got, phi0 + amp * sines[None, None, :] * np.ones_like(got), rtol=0, atol=1e-7)
    assert got.mean() == pytest.approx(phi0, rel=1e-13)
