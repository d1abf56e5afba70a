"""Pins of the CPU oracle (oracle/ssmd_oracle.cpp) -- CPU only.

The reference ships no tests or fixtures, so each oracle function is pinned
by (a) the velocity-set tables extracted from the reference text
(tests/golden/velocity_sets.json), (b) closed-form moment identities that a
wrong coefficient, sign or index would break, (c) IEEE / published known
answers for the storage and noise helpers, and (d) whole-step physics with
closed-form outcomes (rest state, shear-wave decay, conservation,
symmetries).
"""
import json
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from conftest import cpu_threads
from oracle.oracle import Oracle, load
from ssmdlbm_amd import abi, cases
from ssmdlbm_amd.solver import arrival_to_source, phase_moments, phase_populations, source_to_arrival

GOLD = json.loads((Path(__file__).parent / "golden" / "velocity_sets.json").read_text())
RNG = np.random.default_rng(1234)


def lattice(L, q):
    c = np.array([[L.oracle_c(i, a) for a in range(3)] for i in range(q)], dtype=float)
    w = np.array([L.oracle_w(q, i) for i in range(q)])
    return c, w


# -- velocity sets ------------------------------------------------------------
@pytest.mark.parametrize("name,q", [("D3Q7", 7), ("D3Q19", 19), ("D3Q27", 27)])
def test_velocity_sets_match_reference_text(oracle_lib, name, q):
    c, w = lattice(oracle_lib, q)
    assert c.astype(int).tolist() == GOLD[name]["c"]
    np.testing.assert_allclose(w, [float(Fraction(s)) for s in GOLD[name]["w"]], rtol=0, atol=1e-15)
    # package table used by the host driver
    assert [list(d) for d in cases.DIRECTIONS[:q]] == GOLD[name]["c"]


@pytest.mark.parametrize("q,cs2", [(19, 1 / 3), (27, 1 / 3), (7, 1 / 4)])
def test_lattice_isotropy(oracle_lib, q, cs2):
    c, w = lattice(oracle_lib, q)
    assert abs(w.sum() - 1) < 1e-15
    np.testing.assert_allclose(w @ c, 0, atol=1e-16)
    np.testing.assert_allclose(np.einsum("i,ia,ib->ab", w, c, c), cs2 * np.eye(3), atol=1e-15)
    for i in range(1, q, 2):  # opposite pairs (2k-1, 2k)
        np.testing.assert_array_equal(c[i], -c[i + 1])
    if q != 7:
        d = np.eye(3)
        iso = cs2 ** 2 * (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d))
        np.testing.assert_allclose(np.einsum("i,ia,ib,ic,id->abcd", w, c, c, c, c), iso, atol=1e-15)


# -- equilibrium, regularization, forcing -------------------------------------
def _feq(L, q, rho, u):
    return np.array([L.oracle_feq(q, i, rho, *u) for i in range(q)])


@pytest.mark.parametrize("q", [19, 27])
def test_equilibrium_moments(oracle_lib, q):
    c, w = lattice(oracle_lib, q)
    for _ in range(5):
        rho = 1 + 0.1 * RNG.standard_normal()
        u = 0.05 * RNG.standard_normal(3)
        f = _feq(oracle_lib, q, rho, u)  # deviation form fbar = f - w
        assert abs(f.sum() - (rho - 1)) < 1e-14
        np.testing.assert_allclose(c.T @ f, rho * u, atol=1e-14)
        np.testing.assert_allclose(np.einsum("i,ia,ib->ab", f, c, c), rho * np.outer(u, u) + (rho - 1) / 3 * np.eye(3),
                                   atol=1e-14)
        xyz = np.sum(f * c[:, 0] * c[:, 1] * c[:, 2])
        xxy = np.sum(f * c[:, 0] ** 2 * c[:, 1])
        if q == 27:  # third-order Hermite equilibrium
            assert abs(xyz - rho * u[0] * u[1] * u[2]) < 1e-14
            assert abs(xxy - rho * (u[1] / 3 + u[0] ** 2 * u[1])) < 1e-14
        else:  # D3Q19 holds no corner directions and a second-order equilibrium
            assert abs(xyz) < 1e-15
            assert abs(xxy - rho * u[1] / 3) < 1e-14


def _sym(p6):
    xx, yy, zz, xy, xz, yz = p6
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


@pytest.mark.parametrize("q", [19, 27])
def test_regularized_nonequilibrium(oracle_lib, q):
    c, _ = lattice(oracle_lib, q)
    import ctypes as C
    for _ in range(5):
        p6 = 1e-3 * RNG.standard_normal(6)
        tr = p6[:3].sum() / 3
        p6[:3] -= tr  # traceless, as computeMoments leaves it
        u = 0.05 * RNG.standard_normal(3)
        arr = (C.c_double * 6)(*p6)
        f = np.array([oracle_lib.oracle_fneq(q, i, arr, *u) for i in range(q)])
        assert abs(f.sum()) < 1e-17
        np.testing.assert_allclose(c.T @ f, 0, atol=1e-17)
        np.testing.assert_allclose(np.einsum("i,ia,ib->ab", f, c, c), _sym(p6), atol=1e-16)
        xyz = np.sum(f * c[:, 0] * c[:, 1] * c[:, 2])
        xxy = np.sum(f * c[:, 0] ** 2 * c[:, 1])
        if q == 27:
            pxx, pyy, pzz, pxy, pxz, pyz = p6
            assert abs(xyz - (u[0] * pyz + u[1] * pxz + u[2] * pxy)) < 1e-17
            assert abs(xxy - (pxx * u[1] + 2 * u[0] * pxy)) < 1e-17
        else:
            assert abs(xyz) < 1e-18 and abs(xxy) < 1e-18


@pytest.mark.parametrize("q", [19, 27])
def test_guo_force_moments(oracle_lib, q):
    c, _ = lattice(oracle_lib, q)
    for _ in range(5):
        u = 0.05 * RNG.standard_normal(3)
        F = 1e-3 * RNG.standard_normal(3)
        f = np.array([oracle_lib.oracle_force(q, i, *u, *F) for i in range(q)])
        assert abs(f.sum()) < 1e-17
        np.testing.assert_allclose(c.T @ f, F, atol=1e-17)
        np.testing.assert_allclose(np.einsum("i,ia,ib->ab", f, c, c), np.outer(u, F) + np.outer(F, u), atol=1e-17)


def test_phase_equilibrium_and_sharpening(oracle_lib):
    c, _ = lattice(oracle_lib, 7)
    for _ in range(5):
        phi = RNG.uniform(0, 1)
        u = 0.05 * RNG.standard_normal(3)
        g = np.array([oracle_lib.oracle_geq(k, phi, *u) for k in range(7)])
        assert abs(g.sum() - phi) < 1e-15
        np.testing.assert_allclose(c.T @ g, phi * u, atol=1e-16)
        n = RNG.standard_normal(3)
        n /= np.linalg.norm(n)
        sharp = 0.3 * phi * (1 - phi)
        h = np.array([oracle_lib.oracle_antidiff(k, sharp, *n) for k in range(7)])
        assert abs(h.sum()) < 1e-16
        np.testing.assert_allclose(c.T @ h, sharp * n / 4, atol=1e-16)


# -- storage conversions and noise ----------------------------------------------
def test_f16_round_to_nearest_even(oracle_lib):
    L = oracle_lib
    kat = {1.0: 0x3C00, -2.0: 0xC000, 65504.0: 0x7BFF, 65519.0: 0x7BFF, 65520.0: 0x7C00, 2.0 ** -24: 0x0001,
           2.0 ** -25: 0x0000, 3 * 2.0 ** -25: 0x0002, 2.0 ** -14: 0x0400, 1 + 2.0 ** -11: 0x3C00,
           1 + 3 * 2.0 ** -11: 0x3C02, -0.0: 0x8000, float("inf"): 0x7C00}
    for v, h in kat.items():
        assert L.oracle_f32_to_f16(v) == h, (v, hex(L.oracle_f32_to_f16(v)))
    assert L.oracle_f32_to_f16(float("nan")) & 0x7C00 == 0x7C00 and L.oracle_f32_to_f16(float("nan")) & 0x3FF
    # numpy's float16 cast is an independent RNE implementation
    x = np.concatenate([RNG.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** RNG.integers(-9, 5, 20000),
                        (np.arange(-3000, 3000) * 2.0 ** -26).astype(np.float32)]).astype(np.float32)
    mine = np.array([L.oracle_f32_to_f16(float(v)) for v in x], dtype=np.uint16)
    np.testing.assert_array_equal(mine, x.astype(np.float16).view(np.uint16))
    back = np.array([L.oracle_f16_to_f32(int(h)) for h in mine], dtype=np.float32)
    np.testing.assert_array_equal(back, x.astype(np.float16).astype(np.float32))


def test_bf16_round_to_nearest_even(oracle_lib):
    import torch

    L = oracle_lib
    x = (RNG.standard_normal(20000) * 10.0 ** RNG.integers(-30, 30, 20000)).astype(np.float32)
    x = np.concatenate([x, np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -0.0, 3.4e38], dtype=np.float32)])
    mine = np.array([L.oracle_f32_to_bf16(float(v)) for v in x], dtype=np.uint16)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(mine, ref)
    back = np.array([L.oracle_bf16_to_f32(int(h)) for h in mine], dtype=np.float32)
    np.testing.assert_array_equal(back, torch.from_numpy(x).to(torch.bfloat16).float().numpy())


def test_splitmix64_published_stream(oracle_lib):
    # first outputs of SplitMix64 from state 0 (Steele, Lea, Flood 2014)
    g, m = 0x9E3779B97F4A7C15, (1 << 64) - 1
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert [oracle_lib.oracle_splitmix64((k * g) & m) for k in range(3)] == want


def test_noise_is_exact_uniform_grid(oracle_lib):
    vals = np.array([oracle_lib.oracle_noise(7, s, c, a) for s in range(4) for c in range(500) for a in range(3)])
    assert vals.min() >= -1.0 and vals.max() < 1.0
    k = (vals + 1.0) * 2 ** 23
    np.testing.assert_array_equal(k, np.round(k))
    assert abs(vals.mean()) < 0.05 and abs(vals.std() - 1 / np.sqrt(3)) < 0.03
    assert oracle_lib.oracle_noise(7, 1, 2, 0) != oracle_lib.oracle_noise(7, 1, 2, 1)


# -- stencils --------------------------------------------------------------------
def _box(n, q=19, **kw):
    p = cases.closed_box(n, q=q, **kw)
    return p


@pytest.mark.parametrize("q", [19, 27])
def test_forces_vanish_for_linear_phase(q):
    p = _box(12, q=q, sigma=0.1)
    o = Oracle(p, threads=cpu_threads())
    o.init()
    z, y, x = np.meshgrid(*[np.arange(12.0)] * 3, indexing="ij")
    o.set_field("phi", 0.3 + 0.01 * x - 0.02 * y + 0.015 * z)
    o.k_forces()
    inner = (slice(3, -3),) * 3  # normals there see only interior phi
    for n in ("fsx", "fsy", "fsz"):
        assert np.abs(o.field(n)[inner]).max() < 1e-12
    border = o.field("fsx")
    assert np.all(border[0] == 0) and np.all(border[:, :, -1] == 0)


@pytest.mark.parametrize("q", [19, 27])
def test_sphere_curvature(q):
    n, R, W = 40, 9.0, 4.0
    p = _box(n, q=q, sigma=0.01)
    o = Oracle(p, threads=cpu_threads())
    o.init()
    z, y, x = np.meshgrid(*[np.arange(float(n))] * 3, indexing="ij")
    c = (n - 1) / 2
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    phi = 0.5 * (1 - np.tanh(2 * (r - R) / W))
    o.set_field("phi", phi)
    o.k_forces()
    F = np.stack([o.field(k) for k in ("fsx", "fsy", "fsz")])
    # F = -sigma kappa |grad phi| n with n pointing inward: on the shell
    # F_r = -sigma * (-2/R) * |grad phi| * (-1) < 0 ... compare |F| with 2 sigma |grad phi| / R
    shell = np.abs(r - R) < 0.5
    grad = np.abs(-(1 - np.tanh(2 * (r - R) / W) ** 2) / W)  # |d phi / d r|
    ratio = np.linalg.norm(F, axis=0)[shell] / (2 * 0.01 * grad[shell] / R)
    assert abs(np.median(ratio) - 1) < 0.08
    radial = np.einsum("aijk,aijk->ijk", F, np.stack([x - c, y - c, z - c]) / np.maximum(r, 1e-9))
    assert np.all(radial[shell] < 0)  # pulls the interface inward


# -- whole steps -----------------------------------------------------------------
@pytest.mark.parametrize("q", [19, 27])
def test_rest_state_is_fixed_point(q):
    p = _box(10, q=q)
    o = Oracle(p, threads=cpu_threads())
    o.init()
    o.step(20)
    assert np.abs(o.pops("f")).max() == 0.0
    for k in ("ux", "uy", "uz"):
        assert np.abs(o.field(k)).max() == 0.0


@pytest.mark.parametrize("q", [19, 27])
def test_shear_wave_decay(q):
    tau, L, A, T = 0.8, 32, 1e-3, 200
    n = L + 2
    p = _box(n, q=q)
    p.visc_water = p.visc_oil = cases.viscosity_from_tau(tau)
    o = Oracle(p, threads=cpu_threads())
    o.init()
    z = np.arange(n) - 1.0
    ux = A * np.sin(2 * np.pi * z / L)[:, None, None] * np.ones((n, n, n))
    c, _ = lattice(load(), q)
    f = np.stack([np.array([load().oracle_feq(q, i, 1.0, v, 0.0, 0.0) for v in ux[:, 0, 0]])[:, None, None]
                  * np.ones((n, n, n)) for i in range(q)])
    o.set_pops("f", f)
    o.step(T)
    o.step(1)  # one more step materializes the moments of step T
    prof = o.field("ux")[1:-1, n // 2, n // 2]
    amp = 2 / L * np.sum(prof * np.sin(2 * np.pi * z[1:-1] / L))
    nu = (tau - 0.5) / 3
    want = A * np.exp(-nu * (2 * np.pi / L) ** 2 * T)
    assert abs(amp / want - 1) < 0.02, (amp, want)


def _dev(a):
    return a - np.trace(a) / 3 * np.eye(3)


@pytest.mark.parametrize("q", [19, 27])
def test_guo_forcing_uniform_acceleration(q):
    """A uniform flow under a uniform body force F (reference lbm.cu:83-85,
    107 and 253: half the force in u and in Pi, half in the collision).  With
    f = f_eq(u0) at rho = 1 the Guo scheme gives, in closed form,
    u1 = u0 + F/2 and Pi1 = dev(F F)/4 after the first moments pass; one
    collision then adds exactly F to the momentum (u2 = u1 + F) and relaxes
    the stress as Pi2 = (1 - omega) Pi1, the force terms cancelling."""
    L = load()
    n, tau = 8, 0.7
    p = _box(n, q=q, tau=tau, sigma=0.0)
    o = Oracle(p, threads=cpu_threads())
    o.init()
    u0 = np.array([0.04, -0.03, 0.02])
    F = np.array([2e-3, 1e-3, -1.5e-3])
    f = np.array([L.oracle_feq(q, i, 1.0, *u0) for i in range(q)])
    o.set_pops("f", f[:, None, None, None] * np.ones((q,) + o.shape))
    for k, v in zip(("fsx", "fsy", "fsz"), F):
        o.set_field(k, np.full(o.shape, v))
    inner = (slice(2, -2),) * 3  # cells whose pulls all come from interior cells

    def state():
        u = np.array([o.field(k)[inner] for k in ("ux", "uy", "uz")])
        P = {k: o.field(k)[inner] for k in ("pxx", "pyy", "pzz", "pxy", "pxz", "pyz")}
        Pi = np.array([[P["pxx"], P["pxy"], P["pxz"]], [P["pxy"], P["pyy"], P["pyz"]], [P["pxz"], P["pyz"], P["pzz"]]])
        np.testing.assert_allclose(o.field("rho")[inner], 1.0, atol=1e-14)
        return u, Pi

    o.k_moments()
    u1, Pi1 = state()
    np.testing.assert_allclose(u1, (u0 + F / 2)[:, None, None, None] * np.ones_like(u1), atol=1e-15)
    want1 = _dev(np.outer(F, F)) / 4
    np.testing.assert_allclose(Pi1, want1[:, :, None, None, None] * np.ones_like(Pi1), atol=1e-15)
    o.k_collide()
    o.k_moments()
    u2, Pi2 = state()
    np.testing.assert_allclose(u2, (u0 + 3 * F / 2)[:, None, None, None] * np.ones_like(u2), atol=1e-15)
    np.testing.assert_allclose(Pi2, (1 - 1 / tau) * want1[:, :, None, None, None] * np.ones_like(Pi2), atol=1e-15)


@pytest.mark.parametrize("phi", [0.0, 0.5, 1.0])
def test_shear_wave_viscosity_follows_phase(phi):
    """Viscosity is interpolated linearly in phi between the water and oil
    values (reference lbm.cu:229): a shear wave in a uniform phase phi decays
    at nu = nu_w + phi (nu_o - nu_w).  Sharpening is off (gamma = 0): it
    would separate a uniform mixture 0 < phi < 1 from roundoff gradients."""
    tau_w, tau_o, L, A, T = 0.6, 1.0, 32, 1e-3, 200
    n = L + 2
    p = _box(n, q=19, sigma=0.0, gamma=0.0)
    p.visc_water = cases.viscosity_from_tau(tau_w)
    p.visc_oil = cases.viscosity_from_tau(tau_o)
    o = Oracle(p, threads=cpu_threads())
    o.init()
    lib = load()
    o.set_pops("g", np.array([lib.oracle_geq(k, phi, 0, 0, 0) for k in range(7)])[:, None, None, None]
               * np.ones((7,) + o.shape))
    o.set_field("phi", np.full(o.shape, phi))
    z = np.arange(n) - 1.0
    prof0 = A * np.sin(2 * np.pi * z / L)
    f = np.stack([np.array([lib.oracle_feq(19, i, 1.0, v, 0.0, 0.0) for v in prof0]) for i in range(19)])
    o.set_pops("f", f[:, :, None, None] * np.ones((19,) + o.shape))
    o.step(T)
    o.step(1)
    prof = o.field("ux")[1:-1, n // 2, n // 2]
    amp = 2 / L * np.sum(prof * np.sin(2 * np.pi * z[1:-1] / L))
    nu = (tau_w - 0.5) / 3 + phi * ((tau_o - 0.5) / 3 - (tau_w - 0.5) / 3)
    want = A * np.exp(-nu * (2 * np.pi / L) ** 2 * T)
    assert abs(amp / want - 1) < 0.02, (amp, want)
    np.testing.assert_allclose(o.field("phi")[1:-1, 1:-1, 1:-1], phi, atol=1e-12)


def test_interface_sharpening_flux():
    """The phase collision adds the conservative Allen-Cahn sharpening flux
    gamma phi (1 - phi) n (reference lbm.cu:256): on a planar interface at
    rest, the first moment each cell sends out, sum_k c_k g_k, is
    gamma phi (1 - phi) n / 4 with n = grad phi / |grad phi| along +z."""
    lib = load()
    n, gamma = 16, 0.3
    p = _box(n, q=19, sigma=0.0, gamma=gamma)
    o = Oracle(p, threads=cpu_threads())
    o.init()
    zz = np.arange(n, dtype=float)
    prof = 0.5 * (1 + np.tanh((zz - (n - 1) / 2) / 3))
    o.set_field("phi", prof[:, None, None] * np.ones(o.shape))
    o.k_moments()
    o.k_collide()
    c = np.array([[lib.oracle_c(k, a) for a in range(3)] for k in range(7)])
    kp = int(np.flatnonzero((c == [0, 0, 1]).all(1))[0])
    km = int(np.flatnonzero((c == [0, 0, -1]).all(1))[0])
    g = o.pops("g")
    src = np.arange(1, n - 1)
    sent = g[kp][src + 1] - g[km][src - 1]  # c_z-moment sent from planes 1..n-2
    ph = prof[src]
    want = gamma * ph * (1 - ph) / 4
    np.testing.assert_allclose(sent[:, 1:-1, 1:-1], want[:, None, None] * np.ones_like(sent[:, 1:-1, 1:-1]),
                               rtol=1e-6, atol=1e-15)


def _droplet_box(q, n=20, wrap=abi.WRAP_TRUE):
    p = _box(n, q=q, sigma=0.02)
    p.wrap_x = p.wrap_y = p.wrap_z = wrap
    o = Oracle(p, threads=cpu_threads())
    o.init()
    z, y, x = np.meshgrid(*[np.arange(float(n))] * 3, indexing="ij")
    c = (n - 1) / 2
    r = np.sqrt((x - c - 1.3) ** 2 + (y - c) ** 2 + (z - c + 0.7) ** 2)
    phi = 0.5 * (1 - np.tanh(2 * (r - 5) / 4))
    g = np.stack([np.array([load().oracle_geq(k, 1.0, 0, 0, 0) for k in range(7)])[k] * phi for k in range(7)])
    o.set_pops("g", g)
    return o


@pytest.mark.parametrize("q", [19, 27])
def test_true_periodic_box_conserves_mass(q):
    o = _droplet_box(q)
    o.step(1)
    m0 = o.sums()[0]
    rho0 = o.field("rho")[1:-1, 1:-1, 1:-1].sum()
    o.step(150)
    m1 = o.sums()[0]
    assert abs(m1 / m0 - 1) < 1e-12
    assert abs(o.field("rho")[1:-1, 1:-1, 1:-1].sum() / rho0 - 1) < 1e-12
    assert o.field("phi")[1:-1, 1:-1, 1:-1].max() > 0.7  # the droplet survives


def test_reference_periodic_copy_is_not_conservative():
    # periodicX's plane copy (reference BoundaryConditions.cuh:305-332) drops
    # the populations leaving planes 1 and n-2 and duplicates those of 2, n-3
    o = _droplet_box(19, wrap=abi.WRAP_REFERENCE)
    o.step(1)
    m0 = o.sums()[0]
    o.step(60)
    assert abs(o.sums()[0] / m0 - 1) > 1e-6


@pytest.mark.parametrize("wraps", [(1, 1, 1), (1, 1, 0), (0, 1, 1)])
def test_oracle_ghost_wraps_are_sequential_plane_copies(wraps):
    """The oracle's periodic kernels wrap phi and u on the ghost layer as the
    reference does (BoundaryConditions.cuh:335-348): along x, then y, then z,
    [0] <- [n-2] and [n-1] <- [1] on every row of the other two axes, ghost
    rows included -- so a corner column ends up with the doubly wrapped value.
    Restated here with numpy plane copies applied in that order."""
    p = cases.closed_box(12, q=19)
    p.nx, p.ny, p.nz = 13, 11, 9
    p.wrap_x, p.wrap_y, p.wrap_z = (abi.WRAP_REFERENCE if w else 0 for w in wraps)
    o = Oracle(p, fp64=True, threads=1)
    o.init()
    rng = np.random.default_rng(5)
    names = ("phi", "ux", "uy", "uz")
    want = {n: rng.standard_normal(p.shape) for n in names}
    for n in names:
        o.set_field(n, want[n])
    o.k_boundaries(0)
    for axis, on in zip((2, 1, 0), wraps):  # arrays are (z, y, x)
        if not on:
            continue
        for v in want.values():
            n = v.shape[axis]
            v[(slice(None),) * axis + (0,)] = v[(slice(None),) * axis + (n - 2,)]
            v[(slice(None),) * axis + (n - 1,)] = v[(slice(None),) * axis + (1,)]
    for n in names:
        np.testing.assert_array_equal(o.field(n), want[n], err_msg=n)
    # the corner column really is the doubly wrapped interior value
    if wraps[0] and wraps[1]:
        assert o.field("phi")[3, 0, 0] == want["phi"][3, p.ny - 2, p.nx - 2]


def _jet(**kw):
    args = dict(nx=22, ny=22, nz=30, diameter=8, perturb=0.0)
    args.update(kw)
    return cases.JET_D3Q19_SMALL.with_(**args).params()


@pytest.mark.parametrize("q", [19, 27])
def test_jet_mirror_symmetry(q):
    o = Oracle(_jet(q=q), threads=cpu_threads())
    o.init()
    o.step(40)
    f = o.fields()
    np.testing.assert_allclose(f["ux"], -f["ux"][:, :, ::-1], atol=1e-13)
    np.testing.assert_allclose(f["uz"], f["uz"][:, :, ::-1], atol=1e-13)
    np.testing.assert_allclose(f["phi"], f["phi"][:, :, ::-1], atol=1e-13)
    np.testing.assert_allclose(f["uy"], -f["uy"][:, ::-1, :], atol=1e-13)


@pytest.mark.parametrize("q", [19, 27])
def test_jet_xy_transpose_symmetry(q):
    # periodicY is the mirror image of the reference periodicX
    o = Oracle(_jet(q=q), threads=cpu_threads())
    o.init()
    o.step(40)
    f = o.fields()
    t = lambda a: np.swapaxes(a, 1, 2)
    np.testing.assert_allclose(f["ux"], t(f["uy"]), atol=1e-13)
    np.testing.assert_allclose(f["uz"], t(f["uz"]), atol=1e-13)
    np.testing.assert_allclose(f["phi"], t(f["phi"]), atol=1e-13)
    np.testing.assert_allclose(f["rho"], t(f["rho"]), atol=1e-13)


def test_float_oracle_tracks_double():
    p = _jet(perturb=0.01)
    o64, o32 = Oracle(p, fp64=True, threads=cpu_threads()), Oracle(p, fp64=False, threads=cpu_threads())
    for o in (o64, o32):
        o.init()
        o.step(200)
    for k in ("rho", "ux", "uy", "uz", "phi"):
        a, b = o64.field(k), o32.field(k)
        assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-5, k


def test_inlet_noise_enters_and_is_seeded():
    p = _jet(perturb=0.01, seed=5)
    o = Oracle(p, threads=cpu_threads())
    o.init()
    o.step(3)
    ux0 = o.field("ux")[0]
    disc = o.field("phi")[0] == 1.0
    assert np.all(np.abs(ux0[disc]) <= 0.01 * 0.05) and np.abs(ux0[disc]).max() > 0
    q = _jet(perturb=0.01, seed=5)
    o2 = Oracle(q, threads=cpu_threads())
    o2.init()
    o2.step(3)
    np.testing.assert_array_equal(o2.field("ux"), o.field("ux"))


def test_layout_conversion_roundtrip():
    a = RNG.standard_normal((19, 5, 6, 7))
    s = arrival_to_source(a)
    b = source_to_arrival(s)
    ok = ~np.isnan(b)
    np.testing.assert_array_equal(a[ok], b[ok])
    # arrivals whose source is outside are exactly the NaNs
    assert np.isnan(b[1][:, :, 0]).all() and not np.isnan(b[1][:, :, 1:]).any()


def test_phase_moment_form_round_trip_and_rejection(oracle_lib):
    """The 4-plane moment form of the D3Q7 populations: seven populations
    w_k (g_phi + c_k . o) round-trip exactly; the oracle's own post-collision
    populations (equilibrium plus sharpening, reference D3Q7.cuh:124-148) have
    that form; a population set outside it is rejected."""
    rng = np.random.default_rng(8)
    mom = rng.standard_normal((4, 3, 4, 5)).astype(np.float32).astype(np.float64)
    # the populations in exact (double) arithmetic invert exactly
    pops = np.stack([0.25 * mom[0]] + [0.125 * (mom[0] + sg * mom[1 + a]) for a in range(3) for sg in (1, -1)])
    np.testing.assert_array_equal(phase_moments(pops), mom)
    # and phase_populations forms them with one FP32 rounding each
    np.testing.assert_array_equal(phase_populations(mom), pops.astype(np.float32))
    # the oracle's g_eq + anti-diffusion of one cell: g_phi = phi, o = 4 phi u + sharp n
    phi, u, n = 0.3, np.array([0.01, -0.02, 0.03]), np.array([0.6, 0.0, 0.8])
    sharp = 0.4 * phi * (1 - phi)
    g = np.array([oracle_lib.oracle_geq(k, phi, *u) + oracle_lib.oracle_antidiff(k, sharp, *n) for k in range(7)])
    m = phase_moments(g.reshape(7, 1, 1, 1))
    assert m[0].item() == pytest.approx(phi, rel=1e-12)
    np.testing.assert_allclose(m[1:, 0, 0, 0], 4 * phi * u + sharp * n, rtol=1e-12)
    bad = pops.copy()
    bad[1, 1, 1, 1] += 0.01  # one population off the family
    with pytest.raises(ValueError):
        phase_moments(bad)
