"""Pins of the oracle's boundary kernels (inflow, outflow) -- CPU only.

The oracle's inflow()/outflow() restate applyWaterInflow / applyOilInflow
(reference BoundaryConditions.cuh:36-168) and applyOutflowY / applyOutflowZ
(:170-293).  Each test drives one boundary kernel of the FP64 oracle from a
hand-made state and checks what it writes against a closed form:

  * a fluid node at rest with Pi = 0 gives exactly the boundary equilibrium
    fbar_eq(1, u_BC), with no Guo term even where F != 0 (the reference
    reconstructions have none, :94, :163, :228, :288);
  * the non-equilibrium part of every written population is omco * fbar_neq
    with the omco of the face: 1 - 1/(1/2 + 3 nu) of the water viscosity on
    y = 0, of the oil viscosity on z = 0, and of the water viscosity
    (relaxation::omco_ref, globalFunctions.cuh:264-272) on the outflows -- the
    viscosities differ here, so a swapped or rescaled omco fails;
  * hand-evaluated single populations: for the axis direction c = (0,0,+-1) of
    D3Q19 and D3Q27, fbar_eq(1, (0,0,u)) = w (3u +- ... ) and, with traceless
    Pi, fbar_neq = 4.5 w (2/3) Pi_zz - 4.5 w (1/3)(Pi_xx + Pi_yy) = 4.5 w Pi_zz;
  * the one D3Q7 population each face sets is w phi (1 +- 4 u_n) with
    w = 1/8 and cs^2 = 1/4 (D3Q7.cuh:124-132);
  * a uniform equilibrium flowing through the outflow is left unchanged, and
    the outflow's boundary macros are the copy of the plane next to it;
  * the ghost-column rules: the z inlet and the y outflow skip arrivals on a
    ghost column (:145-148, :210-213); the z outflow and the y inlet do not
    check (their geometry keeps them inside the box).

The fbar_eq and fbar_neq used for the composite checks are the oracle's
pointwise functions, themselves pinned by their moment identities in
test_oracle.py.
"""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import Oracle, load
from ssmdlbm_amd import abi, cases

NX, NY, NZ = 12, 11, 14
TAU_WATER, TAU_OIL = 0.6, 0.9
U_OIL, U_WATER = 0.04, 0.03
PHI_JET = 0.7  # the reference's jets carry 0 (water) and 1 (oil); any value exercises the formula


def _omco(tau):
    """1 - omega with omega = 1/(1/2 + 3 nu) and nu = (tau - 1/2)/3."""
    nu = (tau - 0.5) / 3.0
    return 1.0 - 1.0 / (0.5 + 3.0 * nu)


def _params(q, *, oil=False, water=False, out_z=False, out_y=False, oil_center=None):
    p = abi.Params()
    p.nx, p.ny, p.nz = NX, NY, NZ
    p.q = q
    p.visc_water = cases.viscosity_from_tau(TAU_WATER)
    p.visc_oil = cases.viscosity_from_tau(TAU_OIL)
    p.sigma = 0.0
    p.gamma = 0.3
    p.perturb = 0.0
    if oil:
        p.oil.enabled = 1
        ca, cb = oil_center or ((NX - 1) * 0.5, (NY - 1) * 0.5)
        p.oil.center_a, p.oil.center_b, p.oil.r2 = ca, cb, 9.0
        p.oil.u, p.oil.phi = U_OIL, PHI_JET
    if water:
        p.water.enabled = 1
        p.water.center_a, p.water.center_b, p.water.r2 = (NX - 1) * 0.5, 5.0, 4.0
        p.water.u, p.water.phi = U_WATER, PHI_JET
    p.outflow_z = int(out_z)
    p.outflow_y = int(out_y)
    return p


def _dirs(q):
    return np.array(cases.DIRECTIONS[:q])


def _w(q, i):
    return load().oracle_w(q, i)


def _feq(q, i, rho, u):
    return load().oracle_feq(q, i, rho, *u)


def _fneq(q, i, pi6, u):
    return load().oracle_fneq(q, i, (C.c_double * 6)(*pi6), *u)


def _random_state(o, p, rng, *, rest=False):
    """Scrambled populations and force everywhere; Pi traceless and u random
    (or zero, with rest=True) on every cell."""
    shape = p.shape
    o.set_pops("f", rng.standard_normal((p.q,) + shape))
    o.set_pops("g", rng.standard_normal((7,) + shape))
    for name in ("fsx", "fsy", "fsz"):
        o.set_field(name, 1e-3 * rng.standard_normal(shape))
    pi = 1e-3 * rng.standard_normal((6,) + shape)
    pi[:3] -= pi[:3].mean(axis=0)
    if rest:
        pi[:] = 0.0
    for k, name in enumerate(("pxx", "pyy", "pzz", "pxy", "pxz", "pyz")):
        o.set_field(name, pi[k])
    u = np.zeros((3,) + shape) if rest else 0.02 * rng.standard_normal((3,) + shape)
    for k, name in enumerate(("ux", "uy", "uz")):
        o.set_field(name, u[k])
    return pi, u


def _disc(jet, na, nb):
    return [(a, b) for b in range(nb) for a in range(na)
            if (a - jet.center_a) ** 2 + (b - jet.center_b) ** 2 <= jet.r2]


INLETS = {  # face normal axis -> (jet attribute, water/oil tau, arrival maker)
    2: ("oil", TAU_OIL, lambda a, b, c: (a + c[0], b + c[1], 1)),
    1: ("water", TAU_WATER, lambda a, b, c: (a + c[0], 1, b + c[2])),
}


def _inlet_case(q, axis, **kw):
    return _params(q, oil=axis == 2, water=axis == 1, **kw)


def _inlet_writes(p, axis):
    """(disc cell, direction, arrival (x, y, z)) the inlet kernel writes,
    with the reference's ghost-column rule of the z inlet (:145-148)."""
    jet = getattr(p, INLETS[axis][0])
    nb = p.ny if axis == 2 else p.nz
    out = []
    for a, b in _disc(jet, p.nx, nb):
        for i, c in enumerate(_dirs(p.q)):
            if c[axis] != 1:
                continue
            x, y, z = INLETS[axis][2](a, b, c)
            if axis == 2 and (x <= 0 or x >= p.nx - 1 or y <= 0 or y >= p.ny - 1):
                continue
            out.append(((a, b), i, (x, y, z)))
    return out


@pytest.mark.parametrize("axis", [2, 1])
@pytest.mark.parametrize("q", [19, 27])
def test_inlet_from_rest_writes_the_boundary_equilibrium(q, axis):
    """Pi = 0 and u = 0 at the fluid node, F != 0: every written population
    is exactly fbar_eq(1, u_BC); the boundary cell holds rho = 1, u_BC, phi."""
    p = _inlet_case(q, axis)
    o = Oracle(p, fp64=True, threads=1)
    o.init()
    _random_state(o, p, np.random.default_rng(10 + q + axis), rest=True)
    o.k_boundaries(0)
    f = o.pops("f")
    jet = getattr(p, INLETS[axis][0])
    ubc = np.zeros(3)
    ubc[axis] = jet.u
    writes = _inlet_writes(p, axis)
    assert len(writes) > 20
    for _, i, (x, y, z) in writes:
        assert f[i, z, y, x] == pytest.approx(_feq(q, i, 1.0, ubc), rel=0, abs=1e-16)
    # hand value of the axis direction: w (1 + 3u + 4.5u^2 - 1.5u^2) - w = 3 w u (1 + u);
    # the D3Q27 third-order terms cancel on an axis, 4.5 u^3 - 4.5 u^2 u = 0
    k = 5 if axis == 2 else 3
    wk = 1 / 18 if q == 19 else 2 / 27
    hits = [xyz for _, i, xyz in writes if i == k]
    assert hits
    for x, y, z in hits:
        assert f[k, z, y, x] == pytest.approx(3 * wk * jet.u * (1 + jet.u), rel=1e-14)
    # boundary macros on the ghost face
    fld = o.fields(("rho", "ux", "uy", "uz", "phi"))
    nb = p.ny if axis == 2 else p.nz
    for a, b in _disc(jet, p.nx, nb):
        idx = (0, b, a) if axis == 2 else (b, 0, a)
        assert fld["rho"][idx] == 1.0 and fld["phi"][idx] == PHI_JET
        assert (fld["ux"][idx], fld["uy"][idx], fld["uz"][idx]) == tuple(ubc)


@pytest.mark.parametrize("axis", [2, 1])
@pytest.mark.parametrize("q", [19, 27])
def test_inlet_nonequilibrium_uses_the_omco_of_its_phase(q, axis):
    """With Pi and u nonzero at the fluid node: written - fbar_eq(BC) equals
    omco * fbar_neq(Pi, u of the fluid node), omco from the viscosity of the
    phase that inlet carries (water on y = 0, oil on z = 0)."""
    p = _inlet_case(q, axis)
    o = Oracle(p, fp64=True, threads=1)
    o.init()
    pi, u = _random_state(o, p, np.random.default_rng(20 + q + axis))
    o.k_boundaries(0)
    f = o.pops("f")
    jet = getattr(p, INLETS[axis][0])
    ubc = np.zeros(3)
    ubc[axis] = jet.u
    omco = _omco(INLETS[axis][1])
    assert abs(omco - _omco(TAU_WATER if axis == 2 else TAU_OIL)) > 0.1  # the two phases differ
    for _, i, (x, y, z) in _inlet_writes(p, axis):
        neq = f[i, z, y, x] - _feq(q, i, 1.0, ubc)
        want = omco * _fneq(q, i, pi[:, z, y, x], u[:, z, y, x])
        assert neq == pytest.approx(want, rel=1e-12, abs=1e-18)
    # hand value for D3Q19's axis direction: fbar_neq = 4.5 w Pi_nn (traceless Pi)
    if q == 19:
        k = 5 if axis == 2 else 3
        pn = 2 if axis == 2 else 1
        for _, i, (x, y, z) in _inlet_writes(p, axis):
            if i == k:
                neq = f[k, z, y, x] - 3 / 18 * jet.u * (1 + jet.u)
                assert neq == pytest.approx(omco * 4.5 / 18 * pi[pn, z, y, x], rel=1e-12)


@pytest.mark.parametrize("axis", [2, 1])
def test_inlet_phase_population(axis):
    """The inlet sets one D3Q7 population at the first fluid plane:
    g_5(x, y, 1) (z inlet) or g_3(x, 1, z) (y inlet) = (1/8) phi (1 + 4 u)."""
    p = _inlet_case(19, axis)
    o = Oracle(p, fp64=True, threads=1)
    o.init()
    rng = np.random.default_rng(30 + axis)
    _random_state(o, p, rng)
    g0 = o.pops("g")
    o.k_boundaries(0)
    g = o.pops("g")
    jet = getattr(p, INLETS[axis][0])
    k = 5 if axis == 2 else 3
    want = 0.125 * PHI_JET * (1 + 4 * jet.u)
    nb = p.ny if axis == 2 else p.nz
    touched = np.zeros(g.shape, bool)
    for a, b in _disc(jet, p.nx, nb):
        idx = (k, 1, b, a) if axis == 2 else (k, b, 1, a)
        assert g[idx] == pytest.approx(want, rel=1e-15)
        touched[idx] = True
    np.testing.assert_array_equal(g[~touched], g0[~touched])


def test_z_inlet_skips_ghost_columns_and_y_inlet_does_not():
    """A disc against the x = 0 edge: the z inlet leaves arrivals on the ghost
    columns untouched (reference :145-148); inside, it writes."""
    p = _params(19, oil=True, oil_center=(1.0, (NY - 1) * 0.5))
    o = Oracle(p, fp64=True, threads=1)
    o.init()
    rng = np.random.default_rng(40)
    _random_state(o, p, rng, rest=True)
    f0 = o.pops("f")
    o.k_boundaries(0)
    f = o.pops("f")
    ubc = np.array([0.0, 0.0, U_OIL])
    skipped = written = 0
    for a, b in _disc(p.oil, p.nx, p.ny):
        for i, c in enumerate(_dirs(19)):
            if c[2] != 1:
                continue
            x, y = a + c[0], b + c[1]
            if not (0 <= x < p.nx and 0 <= y < p.ny):
                continue
            if x == 0 or y == 0 or y == p.ny - 1:
                assert f[i, 1, y, x] == f0[i, 1, y, x]
                skipped += 1
            else:
                assert f[i, 1, y, x] == pytest.approx(_feq(19, i, 1.0, ubc), abs=1e-16)
                written += 1
    assert skipped > 0 and written > 0


def _outflow_fluid_nodes(p, axis):
    """(boundary cell (x, y, z), direction, fluid node) the outflow writes,
    with the y outflow's ghost-column rule (:210-213)."""
    out = []
    c_all = _dirs(p.q)
    if axis == 2:
        cells = [(x, y, p.nz - 1) for y in range(1, p.ny - 1) for x in range(1, p.nx - 1)]
    else:
        cells = [(x, p.ny - 1, z) for z in range(1, p.nz - 1) for x in range(1, p.nx - 1)]
    for bx, by, bz in cells:
        for i, c in enumerate(c_all):
            if c[axis] != -1:
                continue
            if axis == 2:
                fx, fy, fz = bx + c[0], by + c[1], p.nz - 2
            else:
                fx, fy, fz = bx + c[0], p.ny - 2, bz + c[2]
                if fx <= 0 or fx >= p.nx - 1 or fz <= 0 or fz >= p.nz - 1:
                    continue
            out.append(((bx, by, bz), i, (fx, fy, fz)))
    return out


@pytest.mark.parametrize("axis", [2, 1])
@pytest.mark.parametrize("q", [19, 27])
def test_uniform_flow_through_the_outflow_is_the_identity(q, axis):
    """A uniform equilibrium (fbar = fbar_eq(rho, u) at every arrival, Pi = 0)
    leaves the outflow unchanged: the reconstruction returns the populations
    already there, the boundary macros equal the uniform ones, and the
    D3Q7 population is (1/8) phi (1 - 4 u_n)."""
    p = _params(q, out_z=axis == 2, out_y=axis == 1)
    o = Oracle(p, fp64=True, threads=1)
    o.init()
    rho, u, phi = 1.02, np.array([0.01, -0.02, 0.03]), 0.4
    if axis == 1:
        u = np.array([0.01, 0.03, -0.02])
    shape = p.shape
    feq = np.array([_feq(q, i, rho, u) for i in range(q)])
    o.set_pops("f", np.broadcast_to(feq[:, None, None, None], (q,) + shape))
    for name, v in (("rho", rho), ("ux", u[0]), ("uy", u[1]), ("uz", u[2]), ("phi", phi)):
        o.set_field(name, np.full(shape, v))
    for name in ("pxx", "pyy", "pzz", "pxy", "pxz", "pyz"):
        o.set_field(name, np.zeros(shape))
    o.set_pops("g", np.zeros((7,) + shape))
    o.k_boundaries(0)
    f = o.pops("f")
    np.testing.assert_allclose(f, np.broadcast_to(feq[:, None, None, None], f.shape), rtol=0, atol=1e-17)
    g = o.pops("g")
    k = 6 if axis == 2 else 4
    want = 0.125 * phi * (1 - 4 * u[axis])
    sl = (k, p.nz - 2, slice(1, -1), slice(1, -1)) if axis == 2 else (k, slice(1, -1), p.ny - 2, slice(1, -1))
    np.testing.assert_allclose(g[sl], want, rtol=1e-15)
    assert np.count_nonzero(g) == g[sl].size


@pytest.mark.parametrize("axis", [2, 1])
@pytest.mark.parametrize("q", [19, 27])
def test_outflow_copies_the_plane_and_uses_omco_ref(q, axis):
    """Random state: the boundary macros (rho, phi, u) are copied from the
    plane next to the face on interior rows only, and every written
    population is fbar_eq(copied macros) + omco_ref fbar_neq(fluid node),
    omco_ref being the water viscosity's (oil's differs)."""
    p = _params(q, out_z=axis == 2, out_y=axis == 1)
    o = Oracle(p, fp64=True, threads=1)
    o.init()
    rng = np.random.default_rng(50 + q + axis)
    pi, u = _random_state(o, p, rng)
    shape = p.shape
    rho = 1 + 0.01 * rng.standard_normal(shape)
    phi = rng.uniform(0, 1, shape)
    o.set_field("rho", rho)
    o.set_field("phi", phi)
    o.k_boundaries(0)
    f = o.pops("f")
    fld = o.fields(("rho", "ux", "uy", "uz", "phi"))
    src = {"rho": rho, "ux": u[0], "uy": u[1], "uz": u[2], "phi": phi}
    omco = _omco(TAU_WATER)
    assert abs(omco - _omco(TAU_OIL)) > 0.1
    for name, v in src.items():
        if axis == 2:
            np.testing.assert_array_equal(fld[name][-1, 1:-1, 1:-1], v[-2, 1:-1, 1:-1], err_msg=name)
            np.testing.assert_array_equal(fld[name][-1, 0, :], v[-1, 0, :])  # ghost rows untouched
        else:
            np.testing.assert_array_equal(fld[name][1:-1, -1, 1:-1], v[1:-1, -2, 1:-1], err_msg=name)
            np.testing.assert_array_equal(fld[name][0, -1, :], v[0, -1, :])
    writes = _outflow_fluid_nodes(p, axis)
    assert len(writes) > 50
    ghost_cols = 0
    for (bx, by, bz), i, (fx, fy, fz) in writes:
        ub = np.array([fld["ux"][bz, by, bx], fld["uy"][bz, by, bx], fld["uz"][bz, by, bx]])
        want = _feq(q, i, fld["rho"][bz, by, bx], ub) + omco * _fneq(q, i, pi[:, fz, fy, fx], u[:, fz, fy, fx])
        assert f[i, fz, fy, fx] == pytest.approx(want, rel=1e-12, abs=1e-18)
        ghost_cols += fx in (0, p.nx - 1) or (axis == 2 and fy in (0, p.ny - 1))
    # the z outflow does not check ghost columns (reference :270-283): arrivals there are written
    assert (ghost_cols > 0) == (axis == 2)
    # hand value for D3Q19's axis direction c = -e_n: fbar_neq = 4.5 w Pi_nn
    if q == 19:
        k = 6 if axis == 2 else 4
        for (bx, by, bz), i, (fx, fy, fz) in writes:
            if i != k:
                continue
            ub = np.array([fld["ux"][bz, by, bx], fld["uy"][bz, by, bx], fld["uz"][bz, by, bx]])
            neq = f[k, fz, fy, fx] - _feq(19, k, fld["rho"][bz, by, bx], ub)
            assert neq == pytest.approx(omco * 4.5 / 18 * pi[axis, fz, fy, fx], rel=1e-10)
    g = o.pops("g")
    k = 6 if axis == 2 else 4
    for (bx, by, bz), _, _ in writes[:40]:
        fz, fy = (p.nz - 2, by) if axis == 2 else (bz, p.ny - 2)
        un = fld["uz" if axis == 2 else "uy"][bz, by, bx]
        assert g[k, fz, fy, bx] == pytest.approx(0.125 * fld["phi"][bz, by, bx] * (1 - 4 * un), rel=1e-14)
