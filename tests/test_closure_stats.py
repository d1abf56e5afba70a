"""The closure study's a-posteriori statistics (ssmdlbm_amd.closure_study
Stats and compare) on closed-form fields -- CPU (torch on the host).

  * centreline: u_z = a z + b on the four axis columns (noise elsewhere) is
    recovered exactly, averaged over samples;
  * interface-area density |grad phi| (central differences on the interior):
    a planar phi = s x + t y + r z gives sqrt(s^2 + t^2 + r^2) where all three
    differences exist and the reduced norms on the faces of the interior block;
  * spectrum: the radial power spectrum of the mean-free density in the slices
    equals a direct numpy restatement (np.fft.fft2, integer radius bins) and
    its bins sum to Parseval's n^2 sum |a - mean|^2 when every radius is kept;
  * compare: identical statistics give 0; scaled ones give the scale error.
"""
import numpy as np
import pytest
import torch

from ssmdlbm_amd import abi
from ssmdlbm_amd.closure_study import Stats, compare


def _params(nx=18, ny=18, nz=12):
    p = abi.Params()
    p.nx, p.ny, p.nz = nx, ny, nz
    return p


def test_centreline_is_the_axis_columns_mean():
    p = _params()
    st = Stats(p, (2, 10), kmax=4, device="cpu")
    nz, ny, nx = p.shape
    z = torch.arange(nz, dtype=torch.float64)[:, None, None]
    rng = np.random.default_rng(1)
    for a, b in ((0.01, 0.02), (0.03, -0.01)):
        uz = torch.from_numpy(rng.standard_normal(p.shape)).float()
        uz[:, 8:10, 8:10] = (a * z + b).float().expand(nz, 2, 2)  # centre (nx-1)/2 = 8.5: columns 8, 9
        st.add(uz, torch.zeros(p.shape))
    got = np.array(st.result()["centreline_uz"])
    want = np.array([0.02 * k + 0.005 for k in range(nz)], dtype=np.float32).astype(np.float64)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-8)
    assert st.result()["samples"] == 2


def test_area_density_of_a_planar_phase():
    p = _params(nx=14, ny=12, nz=10)
    s, t, r = 0.05, -0.02, 0.03
    nz, ny, nx = p.shape
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    phi = torch.from_numpy(s * X + t * Y + r * Z)
    st = Stats(p, (1, nz - 1), kmax=4, device="cpu")
    st.add(torch.zeros(p.shape), phi)
    area = np.array(st.result()["area_per_plane"])
    # the interior block (ny-2) x (nx-2): x differences on its inner columns,
    # y differences on its inner rows, z differences on inner planes
    ix = np.zeros((ny - 2, nx - 2))
    ix[:, 1:-1] = 1
    iy = np.zeros((ny - 2, nx - 2))
    iy[1:-1, :] = 1
    for k in range(nz):
        rz = r if 0 < k < nz - 1 else 0.0
        want = np.sum(np.sqrt((s * ix) ** 2 + (t * iy) ** 2 + rz ** 2))
        assert area[k] == pytest.approx(want, rel=1e-12), k


def _radial_spectrum_numpy(a, z0, z1, kmax):
    """Direct restatement: slices z0..z1-1 of the density, mean removed per
    slice, |fft2|^2 averaged over slices, binned by rounded integer radius."""
    my, mx = a.shape[1:]
    out = np.zeros(kmax + 1)
    for zz in range(z0, z1):
        sl = a[zz] - a[zz].mean()
        pw = np.abs(np.fft.fft2(sl)) ** 2 / (z1 - z0)
        for i in range(my):
            for j in range(mx):
                ky = i if i <= my // 2 - (my % 2 == 0) else i - my  # fftfreq's signed mode number
                kx = j if j <= mx // 2 - (mx % 2 == 0) else j - mx
                r = int(np.rint(np.hypot(kx, ky)))
                if r <= kmax:
                    out[r] += pw[i, j]
    return out


@pytest.mark.parametrize("nx,ny", [(18, 18), (18, 13)])
def test_spectrum_matches_numpy_and_parseval(nx, ny):
    p = _params(nx=nx, ny=ny, nz=9)
    nz, ny, nx = p.shape
    rng = np.random.default_rng(2)
    phi = rng.uniform(0, 1, p.shape)
    kmax = 13  # above the largest radius of a 16 x 16 grid (8 sqrt 2 -> 11)
    st = Stats(p, (2, 7), kmax=kmax, device="cpu")
    st.add(torch.zeros(p.shape), torch.from_numpy(phi))
    got = np.array(st.result()["spectrum"])
    # the same density restated in numpy
    inner = phi[:, 1:-1, 1:-1]
    g = np.zeros((3,) + inner.shape)
    g[0][:, :, 1:-1] = 0.5 * (inner[:, :, 2:] - inner[:, :, :-2])
    g[1][:, 1:-1, :] = 0.5 * (inner[:, 2:, :] - inner[:, :-2, :])
    g[2][1:-1] = 0.5 * (inner[2:] - inner[:-2])
    a = np.sqrt((g ** 2).sum(axis=0))
    want = _radial_spectrum_numpy(a, 2, 7, kmax)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12)
    parseval = np.mean([(nx - 2) * (ny - 2) * np.sum((a[z] - a[z].mean()) ** 2) for z in range(2, 7)])
    assert got.sum() == pytest.approx(parseval, rel=1e-10)
    assert got[0] == pytest.approx(0.0, abs=1e-9)  # mean removed


def test_compare_closed_forms():
    z_range, kmax = (1, 5), 3
    base = {"centreline_uz": [0.0, 1.0, 2.0, 3.0, 4.0, 5.0], "spectrum": [9.0, 4.0, 2.0, 1.0],
            "area_per_plane": [1.0, 2.0, 2.0, 2.0, 2.0, 1.0]}
    same = compare(base, base, z_range, kmax)
    assert same == {"centreline_rel_l2": 0.0, "spectrum_rel_l2": 0.0, "interface_area_rel": 0.0}
    scaled = {"centreline_uz": [1.1 * v for v in base["centreline_uz"]],
              "spectrum": [0.9 * v for v in base["spectrum"]],
              "area_per_plane": [1.05 * v for v in base["area_per_plane"]]}
    e = compare(scaled, base, z_range, kmax)
    assert e["centreline_rel_l2"] == pytest.approx(0.1)
    assert e["spectrum_rel_l2"] == pytest.approx(0.1)  # bins 1..kmax only
    assert e["interface_area_rel"] == pytest.approx(0.05)


class _Planes:
    """A stand-in for the filter: planes 1-3 (u), 4 (phi), 5-10 (S) and 17-19
    (J) as given, zero where not given."""

    def __init__(self, phi, j, u=None, strain=None):
        z = torch.zeros_like(phi)
        self.p = {4: phi, 17: j[0], 18: j[1], 19: j[2]}
        for a in range(3):
            self.p[1 + a] = z if u is None else u[a]
        for k in range(6):
            self.p[5 + k] = z if strain is None else strain[k]

    def plane(self, k):
        return self.p[k]


def test_flux_fit_recovers_gradient_diffusion():
    """FluxFit's constant-D least squares: a subgrid flux that is exactly
    -D grad(phi) (central differences) gives D back with all of |J|^2
    explained; a flux orthogonal to grad(phi) gives D = 0."""
    from ssmdlbm_amd.closure_study import FluxFit
    rng = np.random.default_rng(5)
    n = 12
    phi = torch.from_numpy(rng.uniform(0, 1, (n, n, n)))
    c = lambda a, ax: 0.5 * (torch.roll(a, -1, ax) - torch.roll(a, 1, ax))
    g = [c(phi, 2), c(phi, 1), c(phi, 0)]
    D = 0.037
    fit = FluxFit({})
    fit.add({"x": _Planes(phi, [-D * gi for gi in g])})
    r = fit.result()
    assert r["constant"]["D"] == pytest.approx(D, rel=1e-12)
    assert r["constant"]["explained"] == pytest.approx(1.0, rel=1e-12)
    # J = grad(phi) x e_z is orthogonal to grad(phi) everywhere
    fit = FluxFit({})
    fit.add({"x": _Planes(phi, [g[1], -g[0], torch.zeros_like(phi)])})
    r = fit.result()
    assert abs(r["constant"]["D"]) < 1e-12 and r["constant"]["explained"] < 1e-20


def test_flux_fit_tensor_models():
    """A flux that is exactly C (grad phi . grad) u gives C back for the
    gradient model with all of |J|^2 explained, and one that is exactly
    C' S grad phi does the same for the strain model."""
    from ssmdlbm_amd.closure_study import FluxFit
    rng = np.random.default_rng(6)
    n = 12
    phi = torch.from_numpy(rng.uniform(0, 1, (n, n, n)))
    u = [torch.from_numpy(rng.normal(0, 0.05, (n, n, n))) for _ in range(3)]
    S = [torch.from_numpy(rng.normal(0, 1e-3, (n, n, n))) for _ in range(6)]
    c = lambda a, ax: 0.5 * (torch.roll(a, -1, ax) - torch.roll(a, 1, ax))
    g = [c(phi, 2), c(phi, 1), c(phi, 0)]
    grad_model = [sum(g[b] * c(u[a], 2 - b) for b in range(3)) for a in range(3)]
    Sm = ((S[0], S[3], S[4]), (S[3], S[1], S[5]), (S[4], S[5], S[2]))
    strain_model = [sum(Sm[a][b] * g[b] for b in range(3)) for a in range(3)]
    for name, M, C in (("gradient", grad_model, 0.0625), ("strain", strain_model, -0.4)):
        fit = FluxFit({})
        fit.add({"x": _Planes(phi, [C * m for m in M], u=u, strain=S)})
        r = fit.result()[name]
        assert r["C"] == pytest.approx(C, rel=1e-12), name
        assert r["explained"] == pytest.approx(1.0, rel=1e-12), name
