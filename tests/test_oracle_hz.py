"""Pins the oracle's (Hz, Ex, Ey) 2D family (oracle/fdtd_ref.py update_fields_2d_hz,
power_2d_hz, grid_power_2d_hz and RefSim(polarization="Hz")).

The family has no 2D counterpart in the reference; it is defined as the reference's
3D update (beamz/simulation/fdtd.py:200-310) with every d/dz = 0. The pins:
the middle plane of a z-invariant run of the (separately pinned) 3D restatement,
bit for bit; the PEC-cavity eigenfrequency of the Yee scheme (closed form, with
dx != dy so a swapped divisor fails); the loss-free discrete energy invariant;
and brute-force scalar loops."""
import numpy as np
import pytest

from oracle import fdtd_ref as R

EPS0, MU0 = R.EPS_0, R.MU_0


def test_hz_family_is_the_z_invariant_restriction_of_the_3d_step():
    rng = np.random.default_rng(11)
    ny, nx, steps = 9, 11, 12
    nz = 2 * steps + 8                      # the z walls cannot reach the middle plane in `steps` steps
    dx, dy, dz = 1.0e-7, 1.3e-7, 0.9e-7
    eps2 = 1 + 9 * rng.random((ny, nx))
    sig2 = 3e4 * rng.random((ny, nx))
    dt = 0.5 / (R.LIGHT_SPEED * np.sqrt(1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2))
    F2 = {k: rng.standard_normal(s) * (1.0 if k[0] == "E" else 1 / 377) for k, s in
          R.field_shapes(nx, ny, polarization="Hz").items()}
    # PEC walls of the 2D family: Ex rows 0, ny-1 and Ey columns 0, nx-1 are never updated; start them at 0
    F2["Ex"][[0, -1], :] = 0.0
    F2["Ey"][:, [0, -1]] = 0.0
    shp3 = R.field_shapes(nx, ny, nz)
    F3 = {k: np.zeros(s) for k, s in shp3.items()}
    for k in ("Hz", "Ex", "Ey"):
        F3[k][:] = F2[k][None]
    eps3 = np.broadcast_to(eps2, (nz, ny, nx)).copy()
    sig3 = np.broadcast_to(sig2, (nz, ny, nx)).copy()
    for _ in range(steps):
        R.update_fields_2d_hz(F2, eps2, sig2, dx, dy, dt)
        R.update_fields_3d(F3, eps3, sig3, dx, dy, dz, dt)
    mid = nz // 2
    for k in ("Hz", "Ex", "Ey"):
        assert np.array_equal(F3[k][mid], F2[k]), k
    # the other family stays zero in the middle of the 3D run
    for k in ("Ez", "Hx", "Hy"):
        assert not F3[k][mid].any(), k
    assert np.abs(F2["Hz"]).max() > 0 and np.abs(F2["Ex"][1:-1]).max() > 0


def brute_step_hz(Hz, Ex, Ey, sig, eps, dx, dy, dt):
    ny, nx = Ex.shape[0], Ey.shape[1]
    Hz2, Ex2, Ey2 = Hz.copy(), Ex.copy(), Ey.copy()
    for j in range(ny - 1):
        for i in range(nx - 1):
            s = sig[j, i] * dt / (2 * EPS0)
            a, b = (1 - s) / (1 + s), dt / MU0 / (1 + s)
            curl = (Ey[j, i + 1] - Ey[j, i]) / dx - (Ex[j + 1, i] - Ex[j, i]) / dy
            Hz2[j, i] = a * Hz[j, i] - b * curl
    for j in range(ny):
        for i in range(nx):
            s = sig[j, i] * dt / (2 * EPS0 * eps[j, i])
            a, b = (1 - s) / (1 + s), dt / (EPS0 * eps[j, i]) / (1 + s)
            if 1 <= j <= ny - 2 and i <= nx - 2:
                Ex2[j, i] = a * Ex[j, i] + b * (Hz2[j, i] - Hz2[j - 1, i]) / dy
            if j <= ny - 2 and 1 <= i <= nx - 2:
                Ey2[j, i] = a * Ey[j, i] - b * (Hz2[j, i] - Hz2[j, i - 1]) / dx
    return Hz2, Ex2, Ey2


def test_hz_update_matches_scalar_loops():
    rng = np.random.default_rng(12)
    ny, nx = 7, 9
    dx, dy, dt = 1e-7, 1.3e-7, 1.2e-16
    F = {k: rng.standard_normal(s) for k, s in R.field_shapes(nx, ny, polarization="Hz").items()}
    sig = rng.random((ny, nx)) * 3e4
    eps = 1 + 10 * rng.random((ny, nx))
    want = brute_step_hz(F["Hz"], F["Ex"], F["Ey"], sig, eps, dx, dy, dt)
    R.update_fields_2d_hz(F, eps, sig, dx, dy, dt)
    for k, w in zip(("Hz", "Ex", "Ey"), want):
        np.testing.assert_allclose(F[k], w, rtol=1e-12, atol=1e-12 * np.abs(w).max())


def _recurrence_cos(p):
    s = np.asarray(p)
    r = (s[2:] + s[:-2]) / s[1:-1]
    ok = np.abs(s[1:-1]) > 0.3 * np.abs(s).max()
    return np.median(r[ok]) / 2


@pytest.mark.parametrize("mx, my", [(2, 3), (1, 0), (0, 2)])
def test_hz_pec_cavity_dispersion_closed_form(mx, my):
    """Hz = cos(mx pi (i+1/2)/(nx-1)) cos(my pi (j+1/2)/(ny-1)) is an eigenmode of the discrete system with
    PEC walls (Ex = 0 on rows 0, ny-1, Ey = 0 on columns 0, nx-1): cos(w dt) = 1 - 2 (c dt)^2 lambda,
    lambda = (sin(mx pi / (2 (nx-1))) / dx)^2 + (sin(my pi / (2 (ny-1))) / dy)^2."""
    ny, nx = 21, 31
    dx, dy = 1.0e-7, 1.4e-7
    eps_r = 2.1
    c = 1 / np.sqrt(MU0 * EPS0 * eps_r)
    dt = 0.5 / (c * np.sqrt(1 / dx ** 2 + 1 / dy ** 2))
    mode = np.outer(np.cos(np.pi * my * (np.arange(ny - 1) + 0.5) / (ny - 1)),
                    np.cos(np.pi * mx * (np.arange(nx - 1) + 0.5) / (nx - 1)))
    F = {k: np.zeros(s) for k, s in R.field_shapes(nx, ny, polarization="Hz").items()}
    F["Hz"][:] = mode
    sig = np.zeros((ny, nx)); eps = np.full((ny, nx), eps_r)
    proj = []
    for _ in range(60):
        R.update_fields_2d_hz(F, eps, sig, dx, dy, dt)
        proj.append(np.sum(F["Hz"] * mode))
    lam = (np.sin(np.pi * mx / (2 * (nx - 1))) / dx) ** 2 + (np.sin(np.pi * my / (2 * (ny - 1))) / dy) ** 2
    want = 1 - 2 * (c * dt) ** 2 * lam
    assert abs(_recurrence_cos(proj) - want) < 1e-10


def test_hz_energy_conserved_without_loss():
    """sum eps E^2 + mu0 Hz(n+1/2) Hz(n-1/2) is invariant for sigma = 0."""
    rng = np.random.default_rng(13)
    ny, nx = 24, 30
    dx = dy = 1e-7
    eps = 1 + 3 * rng.random((ny, nx)); sig = np.zeros((ny, nx))
    dt = 0.6 * dx / (R.LIGHT_SPEED * np.sqrt(2))
    F = {k: np.zeros(s) for k, s in R.field_shapes(nx, ny, polarization="Hz").items()}
    F["Hz"][:] = rng.standard_normal(F["Hz"].shape) / 377
    energies = []
    for _ in range(200):
        # before the call: E(n), Hz(n-1/2); after it: Hz(n+1/2)
        hz0, ex, ey = F["Hz"].copy(), F["Ex"].copy(), F["Ey"].copy()
        R.update_fields_2d_hz(F, eps, sig, dx, dy, dt)
        energies.append(EPS0 * (np.sum(eps[:, :-1] * ex * ex) + np.sum(eps[:-1, :] * ey * ey))
                        + MU0 * np.sum(F["Hz"] * hz0))
    e = np.array(energies[1:])
    assert np.abs(e - e[0]).max() < 1e-11 * abs(e[0])


def test_hz_power_formulas():
    # monitor power: |S| = |Hz| hypot(Ex, Ey); grid power: (Ey Hz)^2 + (Ex Hz)^2 without a square root
    pm, tot = R.power_2d_hz([2.0, -1.0], [3.0, 0.0], [4.0, 2.0])
    assert np.array_equal(pm, [10.0, 2.0]) and tot == 12.0
    ny, nx = 4, 5
    Hz = np.zeros((ny - 1, nx - 1)); Ex = np.zeros((ny, nx - 1)); Ey = np.zeros((ny - 1, nx))
    Hz[1, 2], Ex[1, 2], Ey[1, 2] = 2.0, 3.0, 4.0
    Ex[3, 2] = 7.0                            # on the padded row: Hz is 0 there
    p = R.grid_power_2d_hz(Hz, Ex, Ey)
    assert p.shape == (ny, nx) and p[1, 2] == 100.0 and p.sum() == 100.0


def test_refsim_hz_source_clipped_to_hz_cells():
    """The Gaussian patch lands on Hz cells (j, i) with j < ny-1, i < nx-1; at the top-right corner the
    patch of a source centred on the last Ez cell keeps only its in-range part."""
    from types import SimpleNamespace
    d = 1e-7
    ny, nx = 6, 7
    src = SimpleNamespace(position=((nx - 1) * d, (ny - 1) * d), width=d, signal=np.array([1.0]))
    sim = R.RefSim(np.ones((ny, nx)), np.zeros((ny, nx)), d, d, 1e-17, 1, sources=[src], polarization="Hz")
    sim.apply_sources()
    y0, y1, x0, x1, w = R.gaussian_patch_2d(src.position, src.width, d, d, nx, ny)
    assert (y0, y1, x0, x1) == (ny - 4, ny, nx - 4, nx)
    want = np.zeros((ny - 1, nx - 1))
    want[y0:, x0:] = w[:-1, :-1]
    assert np.array_equal(sim.F["Hz"], want)
