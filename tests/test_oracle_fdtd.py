"""Pins the FP64 oracle of the FDTD step (oracle/fdtd_ref.py) to closed forms,
invariants and brute force, so that a dropped term, a wrong sign, a wrong
divisor or a shifted index fails here."""
import numpy as np
import pytest

from oracle import fdtd_ref as R
from oracle import signals_ref as S

EPS0, MU0 = R.EPS_0, R.MU_0


def brute_step_2d(Ez, Hx, Hy, sig, eps, dx, dy, dt):
    """Scalar-loop statement of one reference 2D step, written from the update equations."""
    ny, nx = Ez.shape
    Hx2, Hy2, Ez2 = Hx.copy(), Hy.copy(), Ez.copy()
    for j in range(ny):
        for i in range(nx):
            s = sig[j, i] * dt / (2 * EPS0)
            a, b = (1 - s) / (1 + s), dt / MU0 / (1 + s)
            if i < nx - 1:
                Hx2[j, i] = a * Hx[j, i] - b * (Ez[j, i + 1] - Ez[j, i]) / dy
            if j < ny - 1:
                Hy2[j, i] = a * Hy[j, i] + b * (Ez[j + 1, i] - Ez[j, i]) / dx
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            s = sig[j, i] * dt / (2 * EPS0 * eps[j, i])
            a, b = (1 - s) / (1 + s), dt / (EPS0 * eps[j, i]) / (1 + s)
            curl = -(Hx2[j, i] - Hx2[j, i - 1]) / dy + (Hy2[j, i] - Hy2[j - 1, i]) / dx
            Ez2[j, i] = a * Ez[j, i] + b * curl
    return Ez2, Hx2, Hy2


def test_2d_update_matches_scalar_loops():
    rng = np.random.default_rng(1)
    ny, nx = 7, 9
    dx, dy, dt = 1e-7, 1.3e-7, 1.2e-16
    Ez = rng.standard_normal((ny, nx)); Hx = rng.standard_normal((ny, nx - 1)) / 377; Hy = rng.standard_normal((ny - 1, nx)) / 377
    sig = rng.random((ny, nx)) * 3e4; eps = 1 + 10 * rng.random((ny, nx))
    e2, x2, y2 = brute_step_2d(Ez, Hx, Hy, sig, eps, dx, dy, dt)
    hx, hy = R.update_h_fields_2d(Hx.copy(), Hy.copy(), Ez, sig, dx, dy, dt)
    ez = R.update_e_field_2d(Ez.copy(), hx, hy, sig, eps, dx, dy, dt)
    np.testing.assert_allclose(hx, x2, rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(hy, y2, rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(ez, e2, rtol=1e-12, atol=1e-15)


def brute_step_3d(F, eps, sig, dx, dy, dz, dt):
    G = {k: v.copy() for k, v in F.items()}
    nz, ny, nx = eps.shape
    ex, ey, ez = F["Ex"], F["Ey"], F["Ez"]
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                s = sig[k, j, i] * dt / (2 * EPS0)
                a, b = (1 - s) / (1 + s), dt / MU0 / (1 + s)
                if k < nz - 1 and j < ny - 1:
                    c = (ez[k, j + 1, i] - ez[k, j, i]) / dy - (ey[k + 1, j, i] - ey[k, j, i]) / dz
                    G["Hx"][k, j, i] = a * F["Hx"][k, j, i] - b * c
                if k < nz - 1 and i < nx - 1:
                    c = (ex[k + 1, j, i] - ex[k, j, i]) / dz - (ez[k, j, i + 1] - ez[k, j, i]) / dx
                    G["Hy"][k, j, i] = a * F["Hy"][k, j, i] - b * c
                if j < ny - 1 and i < nx - 1:
                    c = (ey[k, j, i + 1] - ey[k, j, i]) / dx - (ex[k, j + 1, i] - ex[k, j, i]) / dy
                    G["Hz"][k, j, i] = a * F["Hz"][k, j, i] - b * c
    hx, hy, hz = G["Hx"], G["Hy"], G["Hz"]
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                s = sig[k, j, i] * dt / (2 * EPS0 * eps[k, j, i])
                a, b = (1 - s) / (1 + s), dt / (EPS0 * eps[k, j, i]) / (1 + s)
                if 1 <= k <= nz - 2 and 1 <= j <= ny - 2 and i <= nx - 2:
                    c = (hz[k, j, i] - hz[k, j - 1, i]) / dy - (hy[k, j, i] - hy[k - 1, j, i]) / dz
                    G["Ex"][k, j, i] = a * ex[k, j, i] + b * c
                if 1 <= k <= nz - 2 and j <= ny - 2 and 1 <= i <= nx - 2:
                    c = (hx[k, j, i] - hx[k - 1, j, i]) / dz - (hz[k, j, i] - hz[k, j, i - 1]) / dx
                    G["Ey"][k, j, i] = a * ey[k, j, i] + b * c
                if k <= nz - 2 and 1 <= j <= ny - 2 and 1 <= i <= nx - 2:
                    c = (hy[k, j, i] - hy[k, j, i - 1]) / dx - (hx[k, j, i] - hx[k, j - 1, i]) / dy
                    G["Ez"][k, j, i] = a * ez[k, j, i] + b * c
    return G


def test_3d_update_matches_scalar_loops():
    rng = np.random.default_rng(2)
    nz, ny, nx = 5, 6, 7
    shp = R.field_shapes(nx, ny, nz)
    F = {k: rng.standard_normal(s) * (1 if k[0] == "E" else 1 / 377) for k, s in shp.items()}
    eps = 1 + 5 * rng.random((nz, ny, nx)); sig = rng.random((nz, ny, nx)) * 1e4
    dx, dy, dz, dt = 2e-8, 2.5e-8, 3e-8, 3e-17
    G = brute_step_3d(F, eps, sig, dx, dy, dz, dt)
    H = {k: v.copy() for k, v in F.items()}
    R.update_fields_3d(H, eps, sig, dx, dy, dz, dt)
    for k in shp:
        np.testing.assert_allclose(H[k], G[k], rtol=1e-12, atol=1e-12 * np.abs(G[k]).max())


def _recurrence_cos(series):
    s = np.asarray(series)
    r = (s[2:] + s[:-2]) / s[1:-1]
    ok = np.abs(s[1:-1]) > 0.3 * np.abs(s).max()
    return np.median(r[ok]) / 2


def test_2d_pec_cavity_dispersion_closed_form():
    """Eigenmode of the discrete 2D system: cos(w dt) from the Yee dispersion relation,
    including the reference's x-difference-over-dy divisor (dx != dy here)."""
    ny, nx = 21, 31
    dx, dy = 1.0e-7, 1.4e-7
    eps_r = 2.1
    c = 1 / np.sqrt(MU0 * EPS0 * eps_r)
    dt = 0.5 / (c * np.sqrt(1 / dx ** 2 + 1 / dy ** 2))
    mx, my = 2, 3
    i = np.arange(nx); j = np.arange(ny)
    shape = np.outer(np.sin(np.pi * my * j / (ny - 1)), np.sin(np.pi * mx * i / (nx - 1)))
    Ez, Hx, Hy = shape.copy(), np.zeros((ny, nx - 1)), np.zeros((ny - 1, nx))
    sig = np.zeros((ny, nx)); eps = np.full((ny, nx), eps_r)
    proj = []
    for _ in range(60):
        Hx, Hy = R.update_h_fields_2d(Hx, Hy, Ez, sig, dx, dy, dt)
        Ez = R.update_e_field_2d(Ez, Hx, Hy, sig, eps, dx, dy, dt)
        proj.append(np.sum(Ez * shape))
    # x-variation sits behind /dy in both curls, y-variation behind /dx (quirk Q2)
    lam = (np.sin(np.pi * mx / (2 * (nx - 1))) / dy) ** 2 + (np.sin(np.pi * my / (2 * (ny - 1))) / dx) ** 2
    want = 1 - 2 * (c * dt) ** 2 * lam          # cos(w dt) = 1 - 2 (c dt)^2 lambda
    assert abs(_recurrence_cos(proj) - want) < 1e-10


@pytest.mark.parametrize("pol", ["z", "x", "y"])
def test_3d_pec_box_dispersion_closed_form(pol):
    nz, ny, nx = 9, 11, 13
    dx, dy, dz = 1e-7, 1.2e-7, 0.9e-7
    eps_r = 1.7
    c = 1 / np.sqrt(MU0 * EPS0 * eps_r)
    dt = 0.5 / (c * np.sqrt(1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2))
    shp = R.field_shapes(nx, ny, nz)
    F = {k: np.zeros(s) for k, s in shp.items()}
    m1, m2 = 2, 1
    if pol == "z":   # Ez(i, j), uniform along z: walls at i=0,nx-1 and j=0,ny-1
        a = np.sin(np.pi * m1 * np.arange(nx) / (nx - 1)); b = np.sin(np.pi * m2 * np.arange(ny) / (ny - 1))
        mode = np.broadcast_to(np.outer(b, a), shp["Ez"]).copy(); key = "Ez"
        lam = (np.sin(np.pi * m1 / (2 * (nx - 1))) / dx) ** 2 + (np.sin(np.pi * m2 / (2 * (ny - 1))) / dy) ** 2
    elif pol == "x":  # Ex(j, k), uniform along x
        a = np.sin(np.pi * m1 * np.arange(ny) / (ny - 1)); b = np.sin(np.pi * m2 * np.arange(nz) / (nz - 1))
        mode = np.broadcast_to(np.outer(b, a)[:, :, None], shp["Ex"]).copy(); key = "Ex"
        lam = (np.sin(np.pi * m1 / (2 * (ny - 1))) / dy) ** 2 + (np.sin(np.pi * m2 / (2 * (nz - 1))) / dz) ** 2
    else:             # Ey(i, k), uniform along y
        a = np.sin(np.pi * m1 * np.arange(nx) / (nx - 1)); b = np.sin(np.pi * m2 * np.arange(nz) / (nz - 1))
        mode = np.broadcast_to(np.outer(b, a)[:, None, :], shp["Ey"]).copy(); key = "Ey"
        lam = (np.sin(np.pi * m1 / (2 * (nx - 1))) / dx) ** 2 + (np.sin(np.pi * m2 / (2 * (nz - 1))) / dz) ** 2
    F[key] = mode.copy()
    eps = np.full((nz, ny, nx), eps_r); sig = np.zeros((nz, ny, nx))
    proj = []
    for _ in range(50):
        R.update_fields_3d(F, eps, sig, dx, dy, dz, dt)
        proj.append(np.sum(F[key] * mode))
    want = 1 - 2 * (c * dt) ** 2 * lam
    assert abs(_recurrence_cos(proj) - want) < 1e-10


def test_3d_h_update_preserves_discrete_divergence():
    rng = np.random.default_rng(3)
    nz, ny, nx = 6, 7, 8
    shp = R.field_shapes(nx, ny, nz)
    F = {k: (rng.standard_normal(s) if k[0] == "E" else np.zeros(s)) for k, s in shp.items()}
    eps = np.ones((nz, ny, nx)); sig = np.zeros((nz, ny, nx))
    dx, dy, dz, dt = 1e-7, 1.1e-7, 0.8e-7, 1e-17
    R.update_fields_3d(F, eps, sig, dx, dy, dz, dt)
    hx, hy, hz = F["Hx"], F["Hy"], F["Hz"]
    div = ((hx[:, :, 1:] - hx[:, :, :-1])[:, :, :] / dx)[:, :, :]
    div = (hx[:nz - 1, :ny - 1, 1:nx - 1] - hx[:nz - 1, :ny - 1, 0:nx - 2]) / dx \
        + (hy[:nz - 1, 1:ny, :nx - 1][:, :ny - 1, :nx - 2] - hy[:nz - 1, 0:ny - 1, :nx - 2]) / dy \
        + (hz[1:nz, :ny - 1, :nx - 2] - hz[0:nz - 1, :ny - 1, :nx - 2]) / dz
    scale = np.abs(hx).max() / dx
    assert np.abs(div).max() < 1e-12 * scale


def test_2d_energy_conserved_without_loss():
    """sum eps E^2 + mu0 H(n+1/2).H(n-1/2) is invariant for sigma = 0 (leapfrog of a skew operator)."""
    rng = np.random.default_rng(4)
    ny, nx = 24, 30
    dx = dy = 1e-7
    eps = 1 + 3 * rng.random((ny, nx)); sig = np.zeros((ny, nx))
    c = 1 / np.sqrt(MU0 * EPS0)
    dt = 0.6 * dx / (c * np.sqrt(2))
    Ez = np.zeros((ny, nx)); Ez[1:-1, 1:-1] = rng.standard_normal((ny - 2, nx - 2))
    Hx, Hy = np.zeros((ny, nx - 1)), np.zeros((ny - 1, nx))
    energies = []
    for _ in range(200):
        Hx0, Hy0 = Hx.copy(), Hy.copy()
        Hx, Hy = R.update_h_fields_2d(Hx, Hy, Ez, sig, dx, dy, dt)
        w = EPS0 * np.sum(eps * Ez * Ez) + MU0 * (np.sum(Hx * Hx0) + np.sum(Hy * Hy0))
        energies.append(w)
        Ez = R.update_e_field_2d(Ez, Hx, Hy, sig, eps, dx, dy, dt)
    energies = np.array(energies[1:])
    assert np.abs(energies - energies[0]).max() < 1e-11 * abs(energies[0])


def test_gaussian_patch_closed_form():
    dx = dy = 1e-7
    y0, y1, x0, x1, w = R.gaussian_patch_2d((5e-7, 3e-7, 0), 1e-7, dx, dy, 20, 20)
    assert (y0, y1, x0, x1) == (0, 7, 2, 9)
    yy, xx = np.mgrid[0:7, 2:9]
    want = np.exp(-((xx - 5) ** 2 + (yy - 3) ** 2) / (2 * (1.0 + 1e-9))) / 4
    np.testing.assert_allclose(w, want, rtol=1e-15)
    # a point source (width << dx) is a 3x3 patch whose centre weight is 1/4 (quirk Q8)
    y0, y1, x0, x1, w = R.gaussian_patch_2d((5e-7, 3e-7, 0), 1e-9, dx, dy, 20, 20)
    assert w.shape == (3, 3) and w[1, 1] == 0.25 and w[0, 0] < 1e-100


def test_dft_closed_form_and_fft_bins():
    dt = 1e-16
    n = np.arange(512)
    t = n * dt
    w0 = 2 * np.pi * 3e14
    f = np.cos(w0 * t + 0.3)
    om = np.array([w0 * 0.9, w0, 2 * np.pi / (512 * dt) * 7])
    F = R.dft_of_series(f[:, None], t, om)[:, 0]
    for q, w in enumerate(om):
        # sum_n cos(a n + p) e^{-i b n} = 1/2 sum (e^{i(a-b)n+ip} + e^{-i(a+b)n-ip}) as geometric series
        a, b = w0 * dt, w * dt
        def geo(z):
            return (1 - z ** 512) / (1 - z) if abs(1 - z) > 1e-14 else 512.0
        want = 0.5 * (np.exp(0.3j) * geo(np.exp(1j * (a - b))) + np.exp(-0.3j) * geo(np.exp(-1j * (a + b))))
        assert abs(F[q] - want) < 1e-9 * 512
    fft = np.fft.fft(f)
    assert abs(F[2] - fft[7]) < 1e-9 * 512


def test_signals_and_stability_closed_forms():
    assert S.sigmoid(0.5, duration=1.0) == pytest.approx(0.5)
    t = np.array([1e-14])                     # envelope centre, and cos(2 pi f t) = 1 there
    g = S.gaussian_pulse(t, 3.0, 1e-14, 1e-15, 1e14, 0.0)
    assert g[0] == pytest.approx(9.0)          # amplitude enters twice (quirk Q11)
    ok, c, lim = S.check_fdtd_stability(1e-17, 1e-8, 1e-8, 1e-8)
    assert lim == pytest.approx(1 / np.sqrt(3)) and c == pytest.approx(R.LIGHT_SPEED * 1e-17 / 1e-8)
    res, dt = S.calc_optimal_fdtd_params(1.55e-6, 2.0, dims=2, safety_factor=1.0, points_per_wavelength=10)
    assert res == pytest.approx(1.55e-6 / 20) and dt == pytest.approx(res / (R.LIGHT_SPEED * np.sqrt(2)))


def test_monitor_points_brute_force():
    for (s, e) in [((1e-6, 1e-6), (1e-6, 3e-6)), ((0.5e-6, 0.7e-6), (2.9e-6, 1.3e-6)), ((1e-6, 2e-6), (3e-6, 2e-6))]:
        pts = R.monitor_points_2d(s, e, 1e-7, 1e-7)
        sx, sy, ex, ey = (int(round(v / 1e-7)) for v in (s[0], s[1], e[0], e[1]))
        n = max(abs(ex - sx), abs(ey - sy)) + 1
        assert len(pts) == n and pts[0] == (sx, sy) and pts[-1] == (ex, ey)
        for q, (x, y) in enumerate(pts):
            fx = sx + (ex - sx) * q / (n - 1) if n > 1 else sx
            fy = sy + (ey - sy) * q / (n - 1) if n > 1 else sy
            assert abs(x - fx) < 1 and abs(y - fy) < 1


def test_refsim_linear_in_source_amplitude():
    ny, nx = 30, 40
    eps = np.full((ny, nx), 2.0); sig = np.zeros((ny, nx)); sig[:, :4] = 1e5
    dx = dy = 5e-8
    dt = 0.9 * dx / (R.LIGHT_SPEED * np.sqrt(2))
    class Src:
        pass
    def run(a):
        s = Src(); s.position = (1e-6, 0.75e-6, 0); s.width = 1e-8
        s.signal = a * np.sin(np.arange(100) * 0.2)
        return R.RefSim(eps, sig, dx, dy, dt, 100, sources=[s]).run().F["Ez"]
    e1, e3 = run(1.0), run(3.0)
    np.testing.assert_allclose(e3, 3 * e1, rtol=1e-12, atol=1e-12 * np.abs(e3).max())
