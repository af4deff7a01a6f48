"""Pins the oracle's power, 3D plane slicing, 3D Gaussian patch and ModeSource
injection (oracle/fdtd_ref.py) to closed forms, brute force and hand-worked
cases, each chosen so that a dropped sqrt, a missing or extra conj, a wrong
cross-product sign, an off-by-one slice or a misplaced direction zeroing fails.

Reference lines restated by the oracle functions under test:
beamz/devices/monitors.py:222-273 (plane slices), :324-360 (slice_field),
:407-439 (monitor power); beamz/simulation/helper.py:8-54 (ModeSource),
:62-95 (3D Gaussian), :120-187 (grid power)."""
import cmath
import math
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import fdtd_ref as R


# ----------------------------------------------------------------------------- monitor power
def test_power_2d_hand_values_and_closed_form():
    # Pythagorean values: |S| = |Ez| * hypot(Hx, Hy); a dropped sqrt gives 64 and 9
    pm, tot = R.power_2d([2.0, -1.0], [0.0, 3.0], [4.0, 0.0])
    assert np.array_equal(pm, [8.0, 3.0]) and tot == 11.0
    rng = np.random.default_rng(0)
    ez, hx, hy = rng.standard_normal((3, 50))
    pm, tot = R.power_2d(ez, hx, hy)
    ref = [abs(a) * math.hypot(b, c) for a, b, c in zip(ez, hx, hy)]
    np.testing.assert_allclose(pm, ref, rtol=1e-14)
    assert tot == pytest.approx(sum(ref), rel=1e-13)


def test_power_2d_complex_has_no_conjugate():
    # monitors.py:411-416 multiplies without conj: S^2 = Ez^2 (Hx^2 + Hy^2), principal square root
    rng = np.random.default_rng(1)
    ez, hx, hy = (rng.standard_normal(20) + 1j * rng.standard_normal(20) for _ in range(3))
    pm, tot = R.power_2d(list(ez), list(hx), list(hy))
    ref = np.array([cmath.sqrt(a * a * (b * b + c * c)) for a, b, c in zip(ez, hx, hy)])
    np.testing.assert_allclose(pm, ref, rtol=1e-12)
    assert tot == pytest.approx(ref.sum(), rel=1e-12)
    # Ez = i, Hy = 1: Sx = -i, Sx^2 = -1, |S| = sqrt(-1) = i (with a conj it would be 1)
    pm, _ = R.power_2d([1j], [0.0], [1.0])
    assert pm[0] == 1j


def test_power_3d_lagrange_identity_and_plane_wave():
    rng = np.random.default_rng(2)
    E = rng.standard_normal((3, 6, 7))
    H = rng.standard_normal((3, 6, 7))
    pm, tot = R.power_3d(*E, *H)
    # |E x H|^2 = |E|^2 |H|^2 - (E.H)^2: independent of the component formulas
    ee, hh, eh = (E * E).sum(0), (H * H).sum(0), (E * H).sum(0)
    np.testing.assert_allclose(pm, np.sqrt(ee * hh - eh * eh), rtol=1e-10)
    np.testing.assert_allclose(pm, np.linalg.norm(np.cross(E, H, axis=0), axis=0), rtol=1e-13)
    assert tot == pytest.approx(pm.sum(), rel=1e-14)
    # x-travelling plane wave: E = z^, H = -y^/eta -> S = x^/eta
    eta = 376.7
    one, zero = np.ones((1, 1)), np.zeros((1, 1))
    pm, _ = R.power_3d(zero, zero, one, zero, -one / eta, zero)
    assert pm[0, 0] == pytest.approx(1 / eta, rel=1e-15)


# ----------------------------------------------------------------------------- grid power
def test_grid_power_2d_brute_force_real_and_complex():
    rng = np.random.default_rng(3)
    ny, nx = 5, 6
    for cplx in (False, True):
        def arr(s):
            a = rng.standard_normal(s)
            return a + 1j * rng.standard_normal(s) if cplx else a
        Ez, Hx, Hy = arr((ny, nx)), arr((ny, nx - 1)), arr((ny - 1, nx))
        got = R.grid_power_2d(Ez, Hx, Hy)
        assert got.shape == (ny, nx)
        for j in range(ny):
            for i in range(nx):
                hx = Hx[j, i] if i < nx - 1 else 0.0
                hy = Hy[j, i] if j < ny - 1 else 0.0
                e = complex(Ez[j, i])
                sx = -(e * np.conj(hy)).real            # Re(Ez conj(Hy)) = ReRe + ImIm
                sy = (e * np.conj(hx)).real
                assert got[j, i] == pytest.approx(sx * sx + sy * sy, rel=1e-12, abs=1e-300)
    # no square root (helper.py:179): Ez = 2, Hx = 3 at one cell gives 36
    Ez = np.zeros((2, 2)); Hx = np.zeros((2, 1)); Hy = np.zeros((1, 2))
    Ez[0, 0], Hx[0, 0] = 2.0, 3.0
    assert R.grid_power_2d(Ez, Hx, Hy)[0, 0] == 36.0


def test_grid_power_3d_brute_force_with_conjugate():
    rng = np.random.default_rng(4)
    nx, ny, nz = 5, 4, 3
    shp = R.field_shapes(nx, ny, nz)
    F = {k: rng.standard_normal(s) + 1j * rng.standard_normal(s) for k, s in shp.items()}
    got = R.grid_power_3d(F)
    assert got.shape == (nz - 1, ny - 1, nx - 1)       # the common corner-aligned box
    for k in range(nz - 1):
        for j in range(ny - 1):
            for i in range(nx - 1):
                E = np.array([F[c][k, j, i] for c in ("Ex", "Ey", "Ez")])
                H = np.array([F[c][k, j, i] for c in ("Hx", "Hy", "Hz")])
                S = np.cross(E, np.conj(H)).real
                assert got[k, j, i] == pytest.approx(float(np.linalg.norm(S)), rel=1e-12)
    # a conj matters: E = (0, 0, 1), H = (0, i, 0): Re(E x conj H) = (-Re(conj(i)), 0, 0) -> |S| = 0
    G = {k: np.zeros(s, complex) for k, s in shp.items()}
    G["Ez"][0, 0, 0], G["Hy"][0, 0, 0] = 1.0, 1j
    assert R.grid_power_3d(G)[0, 0, 0] == 0.0
    G["Hy"][0, 0, 0] = -2.0
    assert R.grid_power_3d(G)[0, 0, 0] == 2.0


# ----------------------------------------------------------------------------- 3D plane slices
def _mon(normal, pos, start, size):
    return SimpleNamespace(plane_normal=normal, plane_position=pos, start=start, size=size)


def test_monitor_slice_3d_hand_worked():
    d = 0.1e-6
    W, H, D = 1.0e-6, 0.8e-6, 0.6e-6                 # base grid 10 x 8 x 6 (x, y, z)
    # z-normal, in range: z = round(2.2) = 2, x 2..7, y 3..7
    sl = R.monitor_slice_3d(_mon("z", 0.22e-6, (0.2e-6, 0.3e-6, 0.0), (0.5e-6, 0.4e-6)), W, H, D, d, d, d)
    assert sl == (slice(3, 7), slice(2, 7), 2)
    # z beyond the top clamps to 5; x start below 0 clamps to 0; ends beyond clamp to 10 and 8
    sl = R.monitor_slice_3d(_mon("z", 9e-6, (-0.3e-6, 0.0, 0.0), (5e-6, 5e-6)), W, H, D, d, d, d)
    assert sl == (slice(0, 8), slice(0, 10), 5)
    # y-normal: (y index, x slice, z slice), size = (x extent, z extent)
    sl = R.monitor_slice_3d(_mon("y", 0.41e-6, (0.1e-6, 0.0, 0.2e-6), (0.3e-6, 0.2e-6)), W, H, D, d, d, d)
    assert sl == (4, slice(1, 4), slice(2, 4))
    # x-normal: (y slice, x index, z slice), size = (y extent, z extent)
    sl = R.monitor_slice_3d(_mon("x", 0.79e-6, (0.0, 0.2e-6, 0.1e-6), (0.5e-6, 0.3e-6)), W, H, D, d, d, d)
    assert sl == (slice(2, 7), 8, slice(1, 4))
    # the start clamps to n - 1, so an all-outside range is empty, not negative
    sl = R.monitor_slice_3d(_mon("z", 0.0, (2e-6, 2e-6, 0.0), (1e-6, 1e-6)), W, H, D, d, d, d)
    assert sl == (slice(7, 8), slice(9, 10), 0)


def test_slice_field_3d_brute_force_codes():
    nz, ny, nx = 5, 6, 7
    code = lambda s: np.fromfunction(lambda k, j, i: 10000 * k + 100 * j + i, s)
    Ez = code((nz - 1, ny, nx))                        # four z-planes
    Hz = code((nz, ny - 1, nx - 1))
    # z-normal plane at 4: Ez has planes 0..3, so the index clamps to 3; Hz keeps 4 and clips y, x to its size
    sl = (slice(1, 6), slice(2, 7), 4)
    got = R.slice_field_3d(Ez, sl, "z")
    assert got.shape == (5, 5)
    assert np.array_equal(got, [[30000 + 100 * j + i for i in range(2, 7)] for j in range(1, 6)])
    got = R.slice_field_3d(Hz, sl, "z")
    assert np.array_equal(got, [[40000 + 100 * j + i for i in range(2, 6)] for j in range(1, 5)])
    # a copy, not a view
    got[0, 0] = -1
    assert Hz[4, 1, 2] == 40102
    # x- and y-normal planes return the whole volume (the reference's fallback, monitors.py:356-358)
    full = R.slice_field_3d(Ez, (slice(1, 3), 2, slice(0, 2)), "x")
    assert np.array_equal(full, Ez) and full is not Ez


# ----------------------------------------------------------------------------- 3D Gaussian patch
def test_gaussian_patch_3d_closed_form_and_clipping():
    d = 50e-9
    nx, ny, nz = 30, 20, 12
    ez_nz = nz - 1
    # width = d: 3 cells each side, a 7^3 box centred on the source's cell, unit weight at the centre (no /4)
    z0, z1, y0, y1, x0, x1, w = R.gaussian_patch_3d((10 * d, 9 * d, 5 * d), d, d, d, d, nx, ny, nz, ez_nz)
    assert (z0, z1, y0, y1, x0, x1) == (2, 9, 6, 13, 7, 14)
    assert w.shape == (7, 7, 7) and w[3, 3, 3] == 1.0
    # separable Gaussian with variance (w/d)^2 + 1e-9 on each axis, from math.exp per axis
    g = lambda c, n: np.array([math.exp(-(q - c) ** 2 / (2 * (1.0 + 1e-9))) for q in n])
    ref = np.einsum("k,j,i->kji", g(5, range(2, 9)), g(9, range(6, 13)), g(10, range(7, 14)))
    np.testing.assert_allclose(w, ref, rtol=1e-13)
    # off-cell centre and anisotropic cells
    dz = 2 * d
    z0, z1, y0, y1, x0, x1, w = R.gaussian_patch_3d((10.4 * d, 9.6 * d, 5.2 * dz), 0.6 * d, d, d, dz, nx, ny, nz,
                                                    ez_nz)
    wx, wz = 0.6, 0.3                                   # width in cells along x (and y) and along z
    assert (x0, x1) == (10 - 2, 10 + 3) and (y0, y1) == (10 - 2, 10 + 3) and (z0, z1) == (5 - 1, 5 + 2)
    gx = lambda c, n, s: np.array([math.exp(-(q - c) ** 2 / (2 * (s * s + 1e-9))) for q in n])
    ref = np.einsum("k,j,i->kji", gx(5.2, range(z0, z1), wz), gx(9.6, range(y0, y1), wx),
                    gx(10.4, range(x0, x1), wx))
    np.testing.assert_allclose(w, ref, rtol=1e-12)
    # clipping: at the low corner the box starts at 0; at the top z the Ez planes end at ez_nz
    z0, z1, y0, y1, x0, x1, w = R.gaussian_patch_3d((d, 0.0, (nz - 1) * d), d, d, d, d, nx, ny, nz, ez_nz)
    assert (x0, x1, y0, y1) == (0, 5, 0, 4)
    assert (z0, z1) == (nz - 1 - 3, ez_nz) and w.shape == (ez_nz - (nz - 4), 4, 5)
    # the last plane kept is z = 10, one cell below the centre plane z = 11 (x = 1, y = 0 on the centre)
    np.testing.assert_allclose(w[-1, 0, 1], math.exp(-1.0 / (2 * (1 + 1e-9))), rtol=1e-14)


# ----------------------------------------------------------------------------- ModeSource points + direction
def _mode(points, direction="+x", enforce=True, signal=(1.0,)):
    return SimpleNamespace(mode_profiles=[points], direction=direction, enforce_direction=enforce,
                           signal=np.asarray(signal))


def test_mode_points_rounding_bounds_and_3d_plane_rule():
    d = 0.1e-6
    pts = [(1.0, 2.4 * d, 3.6 * d), (2.0, 2.5 * d, 1.0 * d), (3.0, -0.6 * d, 1.0 * d), (4.0, 9.6 * d, 1.0 * d),
           (5.0, 9.4 * d, 7.4 * d)]
    got = R.mode_points(_mode(pts), d, d, None, 10, 8, None, None, False)
    # round half to even (2.5 -> 2); x = -1 and x = 10 are dropped; z is None in 2D
    assert got == [(None, 4, 2, 1.0), (None, 1, 2, 2.0), (None, 7, 9, 5.0)]
    # 3D: Ez has nz - 1 = 5 planes; z = 5 lies in the grid but not in Ez: the reference puts it on 5 // 2
    pts3 = [(1.0, d, d, 0.0), (2.0, d, d, 4.0 * d), (3.0, d, d, 5.0 * d), (4.0, d, d, 6.0 * d), (5.0, d, d, -d)]
    got = R.mode_points(_mode(pts3), d, d, d, 10, 8, 6, 5, True)
    assert got == [(0, 1, 1, 1.0), (4, 1, 1, 2.0), (2, 1, 1, 3.0)]


@pytest.mark.parametrize("direction, zeroed", [("+x", (3, 2)), ("-x", (3, 4)), ("+y", (2, 3)), ("-y", (4, 3))])
def test_mode_injection_direction_zeroing_2d(direction, zeroed):
    d = 1.0
    Ez0 = np.arange(1.0, 1.0 + 7 * 6).reshape(6, 7)
    src = _mode([(0.5, 3.0 * d, 3.0 * d)], direction, signal=[2.0])
    sim = R.RefSim(np.ones((6, 7)), np.zeros((6, 7)), d, d, 1.0, 1, sources=[src])
    sim.F["Ez"] = Ez0.copy()
    sim.apply_sources()
    want = Ez0.copy()
    want[3, 3] += 1.0
    want[zeroed] = 0.0
    assert np.array_equal(sim.F["Ez"], want)
    # enforce_direction=False: only the addition
    src.enforce_direction = False
    sim = R.RefSim(np.ones((6, 7)), np.zeros((6, 7)), d, d, 1.0, 1, sources=[src])
    sim.F["Ez"] = Ez0.copy()
    sim.apply_sources()
    want = Ez0.copy(); want[3, 3] += 1.0
    assert np.array_equal(sim.F["Ez"], want)


def test_mode_injection_order_edges_and_complex_amplitude():
    d = 1.0
    # points processed in order: the second point's +x zeroing wipes the first point's cell;
    # at x = 0 there is no upstream cell to zero
    src = _mode([(1.0, 2.0, 1.0), (1.0, 3.0, 1.0), (1.0, 0.0, 2.0)], "+x", signal=[1.0])
    sim = R.RefSim(np.ones((4, 5)), np.zeros((4, 5)), d, d, 1.0, 1, sources=[src])
    sim.F["Ez"][:] = 5.0
    sim.apply_sources()
    want = np.full((4, 5), 5.0)
    want[1, 1] = 0.0; want[1, 2] = 0.0; want[1, 3] = 6.0; want[2, 0] = 6.0
    assert np.array_equal(sim.F["Ez"], want)
    # complex amplitude x signal on a complex field is added unreduced; on a real field the real part
    src = _mode([(1 + 2j, 2.0, 1.0)], "-y", enforce=False, signal=[0.5 - 1j])
    sim = R.RefSim(np.ones((4, 5)), np.zeros((4, 5)), d, d, 1.0, 1, sources=[src], dtype=np.complex128)
    sim.apply_sources()
    assert sim.F["Ez"][1, 2] == (1 + 2j) * (0.5 - 1j) and sim.F["Ez"].sum() == (1 + 2j) * (0.5 - 1j)
    sim = R.RefSim(np.ones((4, 5)), np.zeros((4, 5)), d, d, 1.0, 1, sources=[src])
    sim.apply_sources()
    assert sim.F["Ez"][1, 2] == ((1 + 2j) * (0.5 - 1j)).real


@pytest.mark.parametrize("direction, zeroed", [("+z", (1, 2, 3)), ("-z", (3, 2, 3)), ("+x", (2, 2, 2)),
                                               ("-y", (2, 3, 3))])
def test_mode_injection_direction_zeroing_3d(direction, zeroed):
    d = 1.0
    nz, ny, nx = 6, 5, 6
    src = _mode([(0.25, 3.0, 2.0, 2.0)], direction, signal=[4.0])
    sim = R.RefSim(np.ones((nz, ny, nx)), np.zeros((nz, ny, nx)), d, d, 1.0, 1, dz=d, sources=[src])
    Ez0 = np.arange(1.0, 1.0 + sim.F["Ez"].size).reshape(sim.F["Ez"].shape)
    sim.F["Ez"] = Ez0.copy()
    sim.apply_sources()
    want = Ez0.copy()
    want[2, 2, 3] += 1.0
    want[zeroed] = 0.0
    assert np.array_equal(sim.F["Ez"], want)


# ----------------------------------------------------------------------------- x/y-normal planes and flux
def test_true_plane_3d_brute_force_codes_and_clamping():
    nz, ny, nx = 6, 7, 8
    code = lambda s: np.fromfunction(lambda k, j, i: 10000 * k + 100 * j + i, s)
    Hy = code((nz - 1, ny, nx - 1))          # 5 planes, 7 rows, 7 columns
    Ey = code((nz, ny - 1, nx))
    # x-normal: (y slice, x index, z slice) -> [z][y]; z stops at Hy's 5 planes
    got = R.true_plane_3d(Hy, (slice(1, 5), 3, slice(2, 9)), "x")
    assert np.array_equal(got, [[10000 * k + 100 * j + 3 for j in range(1, 5)] for k in range(2, 5)])
    # the fixed x index clamps to the component's last column (nx - 2 for Hy)
    got = R.true_plane_3d(Hy, (slice(0, 2), nx - 1, slice(0, 1)), "x")
    assert np.array_equal(got, [[100 * j + nx - 2 for j in range(0, 2)]])
    # y-normal: (y index, x slice, z slice) -> [z][x]; y clamps to Ey's last row (ny - 2)
    got = R.true_plane_3d(Ey, (ny - 1, slice(2, 5), slice(1, 3)), "y")
    assert np.array_equal(got, [[10000 * k + 100 * (ny - 2) + i for i in range(2, 5)] for k in range(1, 3)])
    got[0, 0] = -1
    assert Ey[1, ny - 2, 2] != -1


def test_plane_flux_plane_wave_sign_and_conjugate():
    eta = 376.7
    phi = np.linspace(0.3, 2.0, 5)[:, None] * np.ones((1, 12))        # 5 frequencies, 12 plane points
    one = np.exp(1j * phi)
    zero = np.zeros_like(one)
    # +x wave: Ey = e^{i phi}, Hz = e^{i phi} / eta -> Re(Ey conj Hz) = 1/eta per point (no conj: cos(2 phi))
    D = {"Ex": zero, "Ey": one, "Ez": zero, "Hx": zero, "Hy": zero, "Hz": one / eta}
    np.testing.assert_allclose(R.plane_flux(D, "x"), 12 / eta, rtol=1e-14)
    # the other polarisation: Ez = 1, Hy = -1/eta travels along +x too
    D = {"Ex": zero, "Ey": zero, "Ez": one, "Hx": zero, "Hy": -one / eta, "Hz": zero}
    np.testing.assert_allclose(R.plane_flux(D, "x"), 12 / eta, rtol=1e-14)
    # through a y-normal plane it carries nothing; a z-normal plane sees Ex, Hy
    assert np.allclose(R.plane_flux(D, "y"), 0.0)
    D = {"Ex": one, "Ey": zero, "Ez": zero, "Hx": zero, "Hy": one / eta, "Hz": zero}
    np.testing.assert_allclose(R.plane_flux(D, "z"), 12 / eta, rtol=1e-14)
