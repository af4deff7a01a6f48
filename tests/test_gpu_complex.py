"""Complex fields on the GPU against the FP64 oracle run in complex128, as the reference's NumPy path runs
(beamz/simulation/fdtd.py:151-186 allocates complex128 fields; helper.py:29-35 adds amplitude * signal
unreduced). A ModeSource profile with complex amplitudes makes ``FDTD`` run a second, imaginary grid; the
field attributes, monitor series, DFT spectra, monitor power (complex, no conjugate, monitors.py:407-417) and
grid power (helper.py:120-187, from the real and imaginary parts) combine both.

- 2D: a solved waveguide mode (its PML-stretched eigenvector is complex) turned by a phase of 1 rad, so the
  real and imaginary parts are of one size, with and without direction enforcement;
- 3D: a ModeSource plane (sources.py:323-376, helper.py:19-47) in an SOI strip guide, real amplitudes (one
  grid) and the same plane turned by a phase (two grids).

Tolerances (relative L2, complex): fields, series, spectra 1e-4 (also of the imaginary part alone); power
2e-4."""
import numpy as np
import pytest

import beamz_amd as bz
from oracle import fdtd_ref as R

pytestmark = pytest.mark.gpu

um = 1e-6
LAM = 1.55 * um


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128); b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def turn(src, phase):
    for p in src.mode_profiles[0]:
        p[0] = complex(p[0]) * np.exp(1j * phase)


def design_2d(nx=140, ny=70, npml=10):
    dx = LAM / 20
    W, H = nx * dx, ny * dx
    d = bz.Design(W, H, material=bz.Material(1.444 ** 2), auto_pml=False)
    s = 4 * np.log(1e6) / (2 * npml * 1.2)
    for pos, size, o in (((0, 0), (npml * dx, H), "left"), ((W - npml * dx, 0), (npml * dx, H), "right"),
                         ((0, H - npml * dx), (W, npml * dx), "top"), ((0, 0), (W, npml * dx), "bottom")):
        d.boundaries.append(bz.PML("rect", pos, size, o, polynomial_order=3.0, sigma_factor=s))
    d += bz.Rectangle(position=(0, H / 2 - 0.25 * um), width=W, height=0.5 * um, material=bz.Material(3.476 ** 2))
    return d, dx


def compare(sim, ref, mon, names, tag, power=True):
    errs = {}
    for k in names:
        got, want = getattr(sim, k), ref.F[k]
        errs[k] = rel_l2(got, want)
        errs[k + " imag"] = rel_l2(np.imag(got), np.imag(want))
        assert np.linalg.norm(np.imag(want)) > 1e-3 * np.linalg.norm(want), (tag, k)
    ts = ref.mon_series[0]["t"]
    for k in names:
        ser = np.array(ref.mon_series[0][k])
        errs["series " + k] = rel_l2(np.array(mon.fields[k]), ser)
        errs["dft " + k] = rel_l2(mon.dft[k], R.dft_of_series(ser, ts, 2 * np.pi * mon.dft_frequencies))
    errs["monitor power"] = rel_l2(mon.power_history, ref.mon_series[0]["power"])
    if power:
        assert sim.power_accumulation_count == ref.power_count
        errs["grid power"] = rel_l2(sim.power_accumulated, ref.power_accumulated / ref.power_count)
    print(tag, {k: f"{v:.2e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v < (2e-4 if "power" in k else 1e-4), (tag, k, v)


@pytest.mark.parametrize("enforce", [True, False])
def test_complex_mode_source_2d_matches_complex_oracle(hip, enforce):
    d, dx = design_2d()
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(300) * dt
    sig = bz.ramped_cosine(t, 1.0, bz.LIGHT_SPEED / LAM, 0.0, ramp_duration=20 * dt, t_max=t[-1])
    src = bz.ModeSource(design=d, start=(20 * dx, d.height / 2 - 1.0 * um), end=(20 * dx, d.height / 2 + 1.0 * um),
                        wavelength=LAM, signal=sig, num_modes=1, grid_resolution=400)
    turn(src, 1.0)
    src.enforce_direction = enforce
    d += src
    mon = bz.Monitor(design=d, start=(100 * dx, 10 * dx), end=(100 * dx, d.height - 10 * dx),
                     dft_frequencies=bz.LIGHT_SPEED / (np.array([1.5, 1.55, 1.6]) * um), accumulate_power=True)
    d += mon
    sim = bz.FDTD(d, time=t, resolution=dx)
    sim.run(live=False, save=False, accumulate_power=True)
    assert sim.is_complex_backend and sim.grid_im is not None
    ref = R.RefSim(sim.epsilon_r, sim.sigma, dx, dx, dt, len(t), sources=[src], monitors=[mon],
                   dtype=np.complex128, accumulate_power=True).run()
    compare(sim, ref, mon, ("Ez", "Hx", "Hy"), f"2D complex ModeSource (enforce {enforce})")


def test_complex_fields_off_keeps_the_real_part(hip):
    """complex_fields=False: the reference's real-backend branch (helper.py:29-33), Re(amplitude * signal),
    against the oracle run with real fields; and the real part of the complex run is the same field."""
    d, dx = design_2d()
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(200) * dt
    sig = bz.ramped_cosine(t, 1.0, bz.LIGHT_SPEED / LAM, 0.0, ramp_duration=20 * dt, t_max=t[-1])
    src = bz.ModeSource(design=d, start=(20 * dx, d.height / 2 - 1.0 * um), end=(20 * dx, d.height / 2 + 1.0 * um),
                        wavelength=LAM, signal=sig, num_modes=1, grid_resolution=400)
    turn(src, 1.0)
    d += src
    real = bz.FDTD(d, time=t, resolution=dx, complex_fields=False)
    real.run(live=False, save=False)
    assert not real.is_complex_backend
    full = bz.FDTD(d, time=t, resolution=dx)
    full.run(live=False, save=False)
    ref = R.RefSim(real.epsilon_r, real.sigma, dx, dx, dt, len(t), sources=[src]).run()
    for k in ("Ez", "Hx", "Hy"):
        assert np.all(np.imag(getattr(real, k)) == 0)
        assert rel_l2(getattr(real, k), ref.F[k]) < 1e-4, k
        assert np.array_equal(np.real(getattr(real, k)), np.real(getattr(full, k))), k


def _soi_3d(nsteps, nx=40, ny=32, nz=24):
    dx = 50e-9
    d = bz.Design(nx * dx, ny * dx, depth=nz * dx, material=bz.Material(1.444 ** 2), pml_size=6 * dx)
    d += bz.Rectangle(position=(0, ny * dx / 2 - 0.25 * um, nz * dx / 2 - 0.11 * um), width=nx * dx, height=0.5 * um,
                      depth=0.22 * um, material=bz.Material(3.476 ** 2))
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(3))
    t = np.arange(nsteps) * dt
    sig = bz.ramped_cosine(t, 1.0, bz.LIGHT_SPEED / LAM, 0.0, ramp_duration=20 * dt, t_max=t[-1])
    src = bz.ModeSource(design=d, position=(8 * dx, ny * dx / 2, nz * dx / 2), width=1.0 * um, height=0.6 * um,
                        wavelength=LAM, signal=sig, num_modes=1, grid_resolution=200, direction="+x")
    mon = bz.Monitor(design=d, plane_normal="z", plane_position=nz * dx / 2, start=(0, 0, nz * dx / 2),
                     size=(nx * dx, ny * dx), dft_frequencies=bz.LIGHT_SPEED / (np.array([1.5, 1.6]) * um),
                     accumulate_power=True)
    return d, dx, dt, t, src, mon


@pytest.mark.parametrize("phase", [None, 0.7])
def test_mode_source_plane_3d_matches_oracle(hip, phase):
    d, dx, dt, t, src, mon = _soi_3d(150)
    if phase is not None:
        turn(src, phase)
    d += src
    d += mon
    sim = bz.FDTD(d, time=t, resolution=dx)
    sim.run(live=False, save=False, accumulate_power=True)
    assert sim.is_complex_backend == (phase is not None)
    ref = R.RefSim(sim.epsilon_r, sim.sigma, dx, dx, dt, len(t), dz=dx, sources=[src], monitors=[mon],
                   dtype=np.complex128 if phase is not None else np.float64, accumulate_power=True,
                   design_dims=(d.width, d.height, d.depth)).run()
    names = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    if phase is None:
        errs = {k: rel_l2(getattr(sim, k), ref.F[k]) for k in names}
        ts = ref.mon_series[0]["t"]
        for k in names:
            ser = np.array(ref.mon_series[0][k])
            errs["series " + k] = rel_l2(np.array(mon.fields[k]), ser)
            errs["dft " + k] = rel_l2(mon.dft[k], R.dft_of_series(ser, ts, 2 * np.pi * mon.dft_frequencies))
        errs["grid power"] = rel_l2(sim.power_accumulated, ref.power_accumulated / ref.power_count)
        print("3D ModeSource plane, real amplitudes:", {k: f"{v:.2e}" for k, v in errs.items()})
        assert np.abs(ref.F["Ez"]).max() > 0
        for k, v in errs.items():
            assert v < (2e-4 if "power" in k else 1e-4), (k, v)
    else:
        compare(sim, ref, mon, names, f"3D complex ModeSource plane (phase {phase})")
