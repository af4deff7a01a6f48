"""GPU parity of the (Hz, Ex, Ey) polarisation (bz_yee2d_hz_run) against the FP64 oracle
(oracle/fdtd_ref.py RefSim(polarization="Hz"), pinned in tests/test_oracle_hz.py).

Tolerances (relative L2): fields, monitor series and DFT 1e-4; monitor and grid power 2e-4;
a batch of uneven members, graph replays and tile heights: bit-identical to single eager runs."""
import numpy as np
import pytest

import beamz_amd as bz
from beamz_amd.devices.signals import gaussian_pulse
from beamz_amd.simulation.batch import FDTDBatch
from oracle import fdtd_ref as R

pytestmark = pytest.mark.gpu

um = 1e-6


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128); b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def design(nx, ny, npml=10, lam=1.55 * um):
    dx = lam / 20
    W, H = nx * dx, ny * dx
    d = bz.Design(W, H, material=bz.Material(1.444 ** 2), auto_pml=False)
    s = 4 * np.log(1e6) / (2 * npml * 1.2)
    for pos, size, o in (((0, 0), (npml * dx, H), "left"), ((W - npml * dx, 0), (npml * dx, H), "right"),
                         ((0, H - npml * dx), (W, npml * dx), "top"), ((0, 0), (W, npml * dx), "bottom")):
        d.boundaries.append(bz.PML("rect", pos, size, o, polynomial_order=3.0, sigma_factor=s))
    d += bz.Rectangle(position=(0, H / 2 - 0.25 * um), width=W, height=0.5 * um, material=bz.Material(3.476 ** 2))
    d += bz.Rectangle(position=(W / 2, H / 2 - 0.8 * um), width=0.6 * um, height=0.3 * um,
                      material=bz.Material(3.476 ** 2))
    return d, dx


def pulse(t, lam=1.55 * um, dlam=0.2 * um):
    f0 = bz.LIGHT_SPEED / lam
    df = bz.LIGHT_SPEED * dlam / lam ** 2
    st = 1 / (2 * np.pi * df)
    return gaussian_pulse(t, 1.0, 4 * st, st, f0, 0.0)


def run_case(nx=150, ny=84, nsteps=900, power=True, **grid_opts):
    d, dx = design(nx, ny)
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(nsteps) * dt
    d += bz.GaussianSource(position=(25 * dx, d.height / 2), width=dx / 4, signal=pulse(t))
    freqs = bz.LIGHT_SPEED / np.linspace(1.45e-6, 1.65e-6, 6)
    mon = bz.Monitor(design=d, start=((nx - 30) * dx, 12 * dx), end=((nx - 30) * dx, d.height - 12 * dx),
                     dft_frequencies=freqs)
    d += mon
    sim = bz.FDTD(d, time=t, resolution=dx, polarization="Hz")
    for k, v in grid_opts.items():
        setattr(sim.grid, k, v)
    sim.run(live=False, save=False, accumulate_power=power)
    return sim, d, mon, dx, dt, freqs


def test_hz_polarisation_matches_fp64_oracle(hip):
    sim, d, mon, dx, dt, freqs = run_case()
    assert sim.grid.kernel_name().startswith("bz_yee2d_hz_run")
    ref = R.RefSim(sim.epsilon_r, sim.sigma, dx, dx, dt, sim.num_steps, sources=d.sources, monitors=[mon],
                   accumulate_power=True, polarization="Hz").run()
    for k in ("Hz", "Ex", "Ey"):
        got = getattr(sim, k)
        assert got.shape == ref.F[k].shape, k
        err = rel_l2(got, ref.F[k])
        print(f"{k}: rel-L2 {err:.2e}")
        assert err < 1e-4, k
    for k in ("Hz", "Ex", "Ey"):
        ser, ref_ser = np.array(mon.fields[k]), np.array(ref.mon_series[0][k])
        assert ser.shape == ref_ser.shape
        assert rel_l2(ser, ref_ser) < 1e-4, k
        F = R.dft_of_series(ref_ser, ref.mon_series[0]["t"], 2 * np.pi * freqs)
        assert rel_l2(mon.dft[k], F) < 1e-4, k
    assert rel_l2(mon.power_history, ref.mon_series[0]["power"]) < 2e-4
    assert rel_l2(sim.power_accumulated, ref.power_accumulated / ref.power_count) < 2e-4
    # the flux through the line is the sum of Re(Ey conj(Hz)) of the oracle's DFTs
    Fh = R.dft_of_series(np.array(ref.mon_series[0]["Hz"]), ref.mon_series[0]["t"], 2 * np.pi * freqs)
    Fey = R.dft_of_series(np.array(ref.mon_series[0]["Ey"]), ref.mon_series[0]["t"], 2 * np.pi * freqs)
    assert rel_l2(mon.dft_flux(), np.real(Fey * np.conj(Fh)).sum(axis=1)) < 2e-4


def test_hz_tile_heights_graphs_and_batch_are_bit_identical(hip):
    base, *_ = run_case(power=False, seg2d=2)
    for opts in ({"seg2d": 8}, {"seg2d": 16}, {"seg2d": 4, "graph2d": True}):
        other, *_ = run_case(power=False, **opts)
        for k in ("Hz", "Ex", "Ey"):
            assert np.array_equal(getattr(other, k), getattr(base, k)), (opts, k)
    # a batch of uneven members, each bit-identical to its single run
    sims, singles = [], []
    for (nx, ny) in ((150, 84), (97, 131), (203, 66)):
        single, *_ = run_case(nx, ny, power=False)
        singles.append(single)
        d, dx = design(nx, ny)
        dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
        t = np.arange(900) * dt
        d += bz.GaussianSource(position=(25 * dx, d.height / 2), width=dx / 4, signal=pulse(t))
        freqs = bz.LIGHT_SPEED / np.linspace(1.45e-6, 1.65e-6, 6)
        d += bz.Monitor(design=d, start=((nx - 30) * dx, 12 * dx), end=((nx - 30) * dx, d.height - 12 * dx),
                        dft_frequencies=freqs)
        sims.append(bz.FDTD(d, time=t, resolution=dx, polarization="Hz"))
    b = FDTDBatch(sims)
    b.run()
    for s, s1 in zip(sims, singles):
        for k in ("Hz", "Ex", "Ey"):
            assert np.array_equal(getattr(s, k), getattr(s1, k)), k
        assert np.array_equal(s.monitors[0].dft["Hz"], s1.monitors[0].dft["Hz"])


def test_hz_mode_source_matches_oracle(hip):
    lam = 1.55 * um
    d, dx = design(140, 70)
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(300) * dt
    sig = bz.ramped_cosine(t, 1.0, bz.LIGHT_SPEED / lam, 0.0, ramp_duration=20 * dt, t_max=t[-1])
    src = bz.ModeSource(design=d, start=(20 * dx, d.height / 2 - 1.0 * um), end=(20 * dx, d.height / 2 + 1.0 * um),
                        wavelength=lam, signal=sig, num_modes=1, grid_resolution=400)
    d += src
    for enforce in (True, False):
        src.enforce_direction = enforce
        sim = bz.FDTD(d, time=t, resolution=dx, polarization="Hz")
        sim.run(live=False, save=False)
        ref = R.RefSim(sim.epsilon_r, sim.sigma, dx, dx, dt, len(t), sources=[src], polarization="Hz").run()
        for k in ("Hz", "Ex", "Ey"):
            assert rel_l2(np.real(getattr(sim, k)), ref.F[k]) < 1e-4, (enforce, k)
