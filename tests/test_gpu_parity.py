"""GPU parity: the HIP engine (FP32) against the FP64 oracle on identical grids,
sources and monitors. Tolerances are relative L2 errors stated per test."""
import numpy as np
import pytest

import beamz_amd as bz
from beamz_amd.devices.signals import gaussian_pulse
from oracle import fdtd_ref as R

pytestmark = pytest.mark.gpu

um = 1e-6


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128); b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def waveguide_design_2d(nx=160, ny=80, npml=12, lam=1.55 * um):
    dx = lam / 20
    W, H = nx * dx, ny * dx
    d = bz.Design(W, H, material=bz.Material(1.444 ** 2), auto_pml=False)
    s = 4 * np.log(1e6) / (2 * npml * 1.2)
    for pos, size, o in (((0, 0), (npml * dx, H), "left"), ((W - npml * dx, 0), (npml * dx, H), "right"),
                         ((0, H - npml * dx), (W, npml * dx), "top"), ((0, 0), (W, npml * dx), "bottom")):
        d.boundaries.append(bz.PML("rect", pos, size, o, polynomial_order=3.0, sigma_factor=s))
    d += bz.Rectangle(position=(0, H / 2 - 0.25 * um), width=W, height=0.5 * um, material=bz.Material(3.476 ** 2))
    return d, dx


def pulse(t, lam=1.55 * um, dlam=0.2 * um):
    f0 = bz.LIGHT_SPEED / lam
    df = bz.LIGHT_SPEED * dlam / lam ** 2
    sig_t = 1 / (2 * np.pi * df)
    return gaussian_pulse(t, 1.0, 4 * sig_t, sig_t, f0, 0.0)


def test_backend_plugin_updates_match_oracle(hip):
    rng = np.random.default_rng(5)
    ny, nx = 37, 53
    dx, dy, dt = 5e-8, 5e-8, 1e-16
    be = bz.get_backend("hip")
    Ez = rng.standard_normal((ny, nx)); Hx = rng.standard_normal((ny, nx - 1)) / 377; Hy = rng.standard_normal((ny - 1, nx)) / 377
    sig = rng.random((ny, nx)) * 5e4; eps = 1 + 10 * rng.random((ny, nx))
    dEz, dHx, dHy, dsig, deps = (be.from_numpy(a) for a in (Ez, Hx, Hy, sig, eps))
    for _ in range(20):
        dHx, dHy = be.update_h_fields(dHx, dHy, dEz, dsig, dx, dy, dt, bz.MU_0, bz.EPS_0)
        dEz = be.update_e_field(dEz, dHx, dHy, dsig, deps, dx, dy, dt, bz.EPS_0)
        Hx, Hy = R.update_h_fields_2d(Hx, Hy, Ez, sig, dx, dy, dt)
        Ez = R.update_e_field_2d(Ez, Hx, Hy, sig, eps, dx, dy, dt)
    assert rel_l2(be.to_numpy(dEz), Ez) < 2e-6
    assert rel_l2(be.to_numpy(dHx), Hx) < 2e-6
    assert rel_l2(be.to_numpy(dHy), Hy) < 2e-6


def test_fdtd_2d_waveguide_matches_oracle(hip):
    d, dx = waveguide_design_2d()
    nsteps = 1200
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(nsteps) * dt
    sig = pulse(t)
    d += bz.GaussianSource(position=(30 * dx, d.height / 2), width=dx / 4, signal=sig)
    freqs = bz.LIGHT_SPEED / np.linspace(1.45e-6, 1.65e-6, 8)
    mon = bz.Monitor(design=d, start=(120 * dx, 14 * dx), end=(120 * dx, d.height - 14 * dx), dft_frequencies=freqs)
    d += mon
    sim = bz.FDTD(d, time=t, resolution=dx)
    sim.run(live=False, save=False, accumulate_power=True)
    ref = R.RefSim(sim.epsilon_r, sim.sigma, dx, dx, dt, nsteps, sources=d.sources, monitors=[mon],
                   accumulate_power=True).run()
    for k in ("Ez", "Hx", "Hy"):
        assert rel_l2(np.real(getattr(sim, k)), ref.F[k]) < 1e-4, k
    ser = np.array(mon.fields["Ez"]); ref_ser = np.array(ref.mon_series[0]["Ez"])
    assert ser.shape == ref_ser.shape
    assert rel_l2(ser, ref_ser) < 1e-4
    assert rel_l2(mon.power_history, ref.mon_series[0]["power"]) < 2e-4
    F = R.dft_of_series(ref_ser, ref.mon_series[0]["t"], 2 * np.pi * freqs)
    assert rel_l2(mon.dft["Ez"], F) < 1e-4
    assert rel_l2(sim.power_accumulated, ref.power_accumulated / ref.power_count) < 2e-4


def test_fdtd_3d_matches_oracle(hip):
    dx = 50e-9
    nx, ny, nz = 40, 32, 24
    d = bz.Design(nx * dx, ny * dx, depth=nz * dx, material=bz.Material(1.444 ** 2), pml_size=6 * dx)
    d += bz.Rectangle(position=(0, ny * dx / 2 - 0.25 * um, nz * dx / 2), width=nx * dx, height=0.5 * um,
                      depth=0.22 * um, material=bz.Material(3.476 ** 2))
    nsteps = 300
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(3))
    t = np.arange(nsteps) * dt
    d += bz.GaussianSource(position=(10 * dx, ny * dx / 2, nz * dx / 2), width=dx, signal=pulse(t))
    mon = bz.Monitor(design=d, start=(0, 0, nz * dx / 2 + dx), end=(nx * dx, ny * dx, nz * dx / 2 + dx),
                     dft_frequencies=[bz.LIGHT_SPEED / 1.55e-6])
    d += mon
    sim = bz.FDTD(d, time=t, resolution=dx)
    sim.run(live=False, save=False)
    ref = R.RefSim(sim.epsilon_r, sim.sigma, dx, dx, dt, nsteps, dz=dx, sources=d.sources, monitors=[mon],
                   design_dims=(d.width, d.height, d.depth)).run()
    for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        assert rel_l2(np.real(getattr(sim, k)), ref.F[k]) < 1e-4, k
    ser = np.array(mon.fields["Ez"]); ref_ser = np.array(ref.mon_series[0]["Ez"])
    assert rel_l2(ser, ref_ser) < 1e-4
    assert rel_l2(mon.power_history, ref.mon_series[0]["power"]) < 2e-4


def test_fused_3d_kernel_segments_and_oracle(hip):
    """The fused z-march kernel (ping-pong, LDS halos, segment prologues), on a grid whose sizes are
    not multiples of the tile (63 x 7) or of the z-segment: the segment length must not change a single
    bit (each segment recomputes the previous plane's H with the same arithmetic), and both update
    kernels must stay within 1e-4 rel-L2 of the FP64 oracle for every component."""
    from beamz_amd.workloads import soi_mmi_3d
    res = {}
    ref = None
    for kern, kseg, rows in (("basic", 16, 4), ("fused", 5, 4), ("fused", 16, 4), ("fused", 7, 2), ("fused", 37, 2)):
        mesh, d3, t, mons = soi_mmi_3d(nx=131, ny=45, nz=37, steps=120, npml=6, dx=40e-9)
        sim = bz.FDTD(mesh, time=t)
        sim.grid.kernel, sim.grid.kseg, sim.grid.rows = kern, kseg, rows
        sim.run(live=False, save=False)
        res[(kern, kseg, rows)] = {k: np.real(getattr(sim, k)) for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
        res[(kern, kseg, rows)]["dft"] = mons[0].dft["Ey"]
        if ref is None:
            ref = R.RefSim(sim.epsilon_r, sim.sigma, sim.dx, sim.dy, sim.dt, len(t), dz=sim.dz,
                           sources=d3.sources).run()
    for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "dft"):
        assert np.array_equal(res[("fused", 5, 4)][k], res[("fused", 16, 4)][k]), k
    for key in res:
        for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            assert rel_l2(res[key][k], ref.F[k]) < 1e-4, (key, k)


def _wg_sim(kernel, nsteps=400, two_monitors=True, freqs_b=None):
    d, dx = waveguide_design_2d(nx=150, ny=70, npml=10)
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(nsteps) * dt
    d += bz.GaussianSource(position=(25 * dx, d.height / 2 + 0.3 * dx), width=0.6 * dx, signal=pulse(t))
    fa = bz.LIGHT_SPEED / np.array([1.5e-6, 1.55e-6, 1.6e-6])
    m1 = bz.Monitor(design=d, start=(100 * dx, 10 * dx), end=(100 * dx, d.height - 10 * dx), dft_frequencies=fa)
    d += m1
    mons = [m1]
    if two_monitors:
        m2 = bz.Monitor(design=d, start=(40 * dx, 20 * dx), end=(140 * dx, 50 * dx),
                        dft_frequencies=fa if freqs_b is None else freqs_b)
        d += m2
        mons.append(m2)
    sim = bz.FDTD(d, time=t, resolution=dx)
    sim.grid.kernel = kernel
    return sim, d, mons, t, dx, dt


@pytest.mark.parametrize("freqs_b", [None, [2.0e14]])
def test_fused_2d_matches_basic_and_oracle(hip, freqs_b):
    out = {}
    for kern in ("basic", "fused"):
        sim, d, mons, t, dx, dt = _wg_sim(kern, freqs_b=freqs_b)
        sim.initialize_simulation(save=False, live=False, accumulate_power=True)
        for _ in range(7):          # single steps, then a long batched run
            sim.step()
        sim._advance(len(t))
        sim.finalize_simulation()
        out[kern] = (sim, mons)
    sim_b, mons_b = out["basic"]
    sim_f, mons_f = out["fused"]
    ref = R.RefSim(sim_f.epsilon_r, sim_f.sigma, dx, dx, dt, len(t), sources=sim_f.sources, monitors=mons_f,
                   accumulate_power=True).run()
    for k in ("Ez", "Hx", "Hy"):
        assert rel_l2(np.real(getattr(sim_f, k)), ref.F[k]) < 1e-4, k
        assert rel_l2(np.real(getattr(sim_f, k)), np.real(getattr(sim_b, k))) < 1e-5, k
    for q, (mf, mb) in enumerate(zip(mons_f, mons_b)):
        ser = np.array(mf.fields["Ez"])
        assert rel_l2(ser, np.array(ref.mon_series[q]["Ez"])) < 1e-4
        assert rel_l2(ser, np.array(mb.fields["Ez"])) < 1e-5
        for k in ("Ez", "Hx", "Hy"):
            F = R.dft_of_series(np.array(ref.mon_series[q][k]), ref.mon_series[q]["t"], 2 * np.pi * mf.dft_frequencies)
            assert rel_l2(mf.dft[k], F) < 1e-4, (q, k)
    assert rel_l2(sim_f.power_accumulated, ref.power_accumulated / ref.power_count) < 2e-4


def test_mode_source_2d_matches_oracle(hip):
    um = 1e-6
    lam = 1.55 * um
    d, dx = waveguide_design_2d(nx=140, ny=70, npml=10)
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(300) * dt
    sig = bz.ramped_cosine(t, 1.0, bz.LIGHT_SPEED / lam, 0.0, ramp_duration=20 * dt, t_max=t[-1])
    src = bz.ModeSource(design=d, start=(20 * dx, d.height / 2 - 1.0 * um), end=(20 * dx, d.height / 2 + 1.0 * um),
                        wavelength=lam, signal=sig, num_modes=1, grid_resolution=400)
    d += src
    for enforce in (True, False):
        src.enforce_direction = enforce
        sim = bz.FDTD(d, time=t, resolution=dx)
        sim.run(live=False, save=False)
        ref = R.RefSim(sim.epsilon_r, sim.sigma, dx, dx, dt, len(t), sources=[src]).run()
        assert rel_l2(np.real(sim.Ez), ref.F["Ez"]) < 1e-4, enforce


def _two_sources():
    """A 2D waveguide with a ModeSource line (with direction enforcement: its mul zeroes upstream cells) and a
    Gaussian point source of another wavelength, on disjoint cells, and a DFT line monitor."""
    um = 1e-6
    d, dx = waveguide_design_2d(nx=150, ny=80, npml=10)
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(300) * dt
    sig = bz.ramped_cosine(t, 1.0, bz.LIGHT_SPEED / (1.55 * um), 0.0, ramp_duration=20 * dt, t_max=t[-1])
    mode = bz.ModeSource(design=d, start=(20 * dx, d.height / 2 - 1.0 * um), end=(20 * dx, d.height / 2 + 1.0 * um),
                         wavelength=1.55 * um, signal=sig, num_modes=1, grid_resolution=400)
    d += mode
    d += bz.GaussianSource(position=(70 * dx, d.height / 2 + 0.3 * um), width=dx / 2, signal=pulse(t, lam=1.31 * um))
    mon = bz.Monitor(design=d, start=(120 * dx, 12 * dx), end=(120 * dx, d.height - 12 * dx),
                     dft_frequencies=bz.LIGHT_SPEED / (np.array([1.31, 1.55]) * um), name="m")
    d += mon
    return d, dx, t, mon


def test_disjoint_sources_share_one_box_in_every_2d_kernel(hip):
    """Two sources of one grid in one source box (each cell reading its own source's signal): one step per
    launch, two steps per launch and K steps per pass give the same bits, alone and inside a batch, and the
    run matches the FP64 oracle, which applies the sources one after the other."""
    from beamz_amd.simulation.batch import FDTDBatch
    out = {}
    # complex_fields=False: the mode profile is complex, and batch members run real fields (the reference's
    # real-backend branch, helper.py:29-33)
    for key in ("one", "pairs", "kstep", "batch"):
        d, dx, t, mon = _two_sources()
        sim = bz.FDTD(d, time=t, resolution=dx, complex_fields=False)
        if key == "one":
            sim.grid.kstep2d = 0
            sim.run(live=False, save=False)
            assert sim.grid._f2d and not sim.grid._f2d.kstep and not sim.grid._f2d.pairs
        elif key == "kstep":
            sim.run(live=False, save=False)
            assert sim.grid._f2d and sim.grid._f2d.kstep
        else:
            other = _k_case("gauss")
            members = [sim, bz.FDTD(other[0], time=other[2], resolution=other[1])]
            b = FDTDBatch(members, kstep=0 if key == "pairs" else None, pairs=key == "pairs")
            b.run()
            assert bool(b.engine.kstep) == (key == "batch")
        assert len(sim.sources) == 2
        out[key] = {k: np.real(getattr(sim, k)).copy() for k in ("Ez", "Hx", "Hy")}
        out[key]["series"] = np.array(mon.fields["Ez"])
        out[key]["dft"] = np.array(mon.dft["Ez"])
    for key in out:
        for k in out["one"]:
            assert np.array_equal(out[key][k], out["one"][k]), (key, k)
    d, dx, t, mon = _two_sources()
    sim = bz.FDTD(d, time=t, resolution=dx, complex_fields=False)
    sim.run(live=False, save=False)
    ref = R.RefSim(sim.epsilon_r, sim.sigma, dx, dx, sim.dt, len(t), sources=sim.sources, monitors=[mon]).run()
    assert np.abs(ref.F["Ez"]).max() > 0
    for k in ("Ez", "Hx", "Hy"):
        assert rel_l2(np.real(getattr(sim, k)), ref.F[k]) < 1e-4, k
    F = R.dft_of_series(np.array(ref.mon_series[0]["Ez"]), ref.mon_series[0]["t"], 2 * np.pi * mon.dft_frequencies)
    assert rel_l2(mon.dft["Ez"], F) < 1e-4


def test_overlapping_sources_leave_the_fused_engine_and_match_oracle(hip):
    """Two Gaussian sources on overlapping cells cannot share one source box: the grid runs the one-kernel-per-
    update path (each source added in turn, as the reference does) and still matches the oracle."""
    d, dx = waveguide_design_2d(nx=120, ny=60, npml=10)
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(200) * dt
    d += bz.GaussianSource(position=(30 * dx, d.height / 2), width=dx, signal=pulse(t))
    d += bz.GaussianSource(position=(30.5 * dx, d.height / 2), width=dx, signal=pulse(t, lam=1.31 * um))
    sim = bz.FDTD(d, time=t, resolution=dx)
    sim.run(live=False, save=False)
    assert sim.grid._f2d is False
    ref = R.RefSim(sim.epsilon_r, sim.sigma, dx, dx, dt, len(t), sources=sim.sources).run()
    for k in ("Ez", "Hx", "Hy"):
        assert rel_l2(np.real(getattr(sim, k)), ref.F[k]) < 1e-4, k


def _k_case(kind):
    """A single 2D grid with one source and one line monitor (series and DFT): a Gaussian point source, or a
    ModeSource line that zeroes its upstream cells."""
    from beamz_amd.workloads import waveguide_2d
    if kind == "gauss":
        d, dx, t, mon = waveguide_2d(nx=150, ny=90, steps=203, npml=10, src_x=20, mon_x=120, n_freq=4)
        return d, dx, t, mon
    if kind == "tiny":       # smaller than one region's inner cells: the region overhangs every edge
        d, dx, t, mon = waveguide_2d(nx=30, ny=14, steps=77, npml=3, src_x=6, mon_x=25, n_freq=2)
        return d, dx, t, mon
    lam = 1.55 * um
    d, dx = waveguide_design_2d(nx=140, ny=70, npml=10)
    dt = 0.99 * dx / (bz.LIGHT_SPEED * np.sqrt(2))
    t = np.arange(203) * dt
    sig = bz.ramped_cosine(t, 1.0, bz.LIGHT_SPEED / lam, 0.0, ramp_duration=20 * dt, t_max=t[-1])
    d += bz.ModeSource(design=d, start=(20 * dx, d.height / 2 - 1.0 * um), end=(20 * dx, d.height / 2 + 1.0 * um),
                       wavelength=lam, signal=sig, num_modes=1, grid_resolution=400)
    mon = bz.Monitor(design=d, start=(100 * dx, 10 * dx), end=(100 * dx, d.height - 10 * dx),
                     dft_frequencies=bz.LIGHT_SPEED / (np.array([1.5, 1.55, 1.6]) * um), name="m")
    d += mon
    return d, dx, t, mon


@pytest.mark.parametrize("K,rows,waves", [(5, 16, 4), (12, 8, 4), (8, 8, 8)])
def test_k_steps_per_pass_single_grid(hip, K, rows, waves):
    """One grid through the K-steps-per-pass kernel gives the bits of one step per launch: fields, the
    monitor's time series and its FP64 DFT, with a Gaussian point source and with a ModeSource line."""
    for kind in ("gauss", "mode", "tiny"):
        out = []
        for k in (0, K):
            d, dx, t, mon = _k_case(kind)
            sim = bz.FDTD(d, time=t, resolution=dx)
            sim.grid.kstep2d, sim.grid.krows2d, sim.grid.kwaves2d = k, rows, waves
            sim.run(live=False, save=False)
            r = {f: np.real(getattr(sim, f)).copy() for f in ("Ez", "Hx", "Hy")}
            for f in ("Ez", "Hx", "Hy"):
                if f in mon.fields:
                    r["series" + f] = np.array(mon.fields[f])
            for f, v in mon.dft.items():
                r["dft" + f] = np.array(v)
            out.append(r)
        assert sim.grid._f2d.kstep == K
        assert any(f.startswith("series") for f in out[0]) and any(f.startswith("dft") for f in out[0])
        for f in out[0]:
            assert np.array_equal(out[0][f], out[1][f]), (kind, K, rows, f)


def test_two_step_3d_kernel_bit_identical_to_single_steps(hip):
    """The two-step kernel (steps n, n+1 in one pass, LDS row exchange, staged monitor samples) must give
    the very bits of two single-step launches, for every rows/segment shape, on a grid whose sizes are not
    multiples of any tile, with an odd step count (pairs plus one single step), sources and DFT monitors."""
    from beamz_amd.workloads import soi_mmi_3d
    out = {}
    for key in (("single", 0, 0), ("pair", 2, 32), ("pair", 3, 5), ("pair", 4, 2), ("pair", 2, 7)):
        mesh, d3, t, mons = soi_mmi_3d(nx=131, ny=45, nz=37, steps=121, npml=6, dx=40e-9)
        sim = bz.FDTD(mesh, time=t)
        sim.grid.kernel, sim.grid.rows = "fused", 2
        sim.grid.tb2 = key[0] == "pair"
        if sim.grid.tb2:
            sim.grid.rows2, sim.grid.kseg2 = key[1], key[2]
        sim.run(live=False, save=False)
        r = {k: np.real(getattr(sim, k)) for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
        for q, m in enumerate(mons):
            for k, v in m.dft.items():
                r[f"dft{q}{k}"] = np.array(v)
        out[key] = r
        if key[0] == "single":
            ref = R.RefSim(sim.epsilon_r, sim.sigma, sim.dx, sim.dy, sim.dt, len(t), dz=sim.dz,
                           sources=d3.sources).run()
            for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
                assert rel_l2(r[k], ref.F[k]) < 1e-4, k
    base = out[("single", 0, 0)]
    for key, r in out.items():
        for k in base:
            assert np.array_equal(r[k], base[k]), (key, k)


def test_three_step_3d_kernel_bit_identical_to_single_steps(hip):
    """The three-step kernel (steps n, n+1, n+2 in one pass, two staging boxes, one three-step gather) must
    give the very bits of single-step launches for every (rows, waves) shape and segment length, on a grid
    whose sizes are not multiples of any tile, with sources and DFT monitors; 121 steps end with one single
    step, 122 with one two-step pass."""
    from beamz_amd.workloads import soi_mmi_3d
    for steps in (121, 122):
        out = {}
        for key in (("single", 0, 0, 0), ("triple", 1, 16, 0), ("triple", 2, 8, 32), ("triple", 2, 8, 2),
                    ("triple", 1, 16, 5), ("triple", 2, 16, 7), ("triple", 3, 8, 11)):
            mesh, d3, t, mons = soi_mmi_3d(nx=131, ny=45, nz=37, steps=steps, npml=6, dx=40e-9, record_ports=True)
            sim = bz.FDTD(mesh, time=t)
            sim.grid.kernel, sim.grid.rows = "fused", 2
            sim.grid.tb2 = key[0] == "triple"
            sim.grid.tb3 = key[0] == "triple"
            if sim.grid.tb3:
                sim.grid.rows3, sim.grid.waves3, sim.grid.kseg3 = key[1], key[2], key[3] or None
            sim.run(live=False, save=False)
            r = {k: np.real(getattr(sim, k)) for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
            for q, m in enumerate(mons):
                for k, v in m.dft.items():
                    r[f"dft{q}{k}"] = np.array(v)
                for k, v in m.fields.items():
                    if len(v):
                        r[f"series{q}{k}"] = np.array(v)
            assert any(k.startswith("dft") for k in r) and any(k.startswith("series") for k in r)
            out[key] = r
        base = out[("single", 0, 0, 0)]
        for key, r in out.items():
            for k in base:
                assert np.array_equal(r[k], base[k]), (steps, key, k)


def test_watchdog_counts_nonfinite_entries(hip):
    """bz_nonfinite_count counts NaN and infinite entries exactly (across several workgroups); a normal run
    reports no component, and a NaN planted in a field is reported for that component only."""
    torch = hip
    from beamz_amd import _native as N
    st = torch.cuda.current_stream().cuda_stream
    a = torch.zeros(1_000_003, device="cuda")
    a[[5, 77, 999_999, 1_000_002]] = float("nan")
    a[123] = float("inf")
    a[124] = -float("inf")
    a[125] = 3.0e38                       # large but finite
    c = torch.zeros(2, dtype=torch.int32, device="cuda")
    N.call("bz_nonfinite_count", a.data_ptr(), a.numel(), c.data_ptr(), st)
    N.call("bz_nonfinite_count", a.data_ptr(), 100, c.data_ptr() + 4, st)
    assert c.cpu().tolist() == [6, 2]
    sim = _wg_sim("fused", nsteps=60)[0]
    sim.run(live=False, save=False)
    assert sim.check_fields() == {}
    sim.grid.f["Hx"][3, 4] = float("nan")
    assert sim.check_fields() == {"Hx": 1}
